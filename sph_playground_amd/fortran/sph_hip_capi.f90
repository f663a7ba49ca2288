! ISO_C_BINDING interface to libsph_hip.so (include/sph_hip.h) for Fortran
! hosts: one interface per extern "C" entry point, plus the options block.
! Standalone (no dependency on the reference's modules).
module sph_hip_capi
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: SPH_OK = 0

  type, bind(C) :: sph_options
    integer(c_int) :: precision = 0
    integer(c_int) :: tensor_mode = 0
    real(c_double) :: kappa_perp = 0.0_c_double
    integer(c_int) :: max_sweeps = 100
    real(c_double) :: newton_tol = 0.0_c_double
    integer(c_int) :: neigh_cap = 0
    integer(c_int) :: device = 0
    integer(c_int) :: cell_div = 0
  end type

  interface
    integer(c_int) function sph_default_options(opt) bind(C, name="sph_default_options")
      import :: c_int, sph_options
      type(sph_options), intent(out) :: opt
    end function
    integer(c_int) function sph_ctx_create(state, kernel_id, opt, ctx) bind(C, name="sph_ctx_create")
      import :: c_int, c_double, c_ptr, sph_options
      real(c_double), intent(in) :: state(49)
      integer(c_int), value :: kernel_id
      type(sph_options), intent(in) :: opt
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function sph_ctx_destroy(ctx) bind(C, name="sph_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function sph_iterate(ctx, n, store, ld, state, maxconsenrg) bind(C, name="sph_iterate")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n, ld
      real(c_double), intent(inout) :: store(*)
      real(c_double), intent(in) :: state(49)
      real(c_double), intent(out) :: maxconsenrg
    end function
    integer(c_int) function sph_load(ctx, n, store, ld, on_device) bind(C, name="sph_load")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: n, ld
      real(c_double), intent(in) :: store(*)
      integer(c_int), value :: on_device
    end function
    integer(c_int) function sph_save(ctx, store, ld, on_device) bind(C, name="sph_save")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: store(*)
      integer(c_int64_t), value :: ld
      integer(c_int), value :: on_device
    end function
    integer(c_int) function sph_step(ctx, state, maxconsenrg) bind(C, name="sph_step")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: state(49)
      real(c_double), intent(out) :: maxconsenrg
    end function
    integer(c_int) function sph_leapfrog(ctx, state, dtcap, dt_used) bind(C, name="sph_leapfrog")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: state(49)
      real(c_double), value :: dtcap
      real(c_double), intent(out) :: dt_used
    end function
    integer(c_int) function sph_sts(ctx, state, s, dt) bind(C, name="sph_sts")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: state(49)
      integer(c_int), value :: s
      real(c_double), value :: dt
    end function
    integer(c_int) function sph_rad_exchange(ctx, state, dt) bind(C, name="sph_rad_exchange")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: state(49)
      real(c_double), value :: dt
    end function
    integer(c_int) function sph_sts_coeffs(j, s, out) bind(C, name="sph_sts_coeffs")
      import :: c_int, c_double
      integer(c_int), value :: j, s
      real(c_double), intent(out) :: out(4)
    end function
    type(c_ptr) function sph_error_string(code) bind(C, name="sph_error_string")
      import :: c_int, c_ptr
      integer(c_int), value :: code
    end function
  end interface

contains

  ! The message of an SPH_* code as a Fortran string.
  function sph_message(code) result(msg)
    integer(c_int), intent(in) :: code
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: k
    call c_f_pointer(sph_error_string(code), p, [256])
    k = 0
    do while (k < 256)
      if (p(k + 1) == c_null_char) exit
      k = k + 1
    end do
    allocate(character(len=k) :: msg)
    do k = 1, len(msg)
      msg(k:k) = p(k)
    end do
  end function
end module sph_hip_capi


