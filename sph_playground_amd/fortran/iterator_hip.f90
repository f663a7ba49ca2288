! Drop-in for the reference's step kernel on the GPU:
!     call iterate(n, store, maxconsenrg)        (playground/src/iterator.f90:44)
! reads the engine configuration from the reference's own `state` module
! (state.f90:36, via getState) at every call and creates the device context on
! the first call, where the reference freezes its own (iterator.f90:31-42).
! Errors stop with the reference's message, as its `error stop` does.
! Builds against the reference's state.mod: compiled after state.f90 in the
! reference's source list and linked with -lsph_hip.
module iterator_hip
  use iso_c_binding
  use sph_hip_capi
  use state, only: getState
  implicit none

  public :: iterate, iterate_hip_kernel, iterate_hip_options

  private
  save
  type(c_ptr) :: ctx = c_null_ptr
  integer(c_int) :: kernel_id = 1              ! M6, the reference's `make kernel=m6`
  type(sph_options) :: opts

contains

  ! Kernel of the run (0 M4, 1 M6, 2 gauss); call before the first iterate.
  subroutine iterate_hip_kernel(k)
    integer, intent(in) :: k
    kernel_id = int(k, c_int)
  end subroutine

  ! Device-path options (precision, tensor form, ...); before the first iterate.
  subroutine iterate_hip_options(o)
    type(sph_options), intent(in) :: o
    opts = o
  end subroutine

  ! Drop-in for iterator.f90:44 `iterate(n, store, maxconsenrg)`.
  subroutine iterate(n, store, maxconsenrg)
    real, allocatable, intent(inout) :: store(:,:)
    integer, intent(in) :: n
    real, intent(out) :: maxconsenrg
    real(c_double) :: st(49), e
    integer(c_int) :: rc

    call getState(st)
    if (.not. c_associated(ctx)) then
      rc = sph_ctx_create(st, kernel_id, opts, ctx)
      if (rc /= SPH_OK) error stop "sph_ctx_create failed"
    end if
    rc = sph_iterate(ctx, int(n, c_int64_t), store, int(size(store, 1), c_int64_t), st, e)
    if (rc /= SPH_OK) then
      print *, "# <!> ", sph_message(rc)
      error stop
    end if
    maxconsenrg = e
  end subroutine
end module iterator_hip
