! Side-by-side run of the reference's iterate() and the HIP drop-in
! (test infrastructure only).  One Fortran program holds both: the reference
! modules compiled from their sources (oracle/ref.mk) and the product's
! ISO_C_BINDING drop-in (sph_playground_amd/fortran/iterator_hip.f90) calling
! libsph_hip.so.  Both read the same `state` module, the way a Fortran host
! would switch from one to the other.
!
! Input: the ref_harness.f90 file format (mode 0).  Prints, per written store
! row, max |hip - ref| / max |ref| over real particles; writes the HIP store
! to the output file (int32 n_tot, real64 store(32, n_tot)).
program ref_vs_hip
  use const
  use state,        only: setState, getdim
  use kernel,       only: initkernel
  use BC,           only: initBorders
  use iterator,     only: iterate_ref => iterate
  use iterator_hip, only: iterate_hip => iterate, iterate_hip_kernel
  implicit none

  character(len=1024) :: fin, fout, karg
  integer :: mode, nreal, nfixed, sts_s, ntot, dim, f, kid
  real :: dtcap, e1, e2, worst, d, sc
  real :: st(ec_total)
  real, allocatable :: store(:,:), sref(:,:), shost(:,:)
  integer, parameter :: rows(21) = [es_rx, es_ry, es_rz, es_den, es_h, es_om, es_p, es_c, es_ax, es_ay, &
                                    es_az, es_du, es_dh, es_ddt, es_dtdx, es_dtdy, es_dtdz, es_dbx, es_dby, &
                                    es_dbz, es_vfld]

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, karg)
  kid = 1
  if (len_trim(karg) > 0) read(karg, *) kid

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) mode, nreal, nfixed, sts_s
  read(10) dtcap
  read(10) st
  ntot = nreal + nfixed
  allocate(store(es_total, 28*ntot + 64))
  store(:,:) = 0.
  read(10) store(:, 1:ntot)
  close(10)

  call setState(st)
  call getdim(dim)
  call initkernel(dim)
  call initBorders()

  allocate(sref(es_total, size(store, 2)))
  sref(:,:) = store(:,:)
  allocate(shost(es_total, ntot))
  shost(:,:) = store(:, 1:ntot)

  call iterate_ref(ntot, sref, e1)
  call iterate_hip_kernel(kid)
  call iterate_hip(ntot, shost, e2)

  worst = 0.
  do f = 1, size(rows)
    sc = maxval(abs(sref(rows(f), 1:nreal)))
    d = maxval(abs(shost(rows(f), 1:nreal) - sref(rows(f), 1:nreal)))
    if (sc > 0.) d = d / sc
    worst = max(worst, d)
    print '(A,I3,A,ES12.4)', 'row ', rows(f), ' relerr ', d
  end do
  print '(A,ES12.4,A,2ES12.4)', 'worst ', worst, ' maxconsenrg ', e1, e2

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) ntot
  write(11) shost(:, 1:ntot)
  close(11)
end program ref_vs_hip
