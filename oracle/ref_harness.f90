! Driver for the reference Fortran hot path (test infrastructure only).
!
! Reads one problem from a binary file, runs the reference's own modules on it
! and writes the result.  It replaces main.f90's set-up (HDF5 restore, IC
! placement, CLI parsing) with a direct fill of `statevars` and `store`, then
! calls exactly the reference entry points:
!   mode 0: iterate(n, store, dedt)                        iterator.f90:44
!   mode 1: leapfrog steps as main.f90:216-362 (no STS): iterate once, then
!           max(1, sts_s) times calc, predictor, iterate, corrector;
!           dt = min(calc dt, dtcap) (the last step's dt is written)
!   mode 2: sts_integrate(n, s, store, dt)                 sts_integrator.f90:87
!   mode 3: iterate twice, timing only the second call (steady state: the
!           first call also runs initIterate's extra neighbour search)
!   mode 4: mode 1, then update_radenergy(n, store, dt) as main.f90:367-369
!   mode 5: update_radenergy(n, store, dtcap) alone        rad_exchange.f90:9
!   mode 6: createFixedBorders(store, sts_s) on the real particles (BC.f90:477,
!           sts_s = the ebc_* side), as the mti / fld_gauss set-ups call it
!           (IC/IC.f90:170, IC/setups/place_fld_gauss.f90:22); writes the
!           store with the new fixed particles appended, nothing else runs
!   mode + 100: same, without writing the neighbour lists
!
! Input  (stream, little endian): int32 mode, n_real, n_fixed, sts_s;
!        real64 dtcap; real64 state(49); real64 store(32, n_real+n_fixed)
! Output (stream): int32 n_tot; real64 store(32, n_tot); real64 dt_used;
!        int32 n_real; per real i: int32 count, then count pairs
!        (int32 real index of j, int32 image code), image code =
!        (sx+1) + 3*(sy+1) + 9*(sz+1) with s = sign(x_ghost - x_source) per axis.
program ref_harness
  use const
  use state,          only: setState, getdim, getPartNumber
  use kernel,         only: initkernel
  use BC,             only: initBorders, getCrossRef, createFixedBorders
  use iterator,       only: iterate
  use timestep,       only: calc
  use sts_integrator, only: sts_integrate, sts_init => init
  use rad_exchange,   only: update_radenergy
  use neighboursearch, only: xgetneighnumber, xgetneighindex
  use omp_lib
  implicit none

  character(len=1024) :: fin, fout
  integer :: mode, nreal, nfixed, sts_s, ntot, dim, i, k, j, rj, code, cnt, ax
  integer :: sgn(3), dumpnb, step, nsteps, rpn, fpn
  integer(8) :: npairs
  real :: dtcap, dedt, dt, dthydro, dtdiff, dtfld, t0, t1
  real :: st(ec_total)
  real, allocatable :: store(:,:), tr(:,:), tv(:,:), ta(:,:), tdu(:), tdh(:), tddt(:), tdb(:,:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) mode, nreal, nfixed, sts_s
  dumpnb = 1
  if (mode >= 100) then
    dumpnb = 0
    mode = mode - 100
  end if
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

  dt = 0.
  if (mode == 3) call iterate(ntot, store, dedt)
  t0 = omp_get_wtime()
  select case (mode)
  case (0, 3)
    call iterate(ntot, store, dedt)
  case (1, 4)
    ! main.f90:219 initial accelerations, then passes of the loop body
    call iterate(ntot, store, dedt)
    nsteps = max(1, sts_s)
    allocate(tr(3,ntot), tv(3,ntot), ta(3,ntot), tdu(ntot), tdh(ntot), tddt(ntot), tdb(3,ntot))
    do step = 1, nsteps
      call calc(store, dthydro, dtdiff, dtfld, dt)
      dt = min(dt, dtcap)
      tr(:,:)  = store(es_rx:es_rz,1:ntot)
      tv(:,:)  = store(es_vx:es_vz,1:ntot)
      ta(:,:)  = store(es_ax:es_az,1:ntot)
      tdu(:)   = store(es_du,1:ntot)
      tdh(:)   = store(es_dh,1:ntot)
      tddt(:)  = store(es_ddt,1:ntot)
      tdb(:,:) = store(es_dbx:es_dbz,1:ntot)
      store(es_rx:es_rz,1:ntot) = store(es_rx:es_rz,1:ntot) + dt*tv(:,:) + 0.5*dt*dt*ta(:,:)
      store(es_vx:es_vz,1:ntot) = tv(:,:) + dt*ta(:,:)
      store(es_u,1:ntot) = store(es_u,1:ntot) + dt*tdu(:)
      store(es_h,1:ntot) = store(es_h,1:ntot) + dt*tdh(:)
      store(es_t,1:ntot) = store(es_t,1:ntot) + dt*tddt(:)
      store(es_bx:es_bz,1:ntot) = store(es_bx:es_bz,1:ntot) + dt*tdb(:,:)
      call iterate(ntot, store, dedt)
      store(es_vx:es_vz,1:ntot) = store(es_vx:es_vz,1:ntot) + 0.5*dt*(store(es_ax:es_az,1:ntot) - ta(:,:))
      store(es_u,1:ntot) = store(es_u,1:ntot) + 0.5*dt*(store(es_du,1:ntot) - tdu(:))
      store(es_h,1:ntot) = store(es_h,1:ntot) + 0.5*dt*(store(es_dh,1:ntot) - tdh(:))
      store(es_bx:es_bz,1:ntot) = store(es_bx:es_bz,1:ntot) + 0.5*dt*(store(es_dbx:es_dbz,1:ntot) - tdb(:,:))
      store(es_t,1:ntot) = store(es_t,1:ntot) + 0.5*dt*(store(es_ddt,1:ntot) - tddt(:))
      if (mode == 4) call update_radenergy(ntot, store, dt)
    end do
  case (6)
    call createFixedBorders(store, sts_s)
    call getPartNumber(r=rpn, f=fpn)
    ntot = rpn + fpn
  case (5)
    dt = dtcap
    call update_radenergy(ntot, store, dt)
  case (2)
    dt = dtcap
    call sts_init(size(store, 2))
    call sts_integrate(ntot, sts_s, store, dt)
  case default
    error stop "ref_harness: unknown mode"
  end select
  t1 = omp_get_wtime()
  write(0, '(A,F12.6,A,I0)') 'ref_harness: seconds ', t1 - t0, ' threads ', omp_get_max_threads()

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) ntot
  write(11) store(:, 1:ntot)
  write(11) dt
  if (dumpnb == 0 .or. mode == 5 .or. mode == 6) then
    ! only the pair count P = sum_i |N_i| (none: modes 5 and 6 search nothing)
    npairs = 0
    if (mode /= 5 .and. mode /= 6) then
      do i = 1, nreal
        npairs = npairs + xgetneighnumber(i)
      end do
    end if
    write(11) -1
    write(11) npairs
    close(11)
    stop
  end if
  write(11) nreal
  do i = 1, nreal
    cnt = xgetneighnumber(i)
    write(11) cnt
    do k = 1, cnt
      j = xgetneighindex(i, k)
      rj = getCrossRef(j)
      do ax = 1, 3
        sgn(ax) = 0
        if (store(ax, j) > store(ax, rj)) sgn(ax) = 1
        if (store(ax, j) < store(ax, rj)) sgn(ax) = -1
      end do
      code = (sgn(1)+1) + 3*(sgn(2)+1) + 9*(sgn(3)+1)
      write(11) rj, code
    end do
  end do
  close(11)
end program ref_harness
