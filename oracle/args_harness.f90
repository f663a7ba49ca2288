! Test driver: the reference's own fillargs (args.f90:29-176) on this
! program's command line, from a cleared state (state.f90:43-45), then the
! statevars vector, one value per line.  Built by oracle/ref.mk into
! oracle/_ref/<kernel>/args_harness; tests/golden/make_args_golden.py runs it
! to write the fixtures of tests/test_args.py.  Test infrastructure only.
program args_harness
  use const
  use state, only: clearState, getState
  use args, only: fillargs
  implicit none
  real :: s(ec_total)
  integer :: i

  call clearState()
  call fillargs()
  call getState(s)
  print '(A)', 'STATEVARS'
  do i = 1, ec_total
    print '(ES26.17E3)', s(i)
  end do
end program args_harness
