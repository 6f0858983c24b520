! Reference-style drop-in test: the tests/test_measure_shamir.F90 call pattern
! (seed, closed-form starting state, measure_shamir) against the drop-in
! measure_module.  Known answer psibarpsi = 2.504295E-04 and aviter = 5
! (tests/test_measure_shamir.F90:45-46, relative tolerance 0.001); then the
! same measurement through `measure`, the entry dwf3d_lib.F90:355 calls, with
! isweep_total; then measure_wilson's known answer.  Records go to fort.100 /
! fort.200 in the working directory.
program test_measure_drop_in
  use params
  use trial, only: u
  use vector
  use comms
  use random
  use measure_module
  use test_state
  implicit none
  integer :: imass
  real :: res, am, psibarpsi, aviter, pbp2, aviter2
  real(dp) :: seed
  complex(dp), allocatable :: Phi(:, :, :, :, :)
  logical :: ok
  allocate (Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4))
  call init_MPI
  seed = 4139764973254.0
  call init_random(seed)
  res = 0.1
  am = 0.05
  imass = 3
  call starting_state(Phi, u, X=X)
  call measure_shamir(psibarpsi, res, aviter, am, imass)
  ok = abs(psibarpsi - 2.504295E-04) <= 0.001*2.504295E-04 .and. abs(aviter - 5.0) <= 0.001*5.0
  print *, 'psibarpsi', psibarpsi, 'aviter', aviter
  call init_random(seed)
  call measure(pbp2, res, aviter2, am, imass, 7)
  print *, 'measure', pbp2, aviter2
  ok = ok .and. pbp2 == psibarpsi .and. aviter2 == aviter
  ! the Wilson-kernel measurement (tests/test_measure_wilson.F90:24-47): the
  ! reference's value -5.4226685E-02 predates its imass = 3 phase swap
  ! (measure_module.F90:612-618), which negates it
  call init_random(seed)
  call measure_wilson(pbp2, res, aviter2, am, imass)
  print *, 'measure_wilson', pbp2, aviter2
  ok = ok .and. abs(pbp2 - 5.4226685E-02) <= 0.001*5.4226685E-02 .and. &
       abs(aviter2 - 10.9) <= 0.001*10.9
  if (ok) then
    print *, 'PASS measure_shamir'
  else
    print *, 'FAIL measure_shamir'
  end if
end program test_measure_drop_in
