! Reference-style drop-in test: tests/test_congrad.F90 call pattern.
! Known answer itercg == 27 (tests/test_congrad.F90:47); writes X.
program test_congrad_drop_in
  use params
  use trial, only: u
  use vector
  use measure_module
  use comms
  use test_state
  implicit none
  integer :: itercg, imass
  real :: res, am
  complex(dp), allocatable :: Phi(:, :, :, :, :)
  character(len=256) :: outdir
  call get_command_argument(1, outdir)
  allocate (Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4))
  call init_MPI
  res = 0.1
  am = 0.05
  imass = 3
  call starting_state(Phi, u, X=X)
  ! exact inputs, so the oracle solves the same system bit for bit
  call dump(trim(outdir)//'/congrad_phi.bin', reshape(Phi, [size(Phi)]), size(Phi))
  call dump(trim(outdir)//'/congrad_x0.bin', reshape(X, [size(X)]), size(X))
  call dump(trim(outdir)//'/congrad_u.bin', reshape(u, [size(u)]), size(u))
  call congrad(Phi, res, itercg, am, imass)
  if (itercg == 27) then
    print *, 'PASS itercg', itercg
  else
    print *, 'FAIL itercg', itercg
  end if
  call dump(trim(outdir)//'/congrad_x.bin', &
            reshape(X(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :), [kthird_l*kvol*4]), &
            kthird_l*kvol*4)
end program test_congrad_drop_in
