! Reference-style drop-in test: tests/test_qmrherm.F90 call pattern
! (iflag 0..3, Phi0 seeded from the closed form, max_qmr_iters = 2 so
! itercg == 2, tests/test_qmrherm.F90:88) plus tests/test_derivs_shamir.F90.
program test_qmrherm_drop_in
  use qmrherm_module, only: qmrherm, phi0, qmrhprint => printall
  use trial, only: u
  use vector, only: X
  use gforce
  use params
  use comms
  use derivs_module, only: derivs_shamir
  use test_state
  implicit none
  complex, parameter :: iunit = cmplx(0, 1)
  real(dp), parameter :: tau = 8*atan(1.0_8)
  integer :: imass, iflag, i, itercg
  character(len=1) :: fc
  real :: res, am
  complex(dp), allocatable :: Phi(:, :, :, :, :), Phi0_orig(:, :, :, :, :, :), R(:, :, :, :, :)
  real(dp) :: dSdpi_ref(ksizex_l, ksizey_l, ksizet_l, 3)
  real(dp) :: anum(0:ndiag), aden(ndiag)
  character(len=256) :: outdir
  call get_command_argument(1, outdir)
  allocate (Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4))
  allocate (R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4))
  allocate (Phi0_orig(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4, 25))
  call init_MPI
  qmrhprint = .false.
  res = 0.1
  am = 0.05
  imass = 3
  do iflag = 0, 3
    write (fc, '(I1)') iflag
    anum(0) = 0.5
    do i = 1, ndiag
      anum(i) = real(exp(iunit*i*tau/ndiag))
      aden(i) = real(exp(-iunit*0.5*i*tau/ndiag))
    end do
    call starting_state(Phi, u, R=R, X=X, Phi0_orig=Phi0_orig, dSdpi_ref=dSdpi_ref)
    Phi0 = Phi0_orig
    dSdpi = dSdpi_ref
    ! the compiler may fold 1.7*exp(1.0) in higher precision than default
    ! real: hand the checker the exact Phi0 this program starts from
    if (iflag == 0) call dump(trim(outdir)//'/qmrherm_phi0_in.bin', &
                              reshape(Phi0_orig(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :, :), &
                                      [kthird_l*kvol*4*25]), kthird_l*kvol*4*25)
    call qmrherm(Phi, X, res, itercg, am, imass, anum, aden, ndiag, iflag)
    if (itercg == 2) then
      print *, 'PASS itercg iflag ', iflag
    else
      print *, 'FAIL itercg iflag ', iflag, itercg
    end if
    call dump(trim(outdir)//'/qmrherm_x_'//fc//'.bin', &
              reshape(X(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :), [kthird_l*kvol*4]), &
              kthird_l*kvol*4)
    call dump(trim(outdir)//'/qmrherm_phi0_'//fc//'.bin', &
              reshape(Phi0(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :, :), &
                      [kthird_l*kvol*4*25]), kthird_l*kvol*4*25)
    call dump_real(trim(outdir)//'/qmrherm_dsdpi_'//fc//'.bin', &
                   reshape(dSdpi, [kvol*3]), kvol*3)
  end do
  ! tests/test_derivs_shamir.F90: dSdpi = dSdpi_orig; derivs_shamir(R, X2, tau, 0, am, imass)
  call starting_state(Phi, u, R=R, X=X, dSdpi_ref=dSdpi_ref)
  dSdpi = dSdpi_ref
  ! the exact inputs, for the checker (constant folding may differ from the
  ! reference compiler's in the last bit of default-real literals)
  call dump(trim(outdir)//'/derivs_R.bin', reshape(R, [size(R)]), size(R))
  call dump(trim(outdir)//'/derivs_X.bin', reshape(X, [size(X)]), size(X))
  call derivs_shamir(R, X, tau, 0, am, imass)
  call dump_real(trim(outdir)//'/derivs_shamir.bin', reshape(dSdpi, [kvol*3]), kvol*3)
  print *, 'test_qmrherm_drop_in done'
end program test_qmrherm_drop_in
