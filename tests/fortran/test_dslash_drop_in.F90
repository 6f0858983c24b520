! Reference-style drop-in test: tests/test_dslash_shamir.F90 and
! tests/test_dslashd_shamir.F90 call pattern against the GPU modules.
! Writes Phi (halo-free) for imass 1, 3, 5 and both operators to
! <outdir>/dslash_<op>_<imass>.bin for comparison with the oracle.
program test_dslash_drop_in
  use params
  use dirac
  use comms
  use comms5
  use inverter_utils, only: dirac_op_shifted, dirac_operator
  use test_state
  implicit none
  complex(dp), allocatable :: u(:, :, :, :), Phi(:, :, :, :, :), R(:, :, :, :, :)
  integer :: k, imass
  integer, dimension(16) :: reqs
  character(len=256) :: outdir
  character(len=1) :: ic
  real, parameter :: am = 0.05
  call get_command_argument(1, outdir)
  allocate (u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3))
  allocate (Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4))
  allocate (R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4))
  call init_MPI
  do k = 1, 3
    imass = 2*k - 1
    write (ic, '(I1)') imass
    call starting_state(Phi, u, R=R)
    call dslash_shamir(Phi, R, u, am, imass)
    call start_halo_update_5(4, Phi, 2, reqs)
    call complete_halo_update(reqs)
    call dump(trim(outdir)//'/dslash_M_'//ic//'.bin', &
              reshape(Phi(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :), [kthird_l*kvol*4]), &
              kthird_l*kvol*4)
    call starting_state(Phi, u, R=R)
    call dslashd_shamir(Phi, R, u, am, imass, reqs)
    call dump(trim(outdir)//'/dslash_Md_'//ic//'.bin', &
              reshape(Phi(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :), [kthird_l*kvol*4]), &
              kthird_l*kvol*4)
  end do
  ! inverter_utils.F90:9-54 call pattern: M^+ M R + shift R, and shift = 0
  call starting_state(Phi, u, R=R)
  call dirac_op_shifted(Phi, R, u, am, 3, 0.37)
  call dump(trim(outdir)//'/dirac_op_shifted.bin', &
            reshape(Phi(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :), [kthird_l*kvol*4]), &
            kthird_l*kvol*4)
  call dirac_operator(Phi, R, u, am, 3)
  call dump(trim(outdir)//'/dirac_operator.bin', &
            reshape(Phi(1:kthird_l, 1:ksizex_l, 1:ksizey_l, 1:ksizet_l, :), [kthird_l*kvol*4]), &
            kthird_l*kvol*4)
  print *, 'test_dslash_drop_in done'
end program test_dslash_drop_in
