! Closed-form starting state of the reference's unit tests
! (tests/test_utils.F90:35-151), restated for the Fortran drop-in tests.
module test_state
  use params
  use comms4
  use comms5
  use comms6
  use gammamatrices
  implicit none
contains
  subroutine starting_state(Phi, u, R, X, Phi0_orig, dSdpi_ref)
    complex(dp), intent(inout) :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
    complex(dp), intent(inout) :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), intent(inout), optional :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
    complex(dp), intent(inout), optional :: X(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
    complex(dp), intent(inout), optional :: Phi0_orig(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                                      0:ksizet_l + 1, 4, 25)
    real(dp), intent(inout), optional :: dSdpi_ref(ksizex_l, ksizey_l, ksizet_l, 3)
    complex, parameter :: iunit = cmplx(0, 1)
    real(dp), parameter :: tau = 8*atan(1.0_8)
    integer, parameter :: idxmax = 4*ksize*ksize*ksizet*kthird
    integer :: l, j, ix, iy, it, ithird, idx
    do j = 1, 4
      do it = 1, ksizet_l
        do iy = 1, ksizey_l
          do ix = 1, ksizex_l
            do ithird = 1, kthird_l
              idx = ithird + (ix - 1)*kthird + (iy - 1)*kthird*ksize + (it - 1)*kthird*ksize*ksize &
                    + (j - 1)*kthird*ksize*ksize*ksizet
              Phi(ithird, ix, iy, it, j) = 1.1*exp(iunit*idx*tau/idxmax)
              if (present(R)) R(ithird, ix, iy, it, j) = 1.3*exp(iunit*idx*tau/idxmax)
              if (present(X)) X(ithird, ix, iy, it, j) = 0.5*exp(1.0)*exp(iunit*idx*tau/idxmax)
              if (present(Phi0_orig)) then
                do l = 1, 25
                  Phi0_orig(ithird, ix, iy, it, j, l) = 1.7*exp(1.0)*exp(iunit*idx*tau/idxmax) + l
                end do
              end if
            end do
          end do
        end do
      end do
    end do
    do j = 1, 3
      do it = 1, ksizet_l
        do iy = 1, ksizey_l
          do ix = 1, ksizex_l
            idx = ix + (iy - 1)*ksize + (it - 1)*ksize*ksize + (j - 1)*ksize*ksize*ksizet
            u(ix, iy, it, j) = exp(iunit*idx*tau/idxmax)
            if (present(dSdpi_ref)) dSdpi_ref(ix, iy, it, j) = real(tau*exp(iunit*idx*tau/idxmax), dp)
          end do
        end do
      end do
    end do
    call update_halo_4(3, u)
    call update_halo_5(4, Phi)
    if (present(R)) call update_halo_5(4, R)
    if (present(X)) call update_halo_5(4, X)
    if (present(Phi0_orig)) call update_halo_6(4, 25, Phi0_orig)
    beta = 0.4
    am3 = 1.0
    ibound = -1
    call init_gammas()
  end subroutine starting_state

  subroutine dump(name, a, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    complex(dp), intent(in) :: a(n)
    integer :: u
    open (newunit=u, file=name, access='stream', form='unformatted', status='replace')
    write (u) a
    close (u)
  end subroutine dump

  subroutine dump_real(name, a, n)
    character(len=*), intent(in) :: name
    integer, intent(in) :: n
    real(dp), intent(in) :: a(n)
    integer :: u
    open (newunit=u, file=name, access='stream', form='unformatted', status='replace')
    write (u) a
    close (u)
  end subroutine dump_real
end module test_state
