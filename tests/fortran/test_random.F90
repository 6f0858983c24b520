! Per-site random streams and Gaussian noise for the Fortran drop-in tests,
! with the interface of the reference's `random` (site_random.f90) and
! `gaussian` (gaussian.F90) modules, so that measure_shamir draws the
! reference's noise and its known answer (tests/test_measure_shamir.F90:45-46)
! can be checked through the drop-in `measure_module`.  Test infrastructure:
! a caller of the library keeps the reference's own two modules.
!
! Every (x, y, t) site owns a 48-bit linear congruential generator, its state
! held as two 24-bit halves in doubles so the products stay exact, seeded
! with the seed plus the site's global index (site_random.f90:64-72); single-
! precision outputs go through a 97-entry shuffle table (site_random.f90:28-60).
module random
  use params
  implicit none
  save
  real :: yran(ksizex_l, ksizey_l, ksizet_l)
  integer :: idum(ksizex_l, ksizey_l, ksizet_l)
  real :: v(97, ksizex_l, ksizey_l, ksizet_l)
  real(dp), private :: st_lo(ksizex_l, ksizey_l, ksizet_l) = 0
  real(dp), private :: st_hi(ksizex_l, ksizey_l, ksizet_l) = 0
  real(dp), private, parameter :: mul_lo = 15184245.d0, mul_hi = 2651554.d0
  real(dp), private, parameter :: two24 = 16777216.d0, two48 = 281474976710656.d0
contains
  ! one generator step of site (ix, iy, it): state <- mul * state mod 2^48, in [0, 1)
  real(dp) function lcg_step(ix, iy, it)
    integer, intent(in) :: ix, iy, it
    real(dp) :: low, carry, high
    low = st_lo(ix, iy, it)*mul_lo
    carry = aint(low/two24)
    low = low - carry*two24
    high = st_lo(ix, iy, it)*mul_hi + st_hi(ix, iy, it)*mul_lo + carry
    st_hi(ix, iy, it) = high - aint(high/two24)*two24
    st_lo(ix, iy, it) = low
    lcg_step = (st_hi(ix, iy, it)*two24 + st_lo(ix, iy, it))/two48
  end function lcg_step

  real function next_single(ix, iy, it)
    integer, intent(in) :: ix, iy, it
    next_single = real(lcg_step(ix, iy, it))
  end function next_single

  ! shuffled uniform of one site; i < 0 fills the site's table first
  real function rano(y, i, ix, iy, it)
    real, intent(inout) :: y(ksizex_l, ksizey_l, ksizet_l)
    integer, intent(inout) :: i(ksizex_l, ksizey_l, ksizet_l)
    integer, intent(in) :: ix, iy, it
    integer :: k
    real :: skip
    if (i(ix, iy, it) < 0) then
      i(ix, iy, it) = 1
      do k = 1, 97
        skip = next_single(ix, iy, it)
      end do
      do k = 1, 97
        v(k, ix, iy, it) = next_single(ix, iy, it)
      end do
      y(ix, iy, it) = next_single(ix, iy, it)
    end if
    k = min(97, max(1, 1 + int(97.0*y(ix, iy, it))))
    y(ix, iy, it) = v(k, ix, iy, it)
    rano = y(ix, iy, it)
    v(k, ix, iy, it) = next_single(ix, iy, it)
  end function rano

  subroutine init_random(seed)
    real(dp), intent(in) :: seed
    integer :: ix, iy, it
    real(dp) :: s0
    real :: first
    do it = 1, ksizet_l
      do iy = 1, ksizey_l
        do ix = 1, ksizex_l
          ! one process: the global index is the local one; the y and t strides are ksize
          s0 = (ix - 1 + ksize*(iy - 1) + ksize*ksize*(it - 1)) + seed
          st_hi(ix, iy, it) = aint(s0/two24)
          st_lo(ix, iy, it) = s0 - st_hi(ix, iy, it)*two24
          idum(ix, iy, it) = -1
          first = rano(yran, idum, ix, iy, it)
        end do
      end do
    end do
  end subroutine init_random
end module random

module gaussian
  use random
  use params
  use comms4, only: update_halo_4_real
  implicit none
  real, parameter :: tpi = 2.0*acos(-1.0)
contains
  ! unit-variance complex noise, Box-Muller per site in single precision
  ! (gaussian.F90:63-104): modulus from the first uniform, angle from the second
  subroutine gauss0(ps)
    real, intent(out) :: ps(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 2)
    integer :: ix, iy, it
    real :: ang
    do it = 1, ksizet_l
      do iy = 1, ksizey_l
        do ix = 1, ksizex_l
          ps(ix, iy, it, 2) = sqrt(-log(rano(yran, idum, ix, iy, it)))
        end do
      end do
    end do
    do it = 1, ksizet_l
      do iy = 1, ksizey_l
        do ix = 1, ksizex_l
          ang = tpi*rano(yran, idum, ix, iy, it)
          ps(ix, iy, it, 1) = ps(ix, iy, it, 2)*sin(ang)
          ps(ix, iy, it, 2) = ps(ix, iy, it, 2)*cos(ang)
        end do
      end do
    end do
    call update_halo_4_real(2, ps)
  end subroutine gauss0
end module gaussian
