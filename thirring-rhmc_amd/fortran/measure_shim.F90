! Fortran drop-in `measure_module` (measure_module.F90): congrad, measure,
! measure_shamir and measure_wilson with the reference's names, argument
! lists and outputs.
!
! Compiled after shims.F90 and after the caller's own `random` and `gaussian`
! modules: the noise of the psibarpsi measurement comes from the caller's
! gauss0, drawn in the reference's order (two fields per noise vector), so a
! run reproduces the reference's numbers.  The wall-source solves of each
! noise vector (dslashd_shamir or dslashd_wilson, then congrad from the
! source, and for the Wilson kernel the DWilson projection of the read slices)
! run on the GPU in one call of thr_measure_shamir_noise or
! thr_measure_wilson_noise; the normalisation, the
! susceptibility and the fort.100 / fort.200 records are the reference's.

module measure_module
#ifdef MEASURE_OWILSON
  use measure_OWilson
#endif
  implicit none
contains
  ! measure_module.F90:9-140: CG on M^dagger M, initial guess and solution in vector::X
  subroutine congrad(Phi, res, itercg, am, imass, iterations)
    use trial, only: u
    use vector
    use params
    use thr_capi
    complex(dp), intent(in), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                           0:ksizet_l + 1, 4)
    real, intent(in) :: res, am
    integer, intent(out) :: itercg
    integer, intent(in) :: imass
    integer, intent(out), optional :: iterations
    integer(c_int) :: d(4), it, nx
    call thr_ensure_init()
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
    call thr_check(thr_congrad(c_loc(Phi), c_loc(X), c_loc(u), d, real(res, c_float), &
                               real(am, c_float), int(imass, c_int), real(am3, c_float), &
                               int(niterc, c_int), generation_kernel(), it, nx), 'congrad')
    itercg = it
    if (nx > niterc) then
      if (ip_global_zero()) then
        write (7, '(a)') ' # iterations of congrad exceeds niterc'
        write (7, *) "res, Iterations:", res, nx
      end if
    end if
    if (present(iterations)) iterations = nx
  end subroutine congrad

  logical function ip_global_zero()
    use comms_common, only: ip_global
    ip_global_zero = ip_global == 0
  end function ip_global_zero

  ! measure_module.F90:142-162: the measurement the build selects
  subroutine measure(psibarpsi, res, aviter, am, imass, isweep_total)
    real, intent(out) :: psibarpsi, aviter
    real, intent(in) :: res, am
    integer, intent(in) :: imass
    integer, intent(in), optional :: isweep_total
#ifdef MEASURE_SHAMIR
    call measure_shamir(psibarpsi, res, aviter, am, imass, isweep_total)
#endif
#ifdef MEASURE_WILSON
    call measure_wilson(psibarpsi, res, aviter, am, imass, isweep_total)
#endif
#ifdef MEASURE_OWILSON
    call measure_wilsonKDDW(psibarpsi, res, aviter, am, imass, isweep_total)
#endif
  end subroutine measure

  ! measure_module.F90:170-423: noisy psibarpsi with wall sources
  subroutine measure_shamir(psibarpsi, res, aviter, am, imass, isweep_total)
    real, intent(out) :: psibarpsi, aviter
    real, intent(in) :: res, am
    integer, intent(in) :: imass
    integer, intent(in), optional :: isweep_total
    call noisy_measurement(0, psibarpsi, res, aviter, am, imass, isweep_total)
  end subroutine measure_shamir

  ! measure_module.F90:431-640: the same with Wilson-kernel wall sources and
  ! the DWilson projection of the read slices (imass 1 and 3)
  subroutine measure_wilson(psibarpsi, res, aviter, am, imass, isweep_total)
    real, intent(out) :: psibarpsi, aviter
    real, intent(in) :: res, am
    integer, intent(in) :: imass
    integer, intent(in), optional :: isweep_total
    print *, "Wilson measure"
    call noisy_measurement(1, psibarpsi, res, aviter, am, imass, isweep_total)
  end subroutine measure_wilson

  ! the noise loop both measurements share: noise drawn with the caller's
  ! gauss0 in the reference's order, one library call per noise vector for
  ! the wall-source solves, the reference's normalisation and records
  subroutine noisy_measurement(which, psibarpsi, res, aviter, am, imass, isweep_total)
    use trial, only: u
    use params
    use comms_common, only: ip_global
    use gaussian
    use thr_capi
    integer, intent(in) :: which
    real, intent(out) :: psibarpsi, aviter
    real, intent(in) :: res, am
    integer, intent(in) :: imass
    integer, intent(in), optional :: isweep_total
    real, target :: ps(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 2)
    real, target :: pt(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 2)
    real(c_double) :: s1(2), s2(2)
    complex(dp) :: psibarpsi1, psibarpsi2
    real(dp) :: pbp(knoise)
    integer(c_int) :: d(4), itn, rc
    integer :: inoise, iter
    real :: susclsing
#ifdef MPI
    integer :: reqs_ps(12), reqs_pt(12)
#endif
    call thr_ensure_init()
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
    iter = 0
    do inoise = 1, knoise
      ! two noise fields per vector, in the reference's order (measure_module.F90:210-238);
      ! the halo requests complete at once on one rank
#ifdef MPI
      call gauss0(ps, reqs_ps)
      call gauss0(pt, reqs_pt)
#else
      call gauss0(ps)
      call gauss0(pt)
#endif
      if (which == 0) then
        rc = thr_measure_shamir_noise(c_loc(ps), c_loc(pt), c_loc(u), d, real(res, c_float), &
                                      real(am, c_float), int(imass, c_int), real(am3, c_float), &
                                      int(niterc, c_int), generation_kernel(), s1, s2, itn)
        call thr_check(rc, 'measure_shamir')
      else
        rc = thr_measure_wilson_noise(c_loc(ps), c_loc(pt), c_loc(u), d, real(res, c_float), &
                                      real(am, c_float), int(imass, c_int), real(am3, c_float), &
                                      int(niterc, c_int), generation_kernel(), s1, s2, itn)
        call thr_check(rc, 'measure_wilson')
      end if
      iter = iter + itn
      psibarpsi1 = cmplx(s1(1), s1(2), dp)
      psibarpsi2 = cmplx(s2(1), s2(2), dp)
      ! measure_module.F90:394-405, 608-620
      if (imass .eq. 1) then
        psibarpsi1 = psibarpsi1/kvol
        psibarpsi2 = psibarpsi2/kvol
      elseif (imass .eq. 3) then
        psibarpsi1 = cmplx(0.0, -1.0)*psibarpsi1/kvol
        psibarpsi2 = cmplx(0.0, +1.0)*psibarpsi2/kvol
      elseif (imass .eq. 5) then
        psibarpsi1 = cmplx(0.0, -1.0)*psibarpsi1/kvol
        psibarpsi2 = cmplx(0.0, -1.0)*psibarpsi2/kvol
      endif
      pbp(inoise) = real(psibarpsi1 + psibarpsi2)
      if (ip_global .eq. 0) then
        open (unit=100, file='fort.100', action='write', position='append')
        if (present(isweep_total)) then
          write (100, '(I5,4E17.9E3)') isweep_total, real(psibarpsi1), aimag(psibarpsi1), &
            real(psibarpsi2), aimag(psibarpsi2)
        else
          write (100, '(4E17.9E3)') real(psibarpsi1), aimag(psibarpsi1), &
            real(psibarpsi2), aimag(psibarpsi2)
        endif
        close (100)
      end if
    end do
    ! measure_module.F90:408-419
    psibarpsi = real(sum(pbp))
    susclsing = 0.0
    do inoise = 1, knoise
      susclsing = susclsing + real(sum(pbp(inoise)*pbp(inoise + 1:knoise)))
    end do
    psibarpsi = psibarpsi/knoise
    susclsing = 2*kvol*susclsing/(knoise*(knoise - 1))
    if (ip_global .eq. 0) then
      open (unit=200, file='fort.200', action='write', position='append')
      if (present(isweep_total)) then
        write (200, '(I5,2E15.7E3)') isweep_total, psibarpsi, susclsing
      else
        write (200, '(2E15.7E3)') psibarpsi, susclsing
      endif
      close (200)
    end if
    aviter = float(iter)/(4*knoise)
  end subroutine noisy_measurement
end module measure_module
