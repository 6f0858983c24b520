! iso_c_binding interfaces to libthirring_hip.so (include/thirring_hip.h).
module thr_capi
  use iso_c_binding
  implicit none

  interface
    integer(c_int) function thr_init(device) bind(C, name="thr_init")
      import :: c_int
      integer(c_int), value :: device
    end function

    integer(c_int) function thr_finalize() bind(C, name="thr_finalize")
      import :: c_int
    end function

    type(c_ptr) function thr_last_error() bind(C, name="thr_last_error")
      import :: c_ptr
    end function

    integer(c_int) function thr_coef(u, theta, dims, ibound) bind(C, name="thr_coef")
      import :: c_int, c_ptr
      type(c_ptr), value :: u, theta
      integer(c_int), intent(in) :: dims(4)
      integer(c_int), value :: ibound
    end function

    integer(c_int) function thr_dslash(phi, r, u, dims, am, imass, am3, kernel) &
        bind(C, name="thr_dslash")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: phi, r, u
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: am, am3
      integer(c_int), value :: imass, kernel
    end function

    integer(c_int) function thr_dslashd(phi, r, u, dims, am, imass, am3, kernel) &
        bind(C, name="thr_dslashd")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: phi, r, u
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: am, am3
      integer(c_int), value :: imass, kernel
    end function

    integer(c_int) function thr_dirac_op_shifted(xout, xin, u, dims, am, imass, am3, shift, &
        kernel) bind(C, name="thr_dirac_op_shifted")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: xout, xin, u
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: am, am3
      real(c_double), value :: shift
      integer(c_int), value :: imass, kernel
    end function

    integer(c_int) function thr_congrad(phi, x, u, dims, res, am, imass, am3, niterc, kernel, &
        itercg, iterations) bind(C, name="thr_congrad")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: phi, x, u
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: res, am, am3
      integer(c_int), value :: imass, niterc, kernel
      integer(c_int), intent(out) :: itercg, iterations
    end function

    integer(c_int) function thr_measure_shamir_noise(ps, pt, u, dims, res, am, imass, am3, niterc, &
        kernel, pbp1, pbp2, iters) bind(C, name="thr_measure_shamir_noise")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: ps, pt, u
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: res, am, am3
      integer(c_int), value :: imass, niterc, kernel
      real(c_double), intent(out) :: pbp1(2), pbp2(2)
      integer(c_int), intent(out) :: iters
    end function

    integer(c_int) function thr_measure_wilson_noise(ps, pt, u, dims, res, am, imass, am3, niterc, &
        kernel, pbp1, pbp2, iters) bind(C, name="thr_measure_wilson_noise")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: ps, pt, u
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: res, am, am3
      integer(c_int), value :: imass, niterc, kernel
      real(c_double), intent(out) :: pbp1(2), pbp2(2)
      integer(c_int), intent(out) :: iters
    end function

    integer(c_int) function thr_multishift(u, dims, am, imass, am3, ndiagq, aden, anum, output, &
        input, res, maxcg, use_sp, kernel, cg_return, cg_returns) bind(C, name="thr_multishift")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: u, aden, anum, output, input, cg_returns
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: am, am3, res
      integer(c_int), value :: imass, ndiagq, maxcg, use_sp, kernel
      integer(c_int), intent(out) :: cg_return
    end function

    integer(c_int) function thr_derivs(r, x2, anum, iflag, dims, am, imass, am3, kernel, dsdpi) &
        bind(C, name="thr_derivs")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: r, x2, dsdpi
      real(c_double), value :: anum
      integer(c_int), intent(in) :: dims(4)
      real(c_float), value :: am, am3
      integer(c_int), value :: iflag, imass, kernel
    end function

    integer(c_int) function thr_qmrherm(phi, x, res, itercg, am, imass, am3, anum, aden, &
        ndiagq, iflag, use_sp, u, dims, max_qmr_iters, kernel, x1, phi0, dsdpi, cg_returns) &
        bind(C, name="thr_qmrherm")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: phi, x, anum, aden, u, x1, phi0, dsdpi, cg_returns
      real(c_float), value :: res, am, am3
      integer(c_int), intent(out) :: itercg
      integer(c_int), value :: imass, ndiagq, iflag, use_sp, max_qmr_iters, kernel
      integer(c_int), intent(in) :: dims(4)
    end function
  end interface

contains

  ! Fortran side of the reference's error convention: print and exit(1)
  ! (qmrherm_module.F90:66-81, dwf3d_lib.F90:14-20)
  subroutine thr_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: k
    if (rc == 0) return
    call c_f_pointer(thr_last_error(), msg, [512])
    text = ''
    do k = 1, 512
      if (msg(k) == c_null_char) exit
      text(k:k) = msg(k)
    end do
    print *, trim(what), ' failed: ', trim(text)
    call exit(1)
  end subroutine thr_check

  ! the reference's kernel switch (GENERATE_WITH_SHAMIR / GENERATE_WITH_WILSON)
  integer(c_int) function generation_kernel()
#ifdef GENERATE_WITH_WILSON
    generation_kernel = 1
#else
    generation_kernel = 0
#endif
  end function generation_kernel

  subroutine thr_ensure_init()
    logical, save :: done = .false.
    if (.not. done) then
      call thr_check(thr_init(0_c_int), 'thr_init')
      done = .true.
    end if
  end subroutine thr_ensure_init

end module thr_capi
