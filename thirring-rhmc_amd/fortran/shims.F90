! Fortran drop-in modules for j-ostmeyer/thirring-rhmc's hot path.
!
! Same module names, procedure names, argument lists and array shapes as the
! reference (dirac.F90, multishift_module.F90,
! qmrherm_module.F90, derivs_module.F90, gauge.F90, inverter_utils.F90 and the
! state modules trial / vector / gforce / dum1 / gammamatrices), compiled
! against the caller's own `params` module (built with -DMPI and
! NP_X = NP_Y = NP_T = NP_THIRD = 1, the single-rank decomposition).  The
! numerical work runs on the GPU through libthirring_hip.so; halo exchange is
! a periodic host copy and MPI requests are inert.
!
! Module order follows the dependencies.  `measure_module` (congrad, measure,
! measure_shamir) is in measure_shim.F90, compiled after the caller's
! `random` and `gaussian` modules.

module comms_common
  implicit none
  save
  ! one process owns the whole lattice (comms_common.F90, comms.F90:38-95)
  integer :: ip_x = 0, ip_y = 0, ip_t = 0, ip_third = 0, ip_global = 0, np_global = 1
  integer :: comm = 0, comm_grp_third = 0
  integer :: mpiio_type = 0
end module comms_common

module gammamatrices
  use params
  implicit none
  complex(dp) :: gamval(6, 4)
  complex(sp) :: gamvalf(6, 4)
  integer :: gamin(6, 4)
contains
  pure integer function kdelta(nu, mu)
    integer, intent(in) :: nu, mu
    kdelta = merge(1, 0, nu == mu)
  end function kdelta

  ! Dirac algebra (gammamatrices.F90:17-132): row alpha of gamma_mu has the
  ! single entry gamval(mu,alpha) in column gamin(mu,alpha); kappa-scaled.
  subroutine init_gammas()
    complex(dp), parameter :: o = (1.0_dp, 0.0_dp), z = (0.0_dp, 1.0_dp)
    integer :: mu
    gamval(1, :) = [-z, -z, z, z]
    gamval(2, :) = [-o, o, o, -o]
    gamval(3, :) = [-z, z, z, -z]
    gamval(4, :) = [o, o, -o, -o]
    gamval(5, :) = [-o, -o, -o, -o]
    gamval(6, :) = [-o, -o, o, o]
    gamin(1, :) = [4, 3, 2, 1]
    gamin(2, :) = [4, 3, 2, 1]
    gamin(3, :) = [3, 4, 1, 2]
    gamin(4, :) = [1, 2, 3, 4]
    gamin(5, :) = [3, 4, 1, 2]
    gamin(6, :) = [3, 4, 1, 2]
    do mu = 1, 6
      gamvalf(mu, :) = cmplx(gamval(mu, :), kind=sp)
    end do
    gamval = gamval*akappa
    gamvalf = gamvalf*akappaf
  end subroutine init_gammas
end module gammamatrices

module trial
  use params
  implicit none
  save
  complex(dp), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
  real :: theta(ksizex_l, ksizey_l, ksizet_l, 3)
  real :: pp(ksizex_l, ksizey_l, ksizet_l, 3)
end module trial

module vector
  use params
  implicit none
  save
  complex(dp), target :: X(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
end module vector

module gforce
  use params
  implicit none
  save
  real(dp), target :: dSdpi(ksizex_l, ksizey_l, ksizet_l, 3)
end module gforce

module dum1
  use params
  implicit none
  save
  complex(dp) :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
  real :: ps(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 2)
end module dum1

module comms
  use comms_common
  use params
  use thr_capi
  implicit none
contains
  subroutine init_MPI()
    ip_x = 0; ip_y = 0; ip_t = 0; ip_third = 0; ip_global = 0; np_global = 1
    call thr_ensure_init()
  end subroutine init_MPI

  subroutine init_halo_types()
  end subroutine init_halo_types

  subroutine complete_halo_update(reqs)
    integer, intent(inout) :: reqs(16)
    reqs = 0
  end subroutine complete_halo_update
end module comms

module comms4
  use params
  implicit none
contains
  subroutine update_halo_4(size4, Array)
    integer, intent(in) :: size4
    complex(dp), intent(inout) :: Array(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, size4)
    Array(0, :, :, :) = Array(ksizex_l, :, :, :)
    Array(ksizex_l + 1, :, :, :) = Array(1, :, :, :)
    Array(:, 0, :, :) = Array(:, ksizey_l, :, :)
    Array(:, ksizey_l + 1, :, :) = Array(:, 1, :, :)
    Array(:, :, 0, :) = Array(:, :, ksizet_l, :)
    Array(:, :, ksizet_l + 1, :) = Array(:, :, 1, :)
  end subroutine update_halo_4

  subroutine update_halo_4_real(size4, Array)
    integer, intent(in) :: size4
    real, intent(inout) :: Array(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, size4)
    Array(0, :, :, :) = Array(ksizex_l, :, :, :)
    Array(ksizex_l + 1, :, :, :) = Array(1, :, :, :)
    Array(:, 0, :, :) = Array(:, ksizey_l, :, :)
    Array(:, ksizey_l + 1, :, :) = Array(:, 1, :, :)
    Array(:, :, 0, :) = Array(:, :, ksizet_l, :)
    Array(:, :, ksizet_l + 1, :) = Array(:, :, 1, :)
  end subroutine update_halo_4_real

  subroutine start_halo_update_4(size4, Array, tag, reqs)
    integer, intent(in) :: size4, tag
    complex(dp), intent(inout) :: Array(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, size4)
    integer, intent(out) :: reqs(12)
    call update_halo_4(size4, Array)
    reqs = 0
  end subroutine start_halo_update_4

  subroutine start_halo_update_4_real(size4, Array, tag, reqs)
    integer, intent(in) :: size4, tag
    real, intent(inout) :: Array(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, size4)
    integer, intent(out) :: reqs(12)
    call update_halo_4_real(size4, Array)
    reqs = 0
  end subroutine start_halo_update_4_real
end module comms4

module comms5
  use params
  implicit none
contains
  subroutine update_halo_5(size5, Array)
    integer, intent(in) :: size5
    complex(dp), intent(inout) :: Array(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                        0:ksizet_l + 1, size5)
    Array(0, :, :, :, :) = Array(kthird_l, :, :, :, :)
    Array(kthird_l + 1, :, :, :, :) = Array(1, :, :, :, :)
    Array(:, 0, :, :, :) = Array(:, ksizex_l, :, :, :)
    Array(:, ksizex_l + 1, :, :, :) = Array(:, 1, :, :, :)
    Array(:, :, 0, :, :) = Array(:, :, ksizey_l, :, :)
    Array(:, :, ksizey_l + 1, :, :) = Array(:, :, 1, :, :)
    Array(:, :, :, 0, :) = Array(:, :, :, ksizet_l, :)
    Array(:, :, :, ksizet_l + 1, :) = Array(:, :, :, 1, :)
  end subroutine update_halo_5

  subroutine start_halo_update_5(size5, Array, tag, reqs)
    integer, intent(in) :: size5, tag
    complex(dp), intent(inout) :: Array(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                        0:ksizet_l + 1, size5)
    integer, intent(out) :: reqs(16)
    call update_halo_5(size5, Array)
    reqs = 0
  end subroutine start_halo_update_5

  subroutine init_halo_update_5(size5, Array, tag, reqs)
    integer, intent(in) :: size5, tag
    complex(dp), intent(inout) :: Array(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                        0:ksizet_l + 1, size5)
    integer, intent(out) :: reqs(16)
    reqs = 0
  end subroutine init_halo_update_5
end module comms5

module comms6
  use params
  use comms5, only: update_halo_5
  implicit none
contains
  subroutine update_halo_6(size5, size6, Array)
    integer, intent(in) :: size5, size6
    complex(dp), intent(inout) :: Array(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                        0:ksizet_l + 1, size5, size6)
    integer :: k
    do k = 1, size6
      call update_halo_5(size5, Array(:, :, :, :, :, k))
    end do
  end subroutine update_halo_6

  subroutine start_halo_update_6(size5, size6, Array, tag, reqs)
    integer, intent(in) :: size5, size6, tag
    complex(dp), intent(inout) :: Array(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                        0:ksizet_l + 1, size5, size6)
    integer, intent(out) :: reqs(16)
    call update_halo_6(size5, size6, Array)
    reqs = 0
  end subroutine start_halo_update_6
end module comms6

module dirac
  use comms_common
  use params
  use gammamatrices
  use thr_capi
  implicit none
  save
contains
  subroutine dims_of(d)
    integer(c_int), intent(out) :: d(4)
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
  end subroutine dims_of

  subroutine run_op(Phi, R, u, am, imass, kernel, dag)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), target, intent(inout) :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                              0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    integer, intent(in) :: imass, kernel
    real, intent(in) :: am
    logical, intent(in) :: dag
    integer(c_int) :: d(4)
    call thr_ensure_init()
    call dims_of(d)
    if (dag) then
      call thr_check(thr_dslashd(c_loc(Phi), c_loc(R), c_loc(u), d, real(am, c_float), &
                                 int(imass, c_int), real(am3, c_float), int(kernel, c_int)), 'dslashd')
    else
      call thr_check(thr_dslash(c_loc(Phi), c_loc(R), c_loc(u), d, real(am, c_float), &
                                int(imass, c_int), real(am3, c_float), int(kernel, c_int)), 'dslash')
    end if
  end subroutine run_op

  ! dirac.F90:11-27
  subroutine dslash(Phi, R, u, am, imass)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), intent(inout), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                              0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    integer, intent(in) :: imass
    real, intent(in) :: am
    call run_op(Phi, R, u, am, imass, int(generation_kernel()), .false.)
  end subroutine dslash

  ! dirac.F90:29-179
  subroutine dslash_shamir(Phi, R, u, am, imass)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), intent(inout), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                              0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    integer, intent(in) :: imass
    real, intent(in) :: am
    call run_op(Phi, R, u, am, imass, 0, .false.)
  end subroutine dslash_shamir

  ! dirac.F90:226-299
  subroutine dslash_wilson(Phi, R, u, am, imass)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), intent(inout), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                              0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    integer, intent(in) :: imass
    real, intent(in) :: am
    call run_op(Phi, R, u, am, imass, 1, .false.)
  end subroutine dslash_wilson

  ! dirac.F90:409-431 (the MPI request argument has no meaning on one device)
  subroutine dslashd(Phi, R, u, am, imass, reqs_R)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), intent(inout), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                              0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    integer, intent(in) :: imass
    real, intent(in) :: am
    integer, dimension(16), intent(inout), optional :: reqs_R
    call run_op(Phi, R, u, am, imass, int(generation_kernel()), .true.)
  end subroutine dslashd

  ! dirac.F90:433-494
  subroutine dslashd_shamir(Phi, R, u, am, imass, reqs_R)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), intent(inout), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                              0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    integer, intent(in) :: imass
    real, intent(in) :: am
    integer, dimension(16), intent(inout), optional :: reqs_R
    call run_op(Phi, R, u, am, imass, 0, .true.)
  end subroutine dslashd_shamir

  ! dirac.F90:538-625
  subroutine dslashd_wilson(Phi, R, u, am, imass, reqs_R)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    complex(dp), intent(inout), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                              0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    integer, intent(in) :: imass
    real, intent(in) :: am
    integer, dimension(16), intent(inout), optional :: reqs_R
    call run_op(Phi, R, u, am, imass, 1, .true.)
  end subroutine dslashd_wilson
end module dirac

module gauge
  use params
  use thr_capi
  implicit none
  save
  real :: theta(ksizex_l, ksizey_l, ksizet_l, 3)
contains
  ! gauge.F90:12-43
  subroutine coef(u, theta)
    complex(dp), intent(out), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    real, intent(in), target :: theta(ksizex_l, ksizey_l, ksizet_l, 3)
    integer(c_int) :: d(4)
    call thr_ensure_init()
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
    call thr_check(thr_coef(c_loc(u), c_loc(theta), d, int(ibound, c_int)), 'coef')
  end subroutine coef
end module gauge

module inverter_utils
  use params
  use thr_capi
  implicit none
  complex(dp) :: xout(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
contains
  ! inverter_utils.F90:9-18
  subroutine dirac_op_shifted(xout, xin, u, am, imass, shift)
    complex(dp), intent(out), target :: xout(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                             0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: xin(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                           0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    real, intent(in) :: am, shift
    integer, intent(in) :: imass
    integer(c_int) :: d(4)
    call thr_ensure_init()
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
    call thr_check(thr_dirac_op_shifted(c_loc(xout), c_loc(xin), c_loc(u), d, real(am, c_float), &
                                        int(imass, c_int), real(am3, c_float), &
                                        real(shift, c_double), generation_kernel()), &
                   'dirac_op_shifted')
  end subroutine dirac_op_shifted

  ! inverter_utils.F90:20-54
  subroutine dirac_operator(xout, xin, u, am, imass)
    complex(dp), intent(out), target :: xout(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                             0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: xin(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                           0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    real, intent(in) :: am
    integer, intent(in) :: imass
    call dirac_op_shifted(xout, xin, u, am, imass, 0.0)
  end subroutine dirac_operator
end module inverter_utils

module multishift_module
  implicit none
  integer, parameter :: print_every = 100000
contains
  ! multishift_module.F90:13-284
  subroutine multishift_solver(u, am, imass, ndiagq, aden, anum, output, input, &
                               res, maxcg, cg_return, cg_returns)
    use params
    use thr_capi
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    real, intent(in) :: am
    integer, intent(in) :: imass, ndiagq
    real(dp), intent(in), target :: aden(ndiagq), anum(ndiagq)
    complex(dp), target :: output(kthird_l, ksizex_l, ksizey_l, ksizet_l, 4, ndiag)
    complex(dp), target :: input(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
    real, intent(in) :: res
    integer, intent(in) :: maxcg
    integer, intent(out) :: cg_return
    integer, intent(out), optional :: cg_returns(ndiagq)
    call run_ms(u, am, imass, ndiagq, aden, anum, output, input, res, maxcg, cg_return, &
                cg_returns, 0)
  end subroutine multishift_solver

  ! multishift_module.F90:289-534
  subroutine multishift_solver_sp(udp, am, imass, ndiagq, aden, anum, outputdp, inputdp, &
                                  res, maxcg, cg_return, cg_returns)
    use params
    complex(dp), intent(in), target :: udp(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    real, intent(in) :: am
    integer, intent(in) :: imass, ndiagq
    real(dp), intent(in), target :: aden(ndiagq), anum(ndiagq)
    complex(dp), target :: outputdp(kthird_l, ksizex_l, ksizey_l, ksizet_l, 4, ndiag)
    complex(dp), target :: inputdp(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                   0:ksizet_l + 1, 4)
    real, intent(in) :: res
    integer, intent(in) :: maxcg
    integer, intent(out) :: cg_return
    integer, intent(out), optional :: cg_returns(ndiagq)
    call run_ms(udp, am, imass, ndiagq, aden, anum, outputdp, inputdp, res, maxcg, cg_return, &
                cg_returns, 1)
  end subroutine multishift_solver_sp

  subroutine run_ms(u, am, imass, ndiagq, aden, anum, output, input, res, maxcg, cg_return, &
                    cg_returns, use_sp)
    use params
    use thr_capi
    complex(dp), intent(in), target :: u(0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 3)
    real, intent(in) :: am
    integer, intent(in) :: imass, ndiagq, use_sp
    real(dp), intent(in), target :: aden(ndiagq), anum(ndiagq)
    complex(dp), target :: output(kthird_l, ksizex_l, ksizey_l, ksizet_l, 4, ndiag)
    complex(dp), target :: input(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, 0:ksizet_l + 1, 4)
    real, intent(in) :: res
    integer, intent(in) :: maxcg
    integer, intent(out) :: cg_return
    integer, intent(out), optional :: cg_returns(ndiagq)
    integer(c_int), target :: cgr(ndiagq)
    integer(c_int) :: d(4), ret
    call thr_ensure_init()
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
    call thr_check(thr_multishift(c_loc(u), d, real(am, c_float), int(imass, c_int), &
                                  real(am3, c_float), int(ndiagq, c_int), c_loc(aden), &
                                  c_loc(anum), c_loc(output), c_loc(input), real(res, c_float), &
                                  int(maxcg, c_int), int(use_sp, c_int), generation_kernel(), &
                                  ret, c_loc(cgr)), 'multishift_solver')
    cg_return = ret
    if (present(cg_returns)) cg_returns = cgr
  end subroutine run_ms
end module multishift_module

module derivs_module
  use comms_common
  implicit none
contains
  ! derivs_module.F90:8-26
  subroutine derivs(R, X2, anum, iflag, am, imass)
    use params, only: dp
    use thr_capi, only: generation_kernel
    complex(dp), intent(in) :: R(*), X2(*)
    real(dp), intent(in) :: anum
    integer, intent(in) :: iflag, imass
    real, intent(in) :: am
    call run_derivs(R, X2, anum, iflag, am, imass, int(generation_kernel()))
  end subroutine derivs

  ! derivs_module.F90:28-39
  subroutine derivs_shamir(R, X2, anum, iflag, am, imass)
    use params, only: dp
    complex(dp), intent(in) :: R(*), X2(*)
    real(dp), intent(in) :: anum
    integer, intent(in) :: iflag, imass
    real, intent(in) :: am
    call run_derivs(R, X2, anum, iflag, am, imass, 0)
  end subroutine derivs_shamir

  ! derivs_module.F90:41-97
  subroutine derivs_wilson(R, X2, anum, iflag, am, imass)
    use params, only: dp
    complex(dp), intent(in) :: R(*), X2(*)
    real(dp), intent(in) :: anum
    integer, intent(in) :: iflag, imass
    real, intent(in) :: am
    call run_derivs(R, X2, anum, iflag, am, imass, 1)
  end subroutine derivs_wilson

  subroutine run_derivs(R, X2, anum, iflag, am, imass, kernel)
    use params
    use gforce, only: dSdpi
    use thr_capi
    complex(dp), intent(in), target :: R(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                         0:ksizet_l + 1, 4)
    complex(dp), intent(in), target :: X2(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                          0:ksizet_l + 1, 4)
    real(dp), intent(in) :: anum
    integer, intent(in) :: iflag, imass, kernel
    real, intent(in) :: am
    integer(c_int) :: d(4)
    call thr_ensure_init()
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
    call thr_check(thr_derivs(c_loc(R), c_loc(X2), real(anum, c_double), int(iflag, c_int), d, &
                              real(am, c_float), int(imass, c_int), real(am3, c_float), &
                              int(kernel, c_int), c_loc(dSdpi)), 'derivs')
  end subroutine run_derivs
end module derivs_module

module qmrherm_module
  use params
  implicit none
  complex(dp), target :: x1(kthird_l, ksizex_l, ksizey_l, ksizet_l, 4, ndiag)
  complex(dp), target, save :: Phi0(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                    0:ksizet_l + 1, 4, ndiag)
  logical :: printall, no_yet_warned = .true.
contains
  ! qmrherm_module.F90:20-194
  subroutine qmrherm(Phi, x, res, itercg, am, imass, anum, aden, ndiagq, iflag, use_sp, &
                     cg_returns)
    use comms_common, only: ip_global
    use gforce, only: dSdpi
    use trial, only: u
    use thr_capi
    complex(dp), intent(in), target :: Phi(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                           0:ksizet_l + 1, 4)
    complex(dp), intent(out), target :: x(0:kthird_l + 1, 0:ksizex_l + 1, 0:ksizey_l + 1, &
                                          0:ksizet_l + 1, 4)
    integer, intent(in) :: imass, ndiagq, iflag
    real(dp), intent(in), target :: anum(0:ndiagq), aden(ndiagq)
    real, intent(in) :: res, am
    logical, intent(in), optional :: use_sp
    integer, intent(out) :: itercg
    integer, intent(out), optional :: cg_returns(ndiagq)
    integer(c_int), target :: cgr(ndiagq)
    integer(c_int) :: d(4), it, sp
    ! use_sp is honoured only when present (the reference's
    ! "present(use_sp) .and. use_sp" reads an absent argument)
    sp = 0
    if (present(use_sp)) then
      if (use_sp) sp = 1
    end if
    call thr_ensure_init()
    d = [int(ksizex_l, c_int), int(ksizey_l, c_int), int(ksizet_l, c_int), int(kthird_l, c_int)]
    call thr_check(thr_qmrherm(c_loc(Phi), c_loc(x), real(res, c_float), it, real(am, c_float), &
                               int(imass, c_int), real(am3, c_float), c_loc(anum), c_loc(aden), &
                               int(ndiagq, c_int), int(iflag, c_int), sp, c_loc(u), d, &
                               int(max_qmr_iters, c_int), generation_kernel(), c_loc(x1), &
                               c_loc(Phi0), c_loc(dSdpi), c_loc(cgr)), 'qmrherm')
    itercg = it
    if (present(cg_returns)) cg_returns = cgr
    if (ip_global == 0 .and. printall) then
      if (sp == 1 .and. generation_kernel() == 0) then
        print *, "[SP] Qmrherm iterations,res:", itercg, res
      else
        print *, "[DP] Qmrherm iterations,res:", itercg, res
      end if
    end if
  end subroutine qmrherm
end module qmrherm_module
