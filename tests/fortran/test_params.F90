! params module for the Fortran drop-in tests: the interface of the
! reference's params.F90 / tests/test_params.F90 (names, kinds and meaning),
! single-rank decomposition, lattice 12^3 x Ls=24 as the reference's unit tests.
module params
  implicit none
  save
  integer, parameter :: k4b = selected_int_kind(9)
  integer, parameter :: dp = kind(1.d0)
  integer, parameter :: sp = kind(1.)
  integer, parameter :: ksize = 12, ksizet = 12, kthird = 24
  integer, parameter :: kvol = ksize*ksize*ksizet
  integer, parameter :: ndiag = 25, ndiagg = 12
  integer, parameter :: Nf = 1
  real(dp), parameter :: akappa = 0.5d0
  real(sp), parameter :: akappaf = 0.5
  integer, parameter :: np_x = 1, np_y = 1, np_t = 1, np_third = 1
  integer, parameter :: ksizex_l = ksize, ksizey_l = ksize, ksizet_l = ksizet, kthird_l = kthird
  integer, parameter :: kvol_l = kvol
  integer :: max_qmr_iters = 2
  integer :: niterc = kthird*kvol
  real, parameter :: respbp = 1.0e-6, rescgg = 1.0e-6
  real, parameter :: rescga = 1e-9, rescgm = 1e-9
  logical, parameter :: spmd = .false.
  real :: beta
  real :: am3
  integer :: ibound
  integer, parameter :: knoise = 10
end module params
