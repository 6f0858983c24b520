/*
 * CPU oracle for the Thirring-model domain-wall hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the GPU
 * product path is compared against (tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py).  Product code never links or calls it.
 *
 * It is a from-scratch FP64 C restatement of the reference Fortran
 * (j-ostmeyer/thirring-rhmc) and works on the reference's own haloed,
 * column-major layouts:
 *
 *   spinor  Phi(0:Ls+1, 0:Lx+1, 0:Ly+1, 0:Lt+1, 4)  complex128, Ls fastest
 *   links   u  (0:Lx+1, 0:Ly+1, 0:Lt+1, 3)           complex128
 *   theta      (Lx, Ly, Lt, 3)                       float32
 *   force   dSdpi(Lx, Ly, Lt, 3)                     float64
 *   shifts  x1(Ls, Lx, Ly, Lt, 4, nshift)            complex128, halo-free
 *
 * Complex numbers are passed as interleaved (re, im) doubles.  `dims` is
 * always {Lx, Ly, Lt, Ls}.  Scalars that the reference declares as default
 * `real` (am, am3, res, theta, the derivs coefficient) are float32 here too.
 *
 * Pinning: each function is checked in tests/ against closed forms,
 * invariants and brute-force dense constructions, and congrad against the
 * reference's own known answer (itercg == 27, tests/test_congrad.F90:47).
 */
#ifndef THIRRING_ORACLE_H
#define THIRRING_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

int orc_num_threads(void);
void orc_set_threads(int n);

/* periodic halo fills (uncomms5.f90:6-19 / comms5.F90:135-192 semantics) */
void orc_halo5(double *a, const int dims[4], int nspin);
/* 4-D Wilson operator (mass: -am3 in the kernel) on one haloed slice
 * [T+2][Y+2][X+2] x 4 spins, complex interleaved; dag selects DWilsonD */
void orc_dwilson_slice(double *phi, const double *r, const double *u, const int dims[4],
                       float mass, int dag);
void orc_halo4(double *u, const int dims[4]);

/* u = 1 + i theta, antiperiodic t-links, halo (gauge.F90:12-43) */
void orc_coef(double *u, const float *theta, const int dims[4], int ibound);

/* Shamir DWF operator M and M^dagger (dirac.F90:29-179, 303-494) */
void orc_dslash_shamir(double *phi, const double *r, const double *u, const int dims[4],
                       float am, int imass, float am3);
void orc_dslashd_shamir(double *phi, const double *r, const double *u, const int dims[4],
                        float am, int imass, float am3);
/* Wilson-kernel DWF operator (dirac.F90:181-299, 496-625); imass 1 or 3 */
void orc_dslash_wilson(double *phi, const double *r, const double *u, const int dims[4],
                       float am, int imass, float am3);
void orc_dslashd_wilson(double *phi, const double *r, const double *u, const int dims[4],
                        float am, int imass, float am3);
/* FP32 Shamir operator (generated dirac_sp.F90 semantics) on complex64 arrays */
void orc_dslash_shamir_sp(float *phi, const float *r, const float *u, const int dims[4],
                          float am, int imass, float am3);
void orc_dslashd_shamir_sp(float *phi, const float *r, const float *u, const int dims[4],
                           float am, int imass, float am3);

/* x_out = M^dagger M x_in (+ shift x_in) with halos (inverter_utils.F90:9-54) */
void orc_dirac_op_shifted(double *xout, const double *xin, const double *u, const int dims[4],
                          float am, int imass, float am3, double shift, int kernel);

/* CG on M^dagger M, X in/out (measure_module.F90:9-140).  kernel 0 Shamir, 1 Wilson. */
void orc_congrad(const double *phi, double *x, const double *u, const int dims[4], float res,
                 float am, int imass, float am3, int niterc, int kernel, int *itercg,
                 int *iterations);

/* Jegerlehner multishift CG (multishift_module.F90:13-284) */
void orc_multishift(const double *u, const int dims[4], float am, int imass, float am3,
                    int ndiagq, const double *aden, const double *anum, double *output,
                    const double *input, float res, int maxcg, int kernel, int *cg_return,
                    int *cg_returns);
/* FP32-vector multishift (multishift_module.F90:289-534) */
void orc_multishift_sp(const double *u, const int dims[4], float am, int imass, float am3,
                       int ndiagq, const double *aden, const double *anum, double *output,
                       const double *input, float res, int maxcg, int *cg_return,
                       int *cg_returns);

/* fermion force accumulation (derivs_module.F90:99-196) */
void orc_derivs_shamir(const double *r, const double *x2, double anum, int iflag,
                       const int dims[4], double *dsdpi);
/* Wilson-kernel force (derivs_module.F90:41-97, 198-345) */
void orc_derivs_wilson(const double *r, const double *x2, double anum, int iflag,
                       const int dims[4], float am, int imass, float am3, double *dsdpi);

/* rational-function driver (qmrherm_module.F90:20-194).
 * x1 (halo-free, nshift = ndiagq) and phi0 (haloed, 25 slots) are the module
 * arrays; dsdpi is accumulated for iflag 2/3. */
void orc_qmrherm(const double *phi, double *x, float res, int *itercg, float am, int imass,
                 float am3, const double *anum, const double *aden, int ndiagq, int iflag,
                 int use_sp, const double *u, const int dims[4], int max_qmr_iters, int kernel,
                 double *x1, double *phi0, double *dsdpi, int *cg_returns);

/* RHMC force (force_module.F90:9-113, rational branch, Nf = 1) */
void orc_force(const double *phi, float res1, float am, int imass, float am3, float beta,
               int ibound, const float *theta, const double *u, const int dims[4],
               const double *anum4g, const double *aden4g, const double *bnum2g,
               const double *bden2g, const double *anum2g, const double *aden2g, int ndiagg,
               int use_sp, int max_qmr_iters, int kernel, double *dsdpi);
/* Hamiltonian (dwf3d_lib.F90:511-616, rational branch, Nf = 1);
 * out = {h, hg, hp, s, hf} */
void orc_hamilton(const double *phi, float res2, float am, int imass, float am3, float beta,
                  const float *theta, const float *pp, const double *u, const int dims[4],
                  const double *anum4, const double *aden4, const double *bnum2,
                  const double *bden2, int ndiag, int max_qmr_iters, int kernel, double out[5],
                  int itercg_out[2]);

#ifdef __cplusplus
}
#endif
#endif
