// Rational-function driver, RHMC force and Hamiltonian on device fields.
#pragma once
#include "solvers.h"

namespace thr {

int derivs_dispatch(const Dims &d, int kernel, const double2 *R, const double2 *X2, double anum,
                    int iflag, float am, int imass, double *f, hipStream_t s);

// qmrherm_module.F90:20-194 on device fields: phi, x [4 V5]; x1 [n][4 V5];
// phi0 [n][4 V5] (the Phi0 register, halo-free); dsdpi [3 V4].  use_sp: 0
// the reference's FP64 multishift, 1 its FP32 variant, 2 the batched solver
// with mixed-precision storage and FP64 defect correction to true residual
// res |phi| (Shamir kernel; the Wilson kernel always solves in FP64)
int qmrherm_device(const OpDesc &op, const double2 *phi, double2 *x, float res, int *itercg,
                   const double *anum, const double *aden, int n, int iflag, int use_sp,
                   int max_qmr_iters, double2 *x1, double2 *phi0, double *dsdpi,
                   int *cg_returns, hipStream_t s);

struct ForceArgs {
  OpDesc op;             // links, am, imass, am3, kernel
  const double2 *phi;    // device pseudofermion [4 V5]
  const float *theta;    // device theta [3 V4]
  float res1, beta;
  int ibound, ndiagg, use_sp, max_qmr_iters;
  const double *anum4g, *aden4g, *bnum2g, *bden2g, *anum2g, *aden2g;  // host coefficients
};
// force_module.F90:9-113 (rational branch, Nf = 1); dsdpi device [3 V4]
int force_device(const ForceArgs &A, double *dsdpi, int itercg[4], hipStream_t s);

struct HamArgs {
  OpDesc op;
  const double2 *phi;  // device
  const float *theta;  // host theta (3 V4)
  const float *pp;     // host momenta (3 V4)
  float res2, beta;
  int ndiag, max_qmr_iters;
  int use_sp;  // solver of the two qmrherm calls, as qmrherm_device's use_sp
  const double *anum4, *aden4, *bnum2, *bden2;  // host coefficients
};
// dwf3d_lib.F90:511-616 (rational branch, Nf = 1): out = {h, hg, hp, s, hf}
int hamilton_device(const HamArgs &A, double out[5], int itercg[2], hipStream_t s);

// evolution.F90:113-179 (Nf = 1): a half step of pp, then leapfrog steps
// theta += dt pp, links = coef(theta), force, pp -= dt dSdpi; after step k the
// trajectory ends with a closing half step when ytest(ytest_ctx) < 1/iterl
// (the reference draws rano(yran, idum, 1, 1, 1) there), else it runs 4 iterl
// full steps.  theta, pp [3 V4] float, links [3 V4] and dsdpi [3 V4] are
// device arrays updated in place (dsdpi holds the last force); *steps
// receives min(k, 4 iterl).  f.op.u64 and f.theta are replaced by links and theta.
struct EvolveArgs {
  ForceArgs f;
  int iterl;
  float dt;
  float (*ytest)(void *);
  void *ytest_ctx;
};
int evolve_device(const EvolveArgs &A, float *theta, float *pp, double2 *links, double *dsdpi,
                  int *steps, hipStream_t s);

}  // namespace thr
