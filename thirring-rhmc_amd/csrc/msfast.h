// Batched mixed-precision multishift with blocked shift updates.
#pragma once
#include "solvers.h"

namespace thr {

// iterations per shift-update block: at most MSB; a solve runs blocks of 32
// or 64 (see msfast_solve)
constexpr int MSB = 64;
static_assert(MSB % 16 == 0, "coefficient rows are updated in chunks of 16");

struct MsfState {
  double delta, omega, omega_save, minus_omega, gammag, alpha, lambda, bnorm2, tol2;
  double zeta_i[MAXSHIFT], zeta_ii[MAXSHIFT], zeta_iii[MAXSHIFT], omegas[MAXSHIFT],
      gammas[MAXSHIFT];
  double ca[MAXSHIFT], cc[MAXSHIFT];           // p_i = ca P_i + sum cb R ; x_i = X_i + cc P_i + sum cd R
  double cb[MAXSHIFT][MSB], cd[MAXSHIFT][MSB];
  int active[MAXSHIFT], touched[MAXSHIFT], cg_returns[MAXSHIFT];
  int n, mini, maxi, jcount, cg_return, maxcg, stop, first_block;
  int nonfinite;  // alpha or lambda came out NaN/Inf: this source stopped there
  // iterations whose scalars the Dslash passes derive themselves: alpha_j and
  // lambda_j of the block's iteration j, replayed by the block's finalize
  double al_rec[MSB], lam_rec[MSB];
};

struct MsfArgs {
  Dims d;
  int nsrc, nshift;
  const double *aden;  // host, nshift
  const float *am;     // host, nsrc
  int imass;
  float am3;
  const double2 *u64;  // device links FP64
  const float2 *u32;   // device links FP32
  const double2 *b;    // device sources [nsrc][4 V5]
  double2 *x;          // device solutions [nsrc][nshift][4 V5]
  double tol;          // per-shift |r_i| <= tol |b|
  int maxcg;
  double th32, th16;   // switch Krylov storage to FP32 / FP16 below these |r|/|b| (<= 0: never)
  int start_prec;      // 0 FP64 (default), 1 FP32, 2 FP16
  const double *aden_src = nullptr;  // host [nsrc][nshift]: per-source shifts (overrides aden)
  const double *tol_src = nullptr;   // host [nsrc]: per-source tolerances (overrides tol)
  bool note_reliable = true;         // record the reliable-update count in the context
};

struct MsfResult {
  int iterations, blocks;
  int cg_return[MAXRHS];
  int cg_returns[MAXRHS * MAXSHIFT];
  // iteration at which the Krylov storage entered FP64 / FP32 / FP16 (-1: never)
  int switches[3];
  int nonfinite;    // sources whose recurrences produced NaN/Inf
  int unconverged;  // (source, shift) pairs still above the iterated target at maxcg
  int reliable;     // reliable updates of the base residual taken
};

int msfast_solve(const MsfArgs &A, MsfResult *res, hipStream_t s);
// the same solve as one persistent cooperative kernel (FP64 storage, two grid
// barriers per iteration); msfast_solve takes it when msfast_persistent_fits
// (THR_PERSIST=1)
bool msfast_persistent_fits(const MsfArgs &A);
int msfast_persistent(const MsfArgs &A, MsfResult *res, hipStream_t s);
// relres[src * nshift + i] = |b - (M^+M + sigma_i) x_i| / |b| in FP64
int msfast_true_residuals(const MsfArgs &A, double *relres, hipStream_t s);
// FP64 defect correction after msfast_solve: every (source, shift) whose true
// residual exceeds tol_true is solved again for that residual (FP32 Krylov
// storage, one shift, batched over pairs) to tol_true / 2 and the correction
// added to x_i; at most max_rounds rounds.  relres (host, nsrc*nshift)
// receives the final true residuals, corr_iters the iterations spent in
// corrections.
int msfast_correct(const MsfArgs &A, int max_rounds, double tol_true, double *relres,
                   int *corr_iters, hipStream_t s);

}  // namespace thr
