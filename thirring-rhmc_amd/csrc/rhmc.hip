// Rational-function driver and the RHMC force / Hamiltonian on device fields.
//
// qmrherm_device follows qmrherm_module.F90:20-194 (iflag 0..3) with every
// field resident in HBM; force_device follows force_module.F90:9-113 (the
// rational branch, Nf = 1) and hamilton_device dwf3d_lib.F90:511-616.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "msfast.h"
#include "rhmc.h"

namespace thr {

namespace {

inline unsigned nblk(long n) { return (unsigned)((n + BLK - 1) / BLK); }

// x = a0 phi + sum_k a_k x1_k   (qmrherm_module.F90:83, 96-100)
__global__ __launch_bounds__(BLK) void k_lincomb(long ns, int n, double a0, const double *a,
                                                 const double2 *phi, const double2 *x1,
                                                 double2 *x) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= ns) return;
  cx<double> v = a0 * ldc(phi, i);
  for (int k = 0; k < n; k++) v = v + a[k] * ldc(x1, (long)k * ns + i);
  st(x, i, v);
}

__global__ __launch_bounds__(BLK) void k_redot(long ns, const double2 *a, const double2 *b,
                                               double *partial) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  double v = 0.0;
  if (i < ns) {
    const cx<double> p = ldc(a, i), q = ldc(b, i);
    v = p.r * q.r + p.i * q.i;
  }
  const double t = block_sum(v);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(BLK) void k_sum(const double *partial, int n, double *out) {
  const double t = partial_sum(partial, n);
  if (threadIdx.x == 0) *out = t;
}

// dSdpi(:,:,Lt,3) = -dSdpi(:,:,Lt,3) if ibound == -1; dSdpi += beta*Nf*theta
// (force_module.F90:102-110; beta*Nf*theta is a default-real product)
__global__ __launch_bounds__(BLK) void k_force_finish(const Dims d, double *ds, const float *theta,
                                                      float beta, int ibound) {
  const long v4 = d.v4();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v4 * 3) return;
  const int mu = (int)(i / v4);
  const int t = (int)((i % v4) / ((long)d.lx * d.ly));
  double v = ds[i];
  if (ibound == -1 && mu == 2 && t == d.lt - 1) v = -v;
  ds[i] = v + (double)(beta * 1.0f * theta[i]);
}

// pp = pp - real(c*dSdpi) (evolution.F90:130, 171-174): c = 0.5*dt or dt is a
// default-real product, promoted to double by the real(dp) force, and the
// difference is formed in float32
__global__ __launch_bounds__(BLK) void k_pp_step(long n, float *pp, const double *ds, float c) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  pp[i] = pp[i] - (float)((double)c * ds[i]);
}

// thetat = thetat + dt*pp (evolution.F90:152), float32 with a separately
// rounded product as the Fortran statement computes it
__global__ __launch_bounds__(BLK) void k_theta_step(long n, float *theta, const float *pp,
                                                    float dt) {
#pragma clang fp contract(off)
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const float step = dt * pp[i];
  theta[i] = theta[i] + step;
}

}  // namespace

int derivs_dispatch(const Dims &d, int kernel, const double2 *R, const double2 *X2, double anum,
                    int iflag, float am, int imass, double *f, hipStream_t s) {
  if (kernel == 1) {
    void *pb = workspace(450, (size_t)SLICE_PAIR_BYTES * (2 * d.ls + 4));
    THR_CHECK(pb, THR_ERR_HIP, "workspace allocation failed");
    return launch_derivs_wilson(d, R, X2, anum, iflag, am, imass, f, pb, s);
  }
  return launch_derivs(d, R, X2, anum, iflag, f, s);
}

// The shifted solves of qmrherm by the batched solver with mixed-precision
// Krylov storage (FP64, then FP32 below |r|/|b| = 1e-4) and FP64 defect
// correction: every shift reaches FP64 true residual |phi - (M^+M + s_i) x_i|
// <= res |phi| (a relative criterion, unlike the reference's iterated test).
int multishift_mixed(const OpDesc &op, int n, const double *aden, const double2 *phi,
                     double2 *x1, float res, int max_qmr_iters, int *itercg, int *cg_returns,
                     hipStream_t s) {
  const Dims &d = op.d;
  const float2 *u32 = op.u32;
  if (!u32) {
    float2 *w = (float2 *)workspace(453, sizeof(float2) * d.v4() * 3);
    THR_CHECK(w, THR_ERR_HIP, "workspace allocation failed");
    TRY_RC(convert_d2f(d.v4() * 3, op.u64, w, s));
    u32 = w;
  }
  const float am = op.am;
  MsfArgs A{};
  A.d = d;
  A.nsrc = 1;
  A.nshift = n;
  A.aden = aden;
  A.am = &am;
  A.imass = op.imass;
  A.am3 = op.am3;
  A.u64 = op.u64;
  A.u32 = u32;
  A.b = phi;
  A.x = x1;
  A.tol = 0.75 * (double)res;  // iterated target below the true-residual target
  A.maxcg = max_qmr_iters;
  // Krylov storage FP32 below |r|/|b| = th32.  For a loose target (the
  // force's res 1e-8) the switch comes early: the FP64 defect correction
  // only has to reach res (at 32^3 x 64 the force took 7.8 s at th32 = 1e-4,
  // 6.5 s at 1e-2, 5.9 s at 1e-1, all within the same distance of the FP64
  // force).  A tight target (the Hamiltonian's res 1e-10) switches at 1e-4:
  // 1e-2 was 7% faster there but put dH 7e-9 (relative) from the FP64 one
  // instead of 1e-9.  THR_MIXED_TH32 / THR_MIXED_TH16 override the FP32 and
  // FP16 thresholds (0 never).
  A.th32 = (double)res >= 1e-9 ? 0.1 : 1e-4;
  A.th16 = 0.0;
  if (const char *e = std::getenv("THR_MIXED_TH32")) A.th32 = std::atof(e);
  if (const char *e = std::getenv("THR_MIXED_TH16")) A.th16 = std::atof(e);
  MsfResult R{};
  TRY_RC(msfast_solve(A, &R, s));
  THR_CHECK(!R.nonfinite, THR_ERR_NONFINITE,
            "qmrherm (mixed): non-finite dot product (NaN or Inf in Phi or the links)");
  std::vector<double> rr((size_t)n);
  int corr = 0;
  TRY_RC(msfast_correct(A, 4, (double)res, rr.data(), &corr, s));
  if (itercg) *itercg = R.iterations;
  if (cg_returns)
    for (int i = 0; i < n; i++) cg_returns[i] = R.cg_returns[i];
  // every shift must meet the FP64 true residual after the corrections, or
  // the force / Hamiltonian built on it would go on with unconverged shifts
  for (int i = 0; i < n; i++)
    THR_CHECK(rr[i] <= (double)res, THR_ERR_NOCONV,
              "qmrherm (mixed): shift " + std::to_string(i + 1) + " true residual " +
                  std::to_string(rr[i]) + " above res after the correction rounds");
  return THR_OK;
}

int qmrherm_device(const OpDesc &op, const double2 *phi, double2 *x, float res, int *itercg,
                   const double *anum, const double *aden, int n, int iflag, int use_sp,
                   int max_qmr_iters, double2 *x1, double2 *phi0, double *dsdpi,
                   int *cg_returns, hipStream_t s) {
  const Dims &d = op.d;
  const long ns = d.nspinor();
  THR_CHECK(n >= 1 && n <= 25, THR_ERR_ARG,
            "qmrherm requires ndiagq <= ndiag = 25 (qmrherm_module.F90:66-73)");
  THR_CHECK(d.ls % 4 == 0, THR_ERR_ARG,
            "QMRHERM: kthird_l must be divisible by 4 (qmrherm_module.F90:75-81)");
  THR_CHECK(iflag >= 0 && iflag <= 3, THR_ERR_ARG, "iflag must be 0..3");
  THR_CHECK((iflag != 1 && iflag != 3) || phi0, THR_ERR_ARG, "iflag 1/3 need the Phi0 register");
  THR_CHECK(iflag < 2 || dsdpi, THR_ERR_ARG, "iflag 2/3 need dSdpi");
  THR_CHECK(use_sp >= 0 && use_sp <= 2, THR_ERR_ARG, "use_sp must be 0, 1 or 2");
  // single precision is not available with the Wilson kernel (54-63)
  if (use_sp == 2 && op.kernel == 0)
    TRY_RC(multishift_mixed(op, n, aden, phi, x1, res, max_qmr_iters, itercg, cg_returns, s));
  else if (use_sp == 1 && op.kernel == 0)
    TRY_RC(multishift_reference_sp(op, n, aden, anum + 1, phi, x1, res, max_qmr_iters, itercg,
                                   cg_returns, s));
  else
    TRY_RC(multishift_reference(op, n, aden, anum + 1, phi, x1, res, max_qmr_iters, itercg,
                                cg_returns, s));
  double *ad = (double *)workspace(451, sizeof(double) * 32);
  THR_CHECK(ad, THR_ERR_HIP, "workspace allocation failed");
  if (iflag < 2) {
    THR_HIP(hipMemcpyAsync(ad, anum + 1, sizeof(double) * n, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_lincomb, dim3(nblk(ns)), dim3(BLK), 0, s, ns, n, anum[0], ad, phi, x1, x);
    THR_HIP(hipGetLastError());
    if (iflag == 1)
      THR_HIP(hipMemcpyAsync(phi0, x1, sizeof(double2) * ns * n, hipMemcpyDeviceToDevice, s));
    THR_HIP(hipStreamSynchronize(s));  // anum is a host array
    return THR_OK;
  }
  // x = anum(0)*Phi only (83)
  hipLaunchKernelGGL(k_lincomb, dim3(nblk(ns)), dim3(BLK), 0, s, ns, 0, anum[0], ad, phi, x1, x);
  THR_HIP(hipGetLastError());
  if (op.kernel == 0 && d.ls <= 64 && n <= MAXRHS) {
    // Shamir kernel: M x1(k) (and M Phi0(k)) of every pole in one multi-RHS
    // Dslash, then every pole's derivs terms in one pass (same terms as the
    // loop below; dSdpi receives their sum once)
    double2 *mx = (double2 *)workspace(454, sizeof(double2) * ns * n);
    double2 *mr = iflag == 3 ? (double2 *)workspace(455, sizeof(double2) * ns * n) : nullptr;
    DerivPair *pd = (DerivPair *)workspace(456, sizeof(DerivPair) * 2 * n);
    DerivPair *ph = (DerivPair *)host_workspace(11, sizeof(DerivPair) * 2 * n);
    double *part = (double *)workspace(457, sizeof(double) * 3 * d.v4() * 2 * n);
    THR_CHECK(mx && pd && ph && part && (iflag == 2 || mr), THR_ERR_HIP,
              "workspace allocation failed");
    DslashLaunch L{};
    L.d = d;
    L.nrhs = n;
    L.prec = P64;
    L.imass = op.imass;
    L.am3 = op.am3;
    for (int k = 0; k < n; k++) L.am[k] = op.am;
    L.links = op.u64;
    L.mode = EPI_STORE;
    L.sc_stride = 1;
    L.done_stride = 1;
    L.in = x1;
    L.out = mx;
    TRY_RC(launch_dslash(L, s));  // x2 = M x1(k) (143)
    int np = 0;
    if (iflag == 2) {
      for (int k = 0; k < n; k++)
        ph[np++] = derivs_pair(x1 + (long)k * ns, mx + (long)k * ns, anum[k + 1], 0);
    } else {
      L.in = phi0;
      L.out = mr;
      TRY_RC(launch_dslash(L, s));  // M R with R = Phi0(k) (159, 166)
      for (int k = 0; k < n; k++) {
        const double coeff = -anum[k + 1];
        ph[np++] = derivs_pair(phi0 + (long)k * ns, mx + (long)k * ns, coeff, 0);
        ph[np++] = derivs_pair(mr + (long)k * ns, x1 + (long)k * ns, coeff, 1);
      }
    }
    return launch_derivs_batch(d, np, ph, pd, part, dsdpi, s);
  }
  double2 *x2 = (double2 *)workspace(452, sizeof(double2) * ns);
  THR_CHECK(x2, THR_ERR_HIP, "workspace allocation failed");
  for (int k = 0; k < n; k++) {
    const double2 *xk = x1 + (long)k * ns;
    TRY_RC(apply_op(op, 0, xk, x2, s));  // x2 = M x1(k) (143)
    if (iflag == 2) {
      TRY_RC(derivs_dispatch(d, op.kernel, xk, x2, anum[k + 1], 0, op.am, op.imass, dsdpi, s));
    } else {
      const double coeff = -anum[k + 1];
      const double2 *R = phi0 + (long)k * ns;  // R = Phi0(k) (159)
      TRY_RC(derivs_dispatch(d, op.kernel, R, x2, coeff, 0, op.am, op.imass, dsdpi, s));
      TRY_RC(apply_op(op, 0, R, x2, s));  // x2 = M R (166)
      TRY_RC(derivs_dispatch(d, op.kernel, x2, xk, coeff, 1, op.am, op.imass, dsdpi, s));
    }
  }
  return THR_OK;
}

int force_device(const ForceArgs &A, double *dsdpi, int itercg[4], hipStream_t s) {
  const Dims &d = A.op.d;
  const long ns = d.nspinor();
  double2 *X2 = (double2 *)workspace(460, sizeof(double2) * ns);
  double2 *Xr = (double2 *)workspace(461, sizeof(double2) * ns);
  double2 *x1 = (double2 *)workspace(462, sizeof(double2) * ns * A.ndiagg);
  double2 *phi0 = (double2 *)workspace(463, sizeof(double2) * ns * A.ndiagg);
  THR_CHECK(X2 && Xr && x1 && phi0, THR_ERR_HIP, "workspace allocation failed");
  THR_HIP(hipMemsetAsync(dsdpi, 0, sizeof(double) * d.v4() * 3, s));  // dSdpi = 0.0 (52)
  OpDesc one = A.op;  // qmrherm(..., One, 1, ...): am = 1.0, imass = 1
  one.am = 1.0f;
  one.imass = 1;
  OpDesc mop = A.op;
  int it[4] = {0, 0, 0, 0};
  THR_HIP(hipMemcpyAsync(X2, A.phi, sizeof(double2) * ns, hipMemcpyDeviceToDevice, s));
  TRY_RC(qmrherm_device(one, X2, Xr, A.res1, &it[0], A.anum4g, A.aden4g, A.ndiagg, 1, A.use_sp,
                        A.max_qmr_iters, x1, phi0, dsdpi, nullptr, s));  // (80)
  THR_HIP(hipMemcpyAsync(X2, Xr, sizeof(double2) * ns, hipMemcpyDeviceToDevice, s));
  TRY_RC(qmrherm_device(mop, X2, Xr, A.res1, &it[1], A.bnum2g, A.bden2g, A.ndiagg, 0, A.use_sp,
                        A.max_qmr_iters, x1, nullptr, dsdpi, nullptr, s));  // (85)
  THR_HIP(hipMemcpyAsync(X2, Xr, sizeof(double2) * ns, hipMemcpyDeviceToDevice, s));
  TRY_RC(qmrherm_device(mop, X2, Xr, A.res1, &it[2], A.anum2g, A.aden2g, A.ndiagg, 2, A.use_sp,
                        A.max_qmr_iters, x1, nullptr, dsdpi, nullptr, s));  // (91)
  TRY_RC(qmrherm_device(one, X2, Xr, A.res1, &it[3], A.anum4g, A.aden4g, A.ndiagg, 3, A.use_sp,
                        A.max_qmr_iters, x1, phi0, dsdpi, nullptr, s));  // (96)
  hipLaunchKernelGGL(k_force_finish, dim3(nblk(d.v4() * 3)), dim3(BLK), 0, s, d, dsdpi, A.theta,
                     A.beta, A.ibound);
  THR_HIP(hipGetLastError());
  THR_HIP(hipStreamSynchronize(s));
  if (itercg)
    for (int k = 0; k < 4; k++) itercg[k] = it[k];
  return THR_OK;
}

int hamilton_device(const HamArgs &A, double out[5], int itercg[2], hipStream_t s) {
  const Dims &d = A.op.d;
  const long ns = d.nspinor();
  const long n4 = d.v4() * 3;
  // hp = 0.5*sum(pp**2), hg = 0.5*Nf*beta*sum(theta**2): default-real sums
  // (dwf3d_lib.F90:542, 551), accumulated in order on the host
  float spp = 0.0f, sth = 0.0f;
  for (long k = 0; k < n4; k++) spp += A.pp[k] * A.pp[k];
  for (long k = 0; k < n4; k++) sth += A.theta[k] * A.theta[k];
  const double hp = (double)(0.5f * spp);
  const double hg = (double)(0.5f * 1.0f * A.beta * sth);
  double2 *R = (double2 *)workspace(470, sizeof(double2) * ns);
  double2 *Xr = (double2 *)workspace(471, sizeof(double2) * ns);
  double2 *x1 = (double2 *)workspace(472, sizeof(double2) * ns * A.ndiag);
  const int nb = (int)nblk(ns);
  double *partial = (double *)workspace(473, sizeof(double) * nb);
  double *dev = (double *)workspace(474, sizeof(double));
  THR_CHECK(R && Xr && x1 && partial && dev, THR_ERR_HIP, "workspace allocation failed");
  OpDesc one = A.op;
  one.am = 1.0f;
  one.imass = 1;
  int it1 = 0, it2 = 0;
  THR_HIP(hipMemcpyAsync(R, A.phi, sizeof(double2) * ns, hipMemcpyDeviceToDevice, s));
  TRY_RC(qmrherm_device(one, R, Xr, A.res2, &it1, A.anum4, A.aden4, A.ndiag, 0, A.use_sp,
                        A.max_qmr_iters, x1, nullptr, nullptr, nullptr, s));  // (584)
  THR_HIP(hipMemcpyAsync(R, Xr, sizeof(double2) * ns, hipMemcpyDeviceToDevice, s));
  TRY_RC(qmrherm_device(A.op, R, Xr, A.res2, &it2, A.bnum2, A.bden2, A.ndiag, 0, A.use_sp,
                        A.max_qmr_iters, x1, nullptr, nullptr, nullptr, s));  // (590)
  hipLaunchKernelGGL(k_redot, dim3(nb), dim3(BLK), 0, s, ns, R, Xr, partial);
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(BLK), 0, s, partial, nb, dev);
  THR_HIP(hipGetLastError());
  double hf = 0.0;
  THR_HIP(hipMemcpyAsync(&hf, dev, sizeof(double), hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  hf = 0.0 + hf;
  out[0] = hg + hp + hf;  // h
  out[1] = hg;
  out[2] = hp;
  out[3] = hg + hf;  // s
  out[4] = hf;
  if (itercg) {
    itercg[0] = it1;
    itercg[1] = it2;
  }
  return THR_OK;
}

int evolve_device(const EvolveArgs &A, float *theta, float *pp, double2 *links, double *dsdpi,
                  int *steps, hipStream_t s) {
  const Dims &d = A.f.op.d;
  const long n = d.v4() * 3;
  THR_CHECK(A.iterl >= 1, THR_ERR_ARG, "iterl must be positive");
  THR_CHECK(A.ytest, THR_ERR_ARG, "a uniform source for the trajectory end is required");
  ForceArgs F = A.f;
  F.op.u64 = links;
  F.op.u32 = nullptr;
  F.theta = theta;
  const dim3 g(nblk(n));
  // half step forward for p (126-130)
  TRY_RC(force_device(F, dsdpi, nullptr, s));
  hipLaunchKernelGGL(k_pp_step, g, dim3(BLK), 0, s, n, pp, dsdpi, 0.5f * A.dt);
  THR_HIP(hipGetLastError());
  const float proby = 1.0f / (float)A.iterl;
  int iter = 1;
  for (; iter <= 4 * A.iterl; iter++) {
    hipLaunchKernelGGL(k_theta_step, g, dim3(BLK), 0, s, n, theta, pp, A.dt);  // (152)
    THR_HIP(hipGetLastError());
    TRY_RC(launch_coef(d, theta, F.ibound, links, s));  // (155)
    TRY_RC(force_device(F, dsdpi, nullptr, s));          // (156)
    const float ytest = A.ytest(A.ytest_ctx);            // (159-161)
    if (ytest < proby) {
      hipLaunchKernelGGL(k_pp_step, g, dim3(BLK), 0, s, n, pp, dsdpi, 0.5f * A.dt);  // (167)
      THR_HIP(hipGetLastError());
      break;
    }
    hipLaunchKernelGGL(k_pp_step, g, dim3(BLK), 0, s, n, pp, dsdpi, A.dt);  // (170)
    THR_HIP(hipGetLastError());
  }
  THR_HIP(hipStreamSynchronize(s));
  if (steps) *steps = std::min(iter, 4 * A.iterl);  // itot increment (178)
  return THR_OK;
}

}  // namespace thr
