// Batched, mixed-precision multishift CG with blocked shift updates.
//
// Same Krylov recurrences as the reference's Jegerlehner solver
// (multishift_module.F90:103-282), reorganised for HBM bandwidth:
//
//  * alpha = Re<p, M^+M p> = |M p|^2 is taken from the M pass, so omega is
//    known before M^+ runs and the M^+ pass updates r in its epilogue
//    (s = M^+ h is never stored).
//  * The search direction p is never stored: h = M p obeys
//    h_new = M r_new + gamma h_old, so the M pass reads r and the old h and
//    writes h in place (no p update pass: 2 fewer field streams per
//    iteration).
//  * Shift vectors are not touched every iteration.  Every p_i and x_i is a
//    linear combination of its value at the start of a block and the block's
//    residuals R_0..R_{B-1}; the per-iteration recurrences
//        x_i -= omega_i p_i ;  p_i = gamma_i p_i + zeta_i R_j
//    act on the (B+1) FP64 coefficients only, and one flush pass per block
//    applies them, reading each residual once for all shifts.  Shift traffic
//    per iteration drops from 4 streams per active shift to (B + 4 k)/B.
//  * Krylov vectors (residual ring, p, h) are stored in FP64, FP32 or FP16
//    (half2 + per-site scale); precision is lowered as the residual falls
//    (matrix-vector errors may grow as |r_k| shrinks), while all scalars,
//    dot products and the shift solutions x_i, p_i stay FP64.
//  * Up to 32 sources with per-source masses run in one launch each step.
//
// Per-shift convergence: |r_i| = |zeta_i| |r| <= tol |b|.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "msfast.h"

namespace thr {

namespace {

inline unsigned nblk(long n) { return (unsigned)((n + BLK - 1) / BLK); }

template <int P> using TT = typename Store<P>::T;
template <int P> using CC = typename Store<P>::C;

__device__ inline double fin_sum(const double *partial, int n) { return partial_sum(partial, n); }

// one element of a source's field: FP64 stores spin planes (element
// e = spin * V5 + site), FP32 and FP16 the spins of a site together
// ([site][spin], FP16 with a per-site scale)
template <int P> struct VAcc {
  const TT<P> *p;
  const float *nrm;
  long v5;
  __device__ inline cx<CC<P>> ld_at(long e, long site, int spin) const {
    if constexpr (P == P16) {
      const float2 f = __half22float2(p[4 * site + spin]);
      const float sc = nrm[site];
      return {f.x * sc, f.y * sc};
    } else if constexpr (P == P32) {
      return ldc(p, 4 * site + spin);
    } else {
      return ldc(p, e);
    }
  }
};

__global__ __launch_bounds__(BLK) void k_msf_init(MsfState *st, int n, const double *partial,
                                                  int nb, const double *tolv, int maxcg) {
  MsfState *S = st + blockIdx.x;
  const double b2 = fin_sum(partial + (long)blockIdx.x * nb, nb);
  if (threadIdx.x == 0) {
    const double tol = tolv[blockIdx.x];
    S->delta = b2;
    S->bnorm2 = b2;
    S->lambda = b2;
    S->tol2 = tol * tol * b2;
    S->omega = 1.0;
    S->omega_save = 1.0;
    S->minus_omega = -1.0;
    S->gammag = 0.0;
    S->n = n;
    for (int i = 0; i < n; i++) {
      S->zeta_i[i] = 1.0;
      S->zeta_ii[i] = 1.0;
      S->zeta_iii[i] = 1.0;
      S->active[i] = b2 != 0.0 ? 1 : 0;  // x = 0 solves b = 0 exactly
      S->touched[i] = 0;
      S->cg_returns[i] = 0;
      S->ca[i] = 1.0;
      S->cc[i] = 0.0;
      for (int j = 0; j < MSB; j++) S->cb[i][j] = S->cd[i][j] = 0.0;
    }
    S->mini = 1;
    S->maxi = n;
    S->jcount = 0;
    S->cg_return = 0;
    S->maxcg = maxcg;
    S->first_block = 1;
    S->nonfinite = isfinite(b2) ? 0 : 1;
    S->stop = (maxcg <= 0 || b2 == 0.0 || S->nonfinite) ? 1 : 0;
  }
}

// alpha = |h|^2, omega, zeta''' and omega_i (multishift_module.F90:118-135);
// aden holds MAXSHIFT shifts per source.  Thread i updates shift i; every
// shift's arithmetic is the serial recurrence's, so results do not depend on
// the thread count.  Returns false (and stops the source) on a non-finite
// omega.  Block-uniform: every thread of the block calls it.
__device__ bool fin1_step(MsfState *S, const double *aden, double al) {
  const double os = S->omega, gg = S->gammag;
  const double om = -S->delta / al;
  if (!isfinite(om)) {
    // a non-finite alpha can never converge: stop this source before any of
    // its vectors is updated (its later launches exit on the stop flag)
    __syncthreads();
    if (threadIdx.x == 0) {
      S->nonfinite = 1;
      S->stop = 1;
    }
    __syncthreads();
    return false;
  }
  const int i = threadIdx.x;
  if (i >= S->mini - 1 && i < S->maxi && S->active[i]) {
    S->zeta_iii[i] = (S->zeta_i[i] * S->zeta_ii[i] * os) /
                     (om * gg * (S->zeta_i[i] - S->zeta_ii[i]) +
                      S->zeta_i[i] * os * (1.0 - aden[i] * om));
    S->omegas[i] = om * S->zeta_iii[i] / S->zeta_ii[i];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    S->cg_return += 1;
    S->alpha = al;
    S->omega_save = os;
    S->omega = om;
    S->minus_omega = -om;
  }
  __syncthreads();
  return true;
}

// lambda, gamma, gamma_i; block coefficients; convergence (136-176, 230-275).
// Returns false once the source stops (non-finite lambda, every shift
// converged, or maxcg).  Block-uniform.
__device__ bool fin2_step(MsfState *S, double L) {
  __shared__ int s_nmax, s_nmin;
  if (!isfinite(L)) {
    __syncthreads();
    if (threadIdx.x == 0) {
      S->nonfinite = 1;
      S->stop = 1;
    }
    __syncthreads();
    return false;
  }
  if (threadIdx.x == 0) {
    s_nmax = 0;
    s_nmin = S->n + 1;
  }
  __syncthreads();
  const double gam = L / S->delta;
  const double omega = S->omega, tol2 = S->tol2;
  const int j = S->jcount, cgr = S->cg_return;
  const int i = threadIdx.x;
  const bool mine = i >= S->mini - 1 && i < S->maxi && S->active[i];
  __syncthreads();  // every thread has read the scalars before thread 0 writes them
  if (mine) {
    const double zt = S->zeta_iii[i], om = S->omegas[i];
    const double gi = gam * zt * om / (S->zeta_ii[i] * omega);
    S->gammas[i] = gi;
    // x_i -= omega_i p_i  with p_i = ca P + sum cb R ; then
    // p_i = gamma_i p_i + zeta_i''' R_j.  The rows go in chunks of 16 slots:
    // static indices keep a chunk in registers and make its loads independent
    S->cc[i] -= om * S->ca[i];
    S->ca[i] *= gi;
    for (int q0 = 0; q0 < j; q0 += 16) {
      double cbr[16], cdr[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        cbr[u] = S->cb[i][q0 + u];
        cdr[u] = S->cd[i][q0 + u];
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        if (q0 + u < j) {
          S->cd[i][q0 + u] = cdr[u] - om * cbr[u];
          S->cb[i][q0 + u] = cbr[u] * gi;
        }
      }
    }
    S->cb[i][j] = zt;
    S->cd[i][j] = 0.0;
    S->touched[i] = 1;
    S->cg_returns[i] = cgr;
    // shifted residual of the new iterate: |zeta'''| |r_new|
    if (zt * zt * L <= tol2) {
      S->active[i] = 0;
    } else {
      atomicMax(&s_nmax, i + 1);
      atomicMin(&s_nmin, i + 1);
    }
  }
  __syncthreads();
  const int nmax = s_nmax, nmin = s_nmin;
  if (i >= nmin - 1 && i < nmax && S->active[i]) {
    S->zeta_i[i] = S->zeta_ii[i];
    S->zeta_ii[i] = S->zeta_iii[i];
  }
  __syncthreads();
  const bool go = !(nmax == 0 || cgr >= S->maxcg);
  if (threadIdx.x == 0) {
    S->lambda = L;
    S->gammag = gam;
    S->mini = nmin;
    S->maxi = nmax;
    S->delta = L;
    S->jcount = j + 1;
    if (!go) S->stop = 1;
  }
  __syncthreads();
  return go;
}

__global__ __launch_bounds__(BLK) void k_msf_fin1(MsfState *st, const double *adens,
                                                  const double *partial, int stride, int nb) {
  MsfState *S = st + blockIdx.x;
  if (S->stop) return;
  __shared__ double s_al;
  const double al = fin_sum(partial + (long)blockIdx.x * stride, nb);
  if (threadIdx.x == 0) s_al = al;
  __syncthreads();
  fin1_step(S, adens + (long)blockIdx.x * MAXSHIFT, s_al);
}

__global__ __launch_bounds__(BLK) void k_msf_fin2(MsfState *st, const double *partial, int stride,
                                                  int nb) {
  MsfState *S = st + blockIdx.x;
  if (S->stop) return;
  __shared__ double s_lam;
  const double lam = fin_sum(partial + (long)blockIdx.x * stride, nb);
  if (threadIdx.x == 0) s_lam = lam;
  __syncthreads();
  fin2_step(S, s_lam);
}

// The recurrences of a block's iterations replayed from the scalars the
// Dslash passes derived and recorded (al_rec, lam_rec; the last lambda is
// summed here from the last M^dagger pass's partials).  The passes kept
// iterating to the end of the block; the replay stops where the finalize
// kernels would have stopped the source, so x_i, p_i and every count are
// those of the per-iteration finalize.
__global__ __launch_bounds__(BLK) void k_msf_fin_block(MsfState *st, const double *adens,
                                                       const double *lpart, int stride, int nb,
                                                       int nit) {
  MsfState *S = st + blockIdx.x;
  if (S->stop) return;
  __shared__ double s_last;
  const double last = fin_sum(lpart + (long)blockIdx.x * stride, nb);
  if (threadIdx.x == 0) s_last = last;
  __syncthreads();
  const double *aden = adens + (long)blockIdx.x * MAXSHIFT;
  for (int j = 0; j < nit; j++) {
    if (!fin1_step(S, aden, S->al_rec[j])) return;
    if (!fin2_step(S, j + 1 < nit ? S->lam_rec[j] : s_last)) return;
  }
}

// a held residual in FP64.  For FP32 values the empty asm makes the copy
// opaque to the compiler, so the widening stays in the shift loop instead of
// being hoisted out of it as 64 FP64 pairs (twice the registers)
__device__ inline cx<double> widen(const cx<double> &v) { return v; }
__device__ inline cx<double> widen(cx<float> v) {
  asm volatile("" : "+v"(v.r), "+v"(v.i));
  return {(double)v.r, (double)v.i};
}

// apply the block's coefficients of ring slots [q0, q0 + QN) to x_i and p_i
// (FP64), reading each residual once.  A 64-iteration block takes one launch
// with FP32/FP16 residuals (held as such, widened per use) and two with FP64
// ones (q0 = 0, 32: the second adds its slots to the x_i, p_i the first
// wrote).
template <int P, int QN>
__global__ __launch_bounds__(BLK) void k_msf_flush(long v5, int n, TT<P> *const *ring,
                                                   float *const *ring_n, const double2 *b,
                                                   double2 *X, double2 *Pv,
                                                   const MsfState *sts, int final_flush,
                                                   int q0) {
  const int src = blockIdx.y;
  const MsfState *S = sts + src;
  const int jc = S->jcount;
  if (jc <= q0) return;
  __shared__ double sa[MAXSHIFT], sc[MAXSHIFT], sb[MAXSHIFT][QN], sd[MAXSHIFT][QN];
  __shared__ int stouch[MAXSHIFT], sact[MAXSHIFT];
  for (int k = threadIdx.x; k < n; k += BLK) {
    sa[k] = S->ca[k];
    sc[k] = S->cc[k];
    stouch[k] = S->touched[k];
    sact[k] = S->active[k] && !S->stop && !final_flush;
    for (int q = 0; q < QN; q++) {
      sb[k][q] = S->cb[k][q0 + q];
      sd[k][q] = S->cd[k][q0 + q];
    }
  }
  __syncthreads();
  const long ns = 4 * v5;
  const long tau = (long)blockIdx.x * BLK + threadIdx.x;
  if (tau >= ns) return;
  // element e = spin * V5 + site of the FP64 x_i, p_i; FP32 and FP16
  // residuals keep a site's 4 spins together, so there consecutive threads
  // take the 4 spins of one site (the ring reads stay contiguous, the FP64
  // planes 16 lanes wide)
  const long site = tau >> 2;
  const int spin = (int)(tau & 3);
  const long e = P != P64 ? spin * v5 + site : tau;
  const long off = (long)src * ns;
  // 32 slots are held in FP64; 64 in their storage precision (the widening
  // to FP64 is exact, so the sums are the same either way)
  using RT = std::conditional_t<(QN <= 32), double, CC<P>>;
  cx<RT> R[QN];
  if constexpr (P == P16 && QN > 32) {
    // FP16 residuals are 4 bytes per element, so one load per slot and lane
    // moves few bytes per memory instruction.  Instead the 4 lanes of a site
    // (consecutive lanes, spin = lane & 3) each load the whole site (16 B)
    // and scale of one of 4 slots and exchange them through LDS: 2 loads per
    // 4 slots instead of 8.  The values and their scaling are those of
    // VAcc<P16>::ld_at.  A wave only exchanges with itself: LDS operations
    // of a wave run in order, and the wave barriers keep the compiler from
    // moving them.
    __shared__ uint4 xh[BLK / 64][64];
    __shared__ float xn[BLK / 64][64];
    const unsigned wv = threadIdx.x >> 6, ln = threadIdx.x & 63, g = ln & ~3u;
    const long so = (long)src * v5 + site;
#pragma unroll
    for (int b4 = 0; b4 < QN / 4; b4++) {
      const int qb = q0 + 4 * b4;  // slots qb .. qb + 3 (< MSB: all allocated)
      const __half2 *p0 = ring[qb], *p1 = ring[qb + 1], *p2 = ring[qb + 2], *p3 = ring[qb + 3];
      const float *n0 = ring_n[qb], *n1 = ring_n[qb + 1], *n2 = ring_n[qb + 2], *n3 = ring_n[qb + 3];
      const __half2 *pk = spin == 0 ? p0 : spin == 1 ? p1 : spin == 2 ? p2 : p3;
      const float *nk = spin == 0 ? n0 : spin == 1 ? n1 : spin == 2 ? n2 : n3;
      uint4 raw = make_uint4(0u, 0u, 0u, 0u);
      float sk = 0.0f;
      if (qb + spin < jc) {
        raw = *(const uint4 *)(pk + off + 4 * site);
        sk = nk[so];
      }
      __builtin_amdgcn_wave_barrier();
      xh[wv][ln] = raw;
      xn[wv][ln] = sk;
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const uint4 h4 = xh[wv][g + j];
        const unsigned hw = spin == 0 ? h4.x : spin == 1 ? h4.y : spin == 2 ? h4.z : h4.w;
        __half2 hv;
        __builtin_memcpy(&hv, &hw, sizeof(hv));
        const float2 f = __half22float2(hv);
        const float sc = xn[wv][g + j];
        R[4 * b4 + j] = qb + j < jc ? cx<RT>{f.x * sc, f.y * sc} : cx<RT>{RT(0), RT(0)};
      }
      __builtin_amdgcn_wave_barrier();
    }
  } else {
#pragma unroll
    for (int q = 0; q < QN; q++) {
      if (q0 + q < jc) {
        VAcc<P> a{ring[q0 + q] + off, ring_n ? ring_n[q0 + q] + (long)src * v5 : nullptr, v5};
        const cx<CC<P>> v = a.ld_at(e, site, spin);
        R[q] = {(RT)v.r, (RT)v.i};
      } else {
        R[q] = {RT(0), RT(0)};
      }
    }
  }
  const bool head = q0 == 0;
  const int first = S->first_block;
  const cx<double> bv = head && first ? ldc(b, off + e) : cx<double>{0.0, 0.0};
  for (int i = 0; i < n; i++) {
    if (!stouch[i]) continue;
    const long xo = ((long)src * n + i) * ns + e;
    cx<double> xs, ps = {0.0, 0.0};
    if (head) {
      const cx<double> base = first ? bv : ldc(Pv, xo);
      xs = first ? cx<double>{0.0, 0.0} : ldc(X, xo);
      xs = xs + sc[i] * base;
      if (sact[i]) ps = sa[i] * base;
    } else {
      xs = ldc(X, xo);
      if (sact[i]) ps = ldc(Pv, xo);
    }
    if (sact[i]) {
#pragma unroll
      for (int q = 0; q < QN; q++)
        if (q0 + q < jc) {
          const cx<double> r = widen(R[q]);
          xs = xs + sd[i][q] * r;
          ps = ps + sb[i][q] * r;
        }
      st(X, xo, xs);
      st(Pv, xo, ps);
    } else {
#pragma unroll
      for (int q = 0; q < QN; q++)
        if (q0 + q < jc) xs = xs + sd[i][q] * widen(R[q]);
      st(X, xo, xs);
    }
  }
}

__global__ void k_msf_reset(MsfState *st) {
  MsfState *S = st + blockIdx.x;
  if (S->jcount == 0) return;
  const int i = threadIdx.x;
  if (i < S->n) {
    S->ca[i] = 1.0;
    S->cc[i] = 0.0;
    S->touched[i] = 0;
    for (int q = 0; q < MSB; q++) S->cb[i][q] = S->cd[i][q] = 0.0;
  }
  __syncthreads();
  if (i == 0) {
    S->jcount = 0;
    S->first_block = 0;
  }
}

// partial |b - (M^+M + sigma) x|^2 per source: t holds M^+M x
__global__ __launch_bounds__(BLK) void k_true_res(long v5, const double2 *b, const double2 *t,
                                                  const double2 *x, double sigma, long xstride,
                                                  double *partial) {
  const int src = blockIdx.y;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  double red = 0.0;
  if (i < v5) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const long e = (long)src * 4 * v5 + k * v5 + i;
      const cx<double> r = ldc(b, e) - ldc(t, e) - sigma * ldc(x, (long)src * xstride + k * v5 + i);
      red += norm2(r);
    }
  }
  const double tot = block_sum(red);
  if (threadIdx.x == 0) partial[(long)src * gridDim.x + blockIdx.x] = tot;
}

// Reliable update of the base residual (the replacement of the iterated
// residual by a true one, Sleijpen and van der Vorst): for each source with
// rel[src] != 0, r' = (b - (M^+M + sigma) x) / zeta from the FP64 true
// residual of one shift (t holds M^+M x, x that shift's solution, zeta its
// coefficient r_i = zeta r), stored in the Krylov precision P over the
// residual r the next iteration reads.  Every active shift direction holds
// zeta_j r (p_j = gamma_j p_j + zeta_j r at the last iteration), so it gets
// zeta_j (r' - r) added (Pv: [nsrc][n][4 V5]); without that the shifted
// recurrences no longer match the base one and the light shifts stall.
// rel: [nsrc] flags, [nsrc] sigma, [nsrc] 1/zeta, [nsrc][MAXSHIFT] zeta_j of
// the active shifts (0 for the others).
// partial[src * gridDim.x + block] = |r'|^2 of the stored values.
template <int P>
__global__ __launch_bounds__(BLK) void k_msf_reliable(long v5, int n, const double2 *b,
                                                      const double2 *t, const double2 *x,
                                                      const double *rel, TT<P> *r, float *rn,
                                                      double2 *Pv, double *partial) {
  const int src = blockIdx.y, nsrc = gridDim.y;
  const bool on = rel[src] != 0.0;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  double red = 0.0;
  if (on && i < v5) {
    const double sig = rel[nsrc + src], iz = rel[2 * nsrc + src];
    const double *zj = rel + 3 * nsrc + (long)src * MAXSHIFT;
    const long off = (long)src * 4 * v5;
    cx<double> v[4], d[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const long e = off + k * v5 + i;
      v[k] = iz * (ldc(b, e) - ldc(t, e) - sig * ldc(x, e));
    }
    // new and old residual as stored (the process reads the stored values)
    if constexpr (P == P16) {
      const float mo = rn[(long)src * v5 + i];
      float m = 0.f;
#pragma unroll
      for (int k = 0; k < 4; k++) m = fmaxf(m, fmaxf(fabsf((float)v[k].r), fabsf((float)v[k].i)));
      const float inv = m > 0.f ? 1.f / m : 0.f;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const float2 fo = __half22float2(r[off + 4 * i + k]);
        const __half2 h = __floats2half2_rn((float)v[k].r * inv, (float)v[k].i * inv);
        r[off + 4 * i + k] = h;
        const float2 f = __half22float2(h);
        v[k] = {(double)(f.x * m), (double)(f.y * m)};
        d[k] = v[k] - cx<double>{(double)(fo.x * mo), (double)(fo.y * mo)};
      }
      rn[(long)src * v5 + i] = m;
    } else if constexpr (P == P32) {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const long e = off + 4 * i + k;
        const cx<float> fo = ldc(r, e);
        const cx<float> f{(float)v[k].r, (float)v[k].i};
        st(r, e, f);
        v[k] = {(double)f.r, (double)f.i};
        d[k] = v[k] - cx<double>{(double)fo.r, (double)fo.i};
      }
    } else {
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const long e = off + k * v5 + i;
        d[k] = v[k] - ldc(r, e);
        st(r, e, v[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < 4; k++) red += norm2(v[k]);
    for (int j = 0; j < n; j++) {
      const double z = zj[j];
      if (z == 0.0) continue;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const long e = ((long)src * n + j) * 4 * v5 + k * v5 + i;
        st(Pv, e, ldc(Pv, e) + z * d[k]);
      }
    }
  }
  const double tot = block_sum(red);
  if (threadIdx.x == 0) partial[(long)src * gridDim.x + blockIdx.x] = tot;
}

// delta = lambda = |r|^2 of the replaced residual (rel as in k_msf_reliable)
__global__ __launch_bounds__(BLK) void k_msf_reliable_fin(MsfState *st, const double *rel,
                                                          const double *partial, int nb) {
  MsfState *S = st + blockIdx.x;
  const double r2 = fin_sum(partial + (long)blockIdx.x * nb, nb);
  if (threadIdx.x == 0 && rel[blockIdx.x] != 0.0) {
    S->delta = r2;
    S->lambda = r2;
    if (!isfinite(r2)) {
      S->nonfinite = 1;
      S->stop = 1;
    }
  }
}

__global__ __launch_bounds__(BLK) void k_sum_rows(const double *partial, int nb, double *out) {
  const double v = fin_sum(partial + (long)blockIdx.x * nb, nb);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// precision conversion of nsrc batched spinors
template <int SP, int DP>
__global__ __launch_bounds__(BLK) void k_conv(long v5, int nsrc, const TT<SP> *a, const float *an,
                                              TT<DP> *b, float *bn) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  const int src = blockIdx.y;
  if (i >= v5) return;
  const long off = (long)src * 4 * v5;
  cx<float> vf[4];
  cx<double> vd[4];
#pragma unroll
  for (int k = 0; k < 4; k++) {
    if constexpr (SP == P16) {
      const float2 f = __half22float2(a[off + 4 * i + k]);
      const float s = an[(long)src * v5 + i];
      vd[k] = {(double)(f.x * s), (double)(f.y * s)};
    } else {
      const auto v = ldc(a, off + (SP == P32 ? 4 * i + k : k * v5 + i));
      vd[k] = {(double)v.r, (double)v.i};
    }
    vf[k] = {(float)vd[k].r, (float)vd[k].i};
  }
  if constexpr (DP == P16) {
    float m = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++) m = fmaxf(m, fmaxf(fabsf(vf[k].r), fabsf(vf[k].i)));
    const float inv = m > 0.f ? 1.f / m : 0.f;
    bn[(long)src * v5 + i] = m;
#pragma unroll
    for (int k = 0; k < 4; k++) b[off + 4 * i + k] = __floats2half2_rn(vf[k].r * inv, vf[k].i * inv);
  } else if constexpr (DP == P32) {
#pragma unroll
    for (int k = 0; k < 4; k++) st(b, off + 4 * i + k, vf[k]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; k++) st(b, off + k * v5 + i, vd[k]);
  }
}

// out_j = b_j - t_j - sigma_j x_j for a batch of vectors of length ns
__global__ __launch_bounds__(BLK) void k_resvec(long ns, const double2 *b, const double2 *t,
                                                const double2 *x, const double *sig,
                                                double2 *out) {
  const long e = (long)blockIdx.x * BLK + threadIdx.x;
  if (e >= ns) return;
  const long o = (long)blockIdx.y * ns + e;
  st(out, o, ldc(b, o) - ldc(t, o) - sig[blockIdx.y] * ldc(x, o));
}

__global__ __launch_bounds__(BLK) void k_addvec(long ns, double2 *y, const double2 *c) {
  const long e = (long)blockIdx.x * BLK + threadIdx.x;
  if (e >= ns) return;
  st(y, e, ldc(y, e) + ldc(c, e));
}

__global__ __launch_bounds__(BLK) void k_norm_src(long v5, const double2 *x, double *partial) {
  const int src = blockIdx.y;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  double red = 0.0;
  if (i < v5) {
#pragma unroll
    for (int k = 0; k < 4; k++) red += norm2(ldc(x, (long)src * 4 * v5 + k * v5 + i));
  }
  const double t = block_sum(red);
  if (threadIdx.x == 0) partial[(long)src * gridDim.x + blockIdx.x] = t;
}

}  // namespace

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------
static int conv(const Dims &d, int nsrc, int sp, const void *a, const float *an, int dp, void *b,
                float *bn, hipStream_t s) {
  dim3 g(nblk(d.v5()), nsrc);
  const long v5 = d.v5();
#define CV(A, B) hipLaunchKernelGGL((k_conv<A, B>), g, dim3(BLK), 0, s, v5, nsrc, (const TT<A> *)a, an, (TT<B> *)b, bn)
  if (sp == P64 && dp == P64) CV(P64, P64);
  else if (sp == P64 && dp == P32) CV(P64, P32);
  else if (sp == P64 && dp == P16) CV(P64, P16);
  else if (sp == P32 && dp == P64) CV(P32, P64);
  else if (sp == P32 && dp == P32) CV(P32, P32);
  else if (sp == P32 && dp == P16) CV(P32, P16);
  else if (sp == P16 && dp == P64) CV(P16, P64);
  else if (sp == P16 && dp == P32) CV(P16, P32);
  else CV(P16, P16);
#undef CV
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int msfast_solve(const MsfArgs &A, MsfResult *res, hipStream_t s) {
  const Dims &d = A.d;
  const int n = A.nshift, ns_ = A.nsrc;
  THR_CHECK(n >= 1 && n <= MAXSHIFT, THR_ERR_ARG, "number of shifts out of range (1..32)");
  THR_CHECK(ns_ >= 1 && ns_ <= MAXRHS, THR_ERR_ARG, "number of sources out of range (1..32)");
  THR_CHECK(A.imass == 1 || A.imass == 3 || A.imass == 5, THR_ERR_ARG, "imass must be 1, 3 or 5");
  THR_CHECK(A.u64 && A.u32 && A.b && A.x, THR_ERR_ARG, "NULL argument");
  if (msfast_persistent_fits(A)) return msfast_persistent(A, res, s);
  const long v5 = d.v5(), nsp = 4 * v5, nsb = nsp * ns_;
  const int nb = (int)dslash_blocks(d);
  // iterations per block.  Each flush streams x_i and p_i of every shift the
  // block touched once (4 field streams per shift, against one read of each
  // residual), so with many shifts 64-iteration blocks cut the flush traffic
  // per iteration by up to half.  They double the residual ring and may run
  // up to 63 iterations past a source's convergence: solves with fewer than
  // 8 shifts or 2^23 spinor components (all sources; 32^3 x 64 is one
  // source) keep 32.  THR_MSB=32 or 64 forces either.
  int msb = (n >= 8 && nsb >= (1L << 23)) ? 64 : 32;
  if (const char *me = std::getenv("THR_MSB")) {
    const int v = std::atoi(me);
    if (v == 32 || v == 64) msb = v;
  }
  // workspace slots of the Krylov buffers per precision pr: ring slot q,
  // h, r0, and the FP16 scales
  auto w_ring = [](int pr, int q) { return 2000 + pr * MSB + q; };
  auto w_h = [](int pr) { return 1100 + pr; };
  auto w_r0 = [](int pr) { return 1110 + pr; };
  auto w_ringn = [](int q) { return 2200 + q; };
  size_t bytes_p[3] = {16, 8, 4};
  void *ring[3][MSB];
  float *ringn[MSB];
  void *hbuf[3], *r0[3];
  float *hn = nullptr, *r0n = nullptr;
  for (int pr = 0; pr < 3; pr++) {
    // precision only falls during a solve: FP64 buffers only for an FP64 start
    const bool need = (pr == P64 && A.start_prec == P64) || (pr == P32 && A.th32 > 0) ||
                      (pr == P16 && A.th16 > 0);
    for (int q = 0; q < MSB; q++)
      ring[pr][q] = need && q < msb ? workspace(w_ring(pr, q), bytes_p[pr] * nsb) : nullptr;
    hbuf[pr] = need ? workspace(w_h(pr), bytes_p[pr] * nsb) : nullptr;
    r0[pr] = need ? workspace(w_r0(pr), bytes_p[pr] * nsb) : nullptr;
    if (need)
      THR_CHECK(hbuf[pr] && r0[pr], THR_ERR_HIP, "workspace allocation failed");
  }
  if (A.th16 > 0) {
    for (int q = 0; q < MSB; q++)
      ringn[q] = q < msb ? (float *)workspace(w_ringn(q), sizeof(float) * v5 * ns_) : nullptr;
    hn = (float *)workspace(1300, sizeof(float) * v5 * ns_);
    r0n = (float *)workspace(1301, sizeof(float) * v5 * ns_);
  }
  double2 *Pv = (double2 *)workspace(363, sizeof(double2) * nsp * ns_ * n);
  MsfState *st = (MsfState *)workspace(364, sizeof(MsfState) * ns_);
  // per-source shifts [ns_][MAXSHIFT] followed by per-source tolerances [ns_]
  double *aden = (double *)workspace(365, sizeof(double) * (MAXSHIFT + 1) * ns_);
  double *tolv = aden + (long)MAXSHIFT * ns_;
  double *partial = (double *)workspace(366, sizeof(double) * nb * ns_);
  // the M^dagger passes' partial sums when the passes derive their own
  // scalars (the M passes' go to `partial`; a pass reads the other's)
  double *lpart = (double *)workspace(369, sizeof(double) * nb * ns_);
  void **ringptr = (void **)workspace(367, sizeof(void *) * 3 * MSB + sizeof(float *) * MSB);
  double *sums = (double *)workspace(368, sizeof(double) * ns_);
  THR_CHECK(Pv && st && aden && partial && lpart && ringptr && sums, THR_ERR_HIP,
            "workspace allocation failed");
  {
    const int np = 4 * MSB;
    const size_t nad = (size_t)(MAXSHIFT + 1) * ns_;
    void **hp = (void **)host_workspace(0, sizeof(void *) * np);
    double *ad = (double *)host_workspace(1, sizeof(double) * nad);
    THR_CHECK(hp && ad, THR_ERR_HIP, "host workspace allocation failed");
    for (int pr = 0; pr < 3; pr++)
      for (int q = 0; q < MSB; q++) hp[pr * MSB + q] = ring[pr][q];
    for (int q = 0; q < MSB; q++) hp[3 * MSB + q] = A.th16 > 0 ? (void *)ringn[q] : nullptr;
    THR_HIP(hipMemcpyAsync(ringptr, hp, sizeof(void *) * np, hipMemcpyHostToDevice, s));
    for (size_t q = 0; q < nad; q++) ad[q] = 0.0;
    for (int k = 0; k < ns_; k++) {
      for (int i = 0; i < n; i++)
        ad[k * MAXSHIFT + i] = A.aden_src ? A.aden_src[k * n + i] : A.aden[i];
      ad[MAXSHIFT * ns_ + k] = A.tol_src ? A.tol_src[k] : A.tol;
    }
    THR_HIP(hipMemcpyAsync(aden, ad, sizeof(double) * nad, hipMemcpyHostToDevice, s));
    THR_HIP(hipStreamSynchronize(s));
  }
  // |b|^2 per source, state init
  hipLaunchKernelGGL(k_norm_src, dim3(nb, ns_), dim3(BLK), 0, s, v5, A.b, partial);
  hipLaunchKernelGGL(k_msf_init, dim3(ns_), dim3(BLK), 0, s, st, n, partial, nb, tolv, A.maxcg);
  THR_HIP(hipGetLastError());

  THR_HIP(hipMemsetAsync(A.x, 0, sizeof(double2) * nsp * ns_ * n, s));
  // launch shapes are fixed for the whole solve, so the finalize kernels sum
  // exactly the partial entries each precision's Dslash grid writes
  DslashConfig cfg[3];
  int nbl[3];
  for (int pr = 0; pr < 3; pr++) {
    cfg[pr] = dslash_config(d, pr, ns_);
    nbl[pr] = (int)dslash_grid_blocks(d, cfg[pr]);
    THR_CHECK(nbl[pr] <= nb, THR_ERR_STATE, "dslash grid exceeds partial stride");
  }
  int prec = P64;
  if (A.start_prec > 0) prec = A.start_prec;
  if (res) {
    res->switches[0] = res->switches[1] = res->switches[2] = -1;
    res->switches[prec] = 0;
  }
  THR_CHECK(prec == P64 || (prec == P32 && A.th32 > 0) || (prec == P16 && A.th16 > 0), THR_ERR_ARG,
            "start precision needs its buffers (set th32/th16)");
  // r0 and p in the starting precision
  TRY_RC(conv(d, ns_, P64, A.b, nullptr, prec, r0[prec], prec == P16 ? r0n : nullptr, s));

  DslashLaunch L{};
  L.d = d;
  L.nrhs = ns_;
  L.imass = A.imass;
  L.am3 = A.am3;
  for (int k = 0; k < ns_; k++) L.am[k] = A.am[k];
  L.sc_stride = (int)(sizeof(MsfState) / sizeof(double));
  L.done = &st->stop;
  L.done_stride = (int)(sizeof(MsfState) / sizeof(int));
  L.partial = partial;

  MsfState *hs = (MsfState *)host_workspace(2, sizeof(MsfState) * ns_);
  THR_CHECK(hs, THR_ERR_HIP, "host workspace allocation failed");
  long t = 0;  // completed iterations
  int rc = THR_OK;
  int blocks = 0;
  // reliable updates while the Krylov storage is FP32/FP16: each time the
  // iterated residual of the smallest active shift has fallen by rdelta since
  // the last one (worst over the sources), the base
  // residual of every running source is replaced by the FP64 true residual of
  // its smallest active shift divided by that shift's zeta, and the active
  // shift directions are moved with it.  The smallest shift needs the most
  // iterations and is the costliest to correct later; the heavier shifts keep
  // the drift of the low-precision iterations before they converged (the
  // residual gap of a recurrence is the sum of its local rounding errors, so
  // it is set by |r| when the storage precision drops) for the FP64 correction.
  const double rdelta = context_reliable();
  double r_ref = 1.0;  // smallest shift's worst |r_i|/|b| at the last reliable update
  int n_reliable = 0;
  auto reliable_update = [&]() -> int {
    const size_t nrel = (size_t)(3 + MAXSHIFT) * ns_;
    double *hrel = (double *)host_workspace(6, sizeof(double) * nrel);
    int *istar = (int *)host_workspace(7, sizeof(int) * ns_);
    THR_CHECK(hrel && istar, THR_ERR_HIP, "host workspace allocation failed");
    int any = 0;
    for (int k = 0; k < ns_; k++) {
      const double *ak = A.aden_src ? A.aden_src + (long)k * n : A.aden;
      int best = -1;
      if (!hs[k].stop)
        for (int i = 0; i < n; i++)
          if (hs[k].active[i] && (best < 0 || ak[i] < ak[best])) best = i;
      const double z = best >= 0 ? hs[k].zeta_ii[best] : 0.0;
      const bool on = best >= 0 && z != 0.0 && std::isfinite(1.0 / z);
      istar[k] = best;
      hrel[k] = on ? 1.0 : 0.0;
      hrel[ns_ + k] = on ? ak[best] : 0.0;
      hrel[2 * ns_ + k] = on ? 1.0 / z : 0.0;
      for (int j = 0; j < MAXSHIFT; j++)
        hrel[3 * ns_ + k * MAXSHIFT + j] =
            (on && j < n && hs[k].active[j]) ? hs[k].zeta_ii[j] : 0.0;
      any += on ? 1 : 0;
    }
    if (!any) return THR_OK;
    double2 *xg = (double2 *)workspace(380, sizeof(double2) * nsb);
    double2 *t1 = (double2 *)workspace(381, sizeof(double2) * nsb);
    double2 *t2 = (double2 *)workspace(382, sizeof(double2) * nsb);
    double *drel = (double *)workspace(383, sizeof(double) * nrel);
    const int nbr = (int)nblk(v5);
    double *rpart = (double *)workspace(384, sizeof(double) * nbr * ns_);
    THR_CHECK(xg && t1 && t2 && drel && rpart, THR_ERR_HIP, "workspace allocation failed");
    for (int k = 0; k < ns_; k++)
      if (hrel[k] != 0.0)
        THR_HIP(hipMemcpyAsync(xg + k * nsp, A.x + ((long)k * n + istar[k]) * nsp,
                               sizeof(double2) * nsp, hipMemcpyDeviceToDevice, s));
    THR_HIP(hipMemcpyAsync(drel, hrel, sizeof(double) * nrel, hipMemcpyHostToDevice, s));
    DslashLaunch a{};
    a.d = d;
    a.nrhs = ns_;
    a.prec = P64;
    a.imass = A.imass;
    a.am3 = A.am3;
    for (int k = 0; k < ns_; k++) a.am[k] = A.am[k];
    a.links = A.u64;
    a.mode = EPI_STORE;
    a.sc_stride = 1;
    a.done_stride = 1;
    a.in = xg;
    a.out = t1;
    TRY_RC(launch_dslash(a, s));
    DslashLaunch c = a;
    c.dagger = 1;
    c.in = t1;
    c.out = t2;
    TRY_RC(launch_dslash(c, s));
    const long slot = (t - 1) % msb;
    dim3 g(nbr, ns_);
    if (prec == P32)
      hipLaunchKernelGGL(k_msf_reliable<P32>, g, dim3(BLK), 0, s, v5, n, A.b, t2, xg, drel,
                         (float2 *)ring[P32][slot], nullptr, Pv, rpart);
    else if (prec == P16)
      hipLaunchKernelGGL(k_msf_reliable<P16>, g, dim3(BLK), 0, s, v5, n, A.b, t2, xg, drel,
                         (__half2 *)ring[P16][slot], ringn[slot], Pv, rpart);
    else
      hipLaunchKernelGGL(k_msf_reliable<P64>, g, dim3(BLK), 0, s, v5, n, A.b, t2, xg, drel,
                         (double2 *)ring[P64][slot], nullptr, Pv, rpart);
    hipLaunchKernelGGL(k_msf_reliable_fin, dim3(ns_), dim3(BLK), 0, s, st, drel, rpart, nbr);
    THR_HIP(hipGetLastError());
    n_reliable++;
    return THR_OK;
  };
  // one block = msb iterations + flush + reset; every block after the first
  // issues identical launches, so it can be captured once per precision as a
  // hipGraph and replayed (THR_GRAPHS=0 disables)
  const char *ge = std::getenv("THR_GRAPHS");
  const bool use_graph = !(ge && ge[0] == '0') && context_graphs();
  // the Dslash passes derive omega and gamma from each other's partial sums
  // and record alpha and lambda; one finalize per block replays the shift
  // recurrences (2 launches per iteration instead of 4; THR_FUSED=0: a
  // finalize kernel after every pass).  Results are bit-identical.  Every
  // Dslash block then sums the other pass's partials before its sites, which
  // costs more than the two saved launches once a pass is many blocks: the
  // passes fuse only up to FUSE_MAX_BLOCKS blocks (all right-hand sides),
  // decided per precision (THR_FUSED=1 / 0 forces either way).
  constexpr long FUSE_MAX_BLOCKS = 2048;
  const char *fe = std::getenv("THR_FUSED");
  const int fuse_env = fe ? (fe[0] == '0' ? 0 : 1) : -1;
  auto fused_for = [&](int pr) {
    return fuse_env >= 0 ? fuse_env == 1 : (long)nbl[pr] * ns_ <= FUSE_MAX_BLOCKS;
  };
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  auto enqueue_block = [&](long t0, int pr) -> int {
    const bool fused = fused_for(pr);
    for (int j = 0; j < msb; j++) {
      const long tt = t0 + j;
      void *rin = (tt == 0) ? r0[pr] : ring[pr][(tt - 1) % msb];
      float *rinn = (pr == P16) ? ((tt == 0) ? r0n : ringn[(tt - 1) % msb]) : nullptr;
      void *rout = ring[pr][tt % msb];
      float *routn = (pr == P16) ? ringn[tt % msb] : nullptr;
      // h = M p ; alpha = |h|^2.  p = r + gamma p is never formed: h obeys
      // h_new = M r_new + gamma h_old (M is linear), so the pass reads r and
      // the old h and writes h in place
      DslashLaunch a = L;
      a.prec = pr;
      a.links = pr == P64 ? (const void *)A.u64 : (const void *)A.u32;
      a.dagger = 0;
      a.in = rin;
      a.in_norm = rinn;
      a.out = hbuf[pr];
      a.out_norm = pr == P16 ? hn : nullptr;
      if (tt == 0) {
        a.mode = EPI_NORM;  // p_0 = r_0 = b
      } else {
        a.mode = EPI_HREC;
        a.b_ptr = &st->gammag;
        if (fused && j > 0) {
          // gamma = lambda_{k-1} / delta_{k-1} from the last M^+ pass's partials
          a.fsum = lpart;
          a.fsum_n = nbl[pr];
          a.fsum_stride = nb;
          a.fdelta = j >= 2 ? &st->lam_rec[j - 2] : &st->delta;
          a.frec = &st->lam_rec[j - 1];
        }
      }
      if (fused) a.partial = partial;  // alpha partials; lambda ones go to lpart
      TRY_RC(launch_dslash_cfg(a, cfg[pr], s));
      if (!fused)
        hipLaunchKernelGGL(k_msf_fin1, dim3(ns_), dim3(BLK), 0, s, st, aden, partial, nb,
                           nbl[pr]);
      // r_new = r + omega M^+ h ; lambda = |r_new|^2
      DslashLaunch b = L;
      b.prec = pr;
      b.links = a.links;
      b.dagger = 1;
      b.in = hbuf[pr];
      b.in_norm = a.out_norm;
      b.mode = EPI_CG;
      b.r = rin;
      b.r_norm = rinn;
      b.r_out = rout;
      b.r_out_norm = routn;
      b.a_ptr = &st->minus_omega;
      b.b_ptr = nullptr;
      if (fused) {
        // omega = -delta_k / alpha_k from the M pass's partials
        b.fsum = partial;
        b.fsum_n = nbl[pr];
        b.fsum_stride = nb;
        b.fdelta = j >= 1 ? &st->lam_rec[j - 1] : &st->delta;
        b.frec = &st->al_rec[j];
        b.partial = lpart;
      }
      TRY_RC(launch_dslash_cfg(b, cfg[pr], s));
      if (!fused)
        hipLaunchKernelGGL(k_msf_fin2, dim3(ns_), dim3(BLK), 0, s, st, partial, nb, nbl[pr]);
      THR_HIP(hipGetLastError());
    }
    if (fused)
      hipLaunchKernelGGL(k_msf_fin_block, dim3(ns_), dim3(BLK), 0, s, st, aden, lpart, nb, nbl[pr],
                         msb);
    // flush the block into x_i, p_i
    dim3 gf(nblk(nsp), ns_);
    void *const *rp = (void *const *)ringptr + pr * MSB;
    float *const *rnp = (float *const *)((void **)ringptr + 3 * MSB);
    if (pr == P64) {
      // FP64 residuals: 32 slots per launch
      for (int q0 = 0; q0 < msb; q0 += 32)
        hipLaunchKernelGGL((k_msf_flush<P64, 32>), gf, dim3(BLK), 0, s, v5, n,
                           (double2 *const *)rp, nullptr, A.b, A.x, Pv, st, 0, q0);
    } else if (pr == P32) {
      if (msb == 64)
        hipLaunchKernelGGL((k_msf_flush<P32, 64>), gf, dim3(BLK), 0, s, v5, n,
                           (float2 *const *)rp, nullptr, A.b, A.x, Pv, st, 0, 0);
      else
        hipLaunchKernelGGL((k_msf_flush<P32, 32>), gf, dim3(BLK), 0, s, v5, n,
                           (float2 *const *)rp, nullptr, A.b, A.x, Pv, st, 0, 0);
    } else {
      if (msb == 64)
        hipLaunchKernelGGL((k_msf_flush<P16, 64>), gf, dim3(BLK), 0, s, v5, n,
                           (__half2 *const *)rp, rnp, A.b, A.x, Pv, st, 0, 0);
      else
        hipLaunchKernelGGL((k_msf_flush<P16, 32>), gf, dim3(BLK), 0, s, v5, n,
                           (__half2 *const *)rp, rnp, A.b, A.x, Pv, st, 0, 0);
    }
    hipLaunchKernelGGL(k_msf_reset, dim3(ns_), dim3(MAXSHIFT), 0, s, st);
    THR_HIP(hipGetLastError());
    return THR_OK;
  };
  auto free_graphs = [&]() {
    for (int k = 0; k < 3; k++)
      if (gexec[k]) (void)hipGraphExecDestroy(gexec[k]);
  };
  // graph capture is an optimisation: if another thread's work invalidates a
  // capture (or instantiation fails), the block runs eagerly instead and the
  // solve continues without graphs (nothing of a failed capture executed)
  bool graphs_ok = use_graph;
  for (;;) {
    bool done_block = false;
    if (graphs_ok && t > 0) {
      if (!gexec[prec]) {
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed) == hipSuccess) {
          rc = enqueue_block(t, prec);
          const hipError_t e = hipStreamEndCapture(s, &g);
          if (rc && rc != THR_ERR_HIP) {  // argument error, not a capture problem
            if (g) (void)hipGraphDestroy(g);
            free_graphs();
            return rc;
          }
          if (rc || e != hipSuccess ||
              hipGraphInstantiate(&gexec[prec], g, nullptr, nullptr, 0) != hipSuccess)
            gexec[prec] = nullptr;
          if (g) (void)hipGraphDestroy(g);
        }
        if (!gexec[prec]) {
          (void)hipGetLastError();  // clear the failed capture's error state
          graphs_ok = false;
          rc = THR_OK;
        }
      }
      if (gexec[prec]) {
        if (hipGraphLaunch(gexec[prec], s) != hipSuccess) {
          free_graphs();
          set_error("hipGraphLaunch failed");
          return THR_ERR_HIP;
        }
        done_block = true;
      }
    }
    if (!done_block) {
      rc = enqueue_block(t, prec);
      if (rc) {
        free_graphs();
        return rc;
      }
    }
    t += msb;
    blocks++;
    if (hipMemcpyAsync(hs, st, sizeof(MsfState) * ns_, hipMemcpyDeviceToHost, s) !=
            hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      free_graphs();
      set_error("multishift: state read-back failed");
      return THR_ERR_HIP;
    }
    bool all = true;
    double worst = 0.0;  // base |r|/|b|
    double worst_lo = 0.0;  // |r_i|/|b| of the smallest active shift
    for (int k = 0; k < ns_; k++) {
      all = all && hs[k].stop;
      if (hs[k].stop) continue;
      const double rb = std::sqrt(hs[k].lambda / hs[k].bnorm2);
      worst = std::max(worst, rb);
      const double *ak = A.aden_src ? A.aden_src + (long)k * n : A.aden;
      int lo = -1;
      for (int i = 0; i < n; i++)
        if (hs[k].active[i] && (lo < 0 || ak[i] < ak[lo])) lo = i;
      if (lo >= 0) worst_lo = std::max(worst_lo, std::fabs(hs[k].zeta_ii[lo]) * rb);
    }
    if (all) break;
    if (t > (long)A.maxcg + 2 * msb) {
      free_graphs();
      set_error("multishift: stop flag not raised");
      return THR_ERR_STATE;
    }
    // the trigger follows the smallest shift's own residual: the base
    // residual of a light-mass operator may hardly fall while the shifted
    // ones converge through zeta, and the drift of shift i accumulates with
    // its own residual history
    if (rdelta > 0 && prec != P64 && worst_lo < rdelta * r_ref) {
      rc = reliable_update();
      if (rc) {
        free_graphs();
        return rc;
      }
      r_ref = worst_lo;
    }
    if (prec == P64) r_ref = worst_lo;  // drift starts when the storage precision drops
    // lower the Krylov precision as the residual falls
    int want = prec;
    if (A.th32 > 0 && worst < A.th32 && prec == P64) want = P32;
    if (A.th16 > 0 && worst < A.th16 && prec != P16) want = P16;
    if (want != prec) {
      void *rl = ring[prec][(t - 1) % msb];
      float *rln = prec == P16 ? ringn[(t - 1) % msb] : nullptr;
      int rcc = conv(d, ns_, prec, rl, rln, want, ring[want][(t - 1) % msb],
                     want == P16 ? ringn[(t - 1) % msb] : nullptr, s);
      if (!rcc)
        rcc = conv(d, ns_, prec, hbuf[prec], prec == P16 ? hn : nullptr, want, hbuf[want],
                   want == P16 ? hn : nullptr, s);
      if (rcc) {
        free_graphs();
        return rcc;
      }
      prec = want;
      if (res) res->switches[prec] = (int)t;
    }
  }
  free_graphs();
  if (A.note_reliable) context_note_reliable(n_reliable);
  if (res) {
    res->reliable = n_reliable;
    res->iterations = (int)t;
    res->blocks = blocks;
    res->nonfinite = 0;
    res->unconverged = 0;
    for (int k = 0; k < ns_; k++) {
      res->cg_return[k] = hs[k].cg_return;
      for (int i = 0; i < n; i++) res->cg_returns[k * MAXSHIFT + i] = hs[k].cg_returns[i];
      res->nonfinite += hs[k].nonfinite ? 1 : 0;
      // a shift leaves the active set only once it meets its target
      for (int i = 0; i < n; i++) res->unconverged += hs[k].active[i] ? 1 : 0;
    }
  }
  return rc;
}

int msfast_true_residuals(const MsfArgs &A, double *relres, hipStream_t s) {
  const Dims &d = A.d;
  const int n = A.nshift, ns_ = A.nsrc;
  const long v5 = d.v5(), nsp = 4 * v5;
  const int nb = (int)dslash_blocks(d);
  double2 *t1 = (double2 *)workspace(370, sizeof(double2) * nsp * ns_);
  double2 *t2 = (double2 *)workspace(371, sizeof(double2) * nsp * ns_);
  double *partial = (double *)workspace(366, sizeof(double) * nb * ns_);
  double *sums = (double *)workspace(368, sizeof(double) * ns_);
  double2 *xs = (double2 *)workspace(372, sizeof(double2) * nsp * ns_);
  THR_CHECK(t1 && t2 && partial && sums && xs, THR_ERR_HIP, "workspace allocation failed");
  double *bn = (double *)host_workspace(3, sizeof(double) * ns_);
  double *hsum = (double *)host_workspace(4, sizeof(double) * ns_);
  THR_CHECK(bn && hsum, THR_ERR_HIP, "host workspace allocation failed");
  hipLaunchKernelGGL(k_norm_src, dim3(nb, ns_), dim3(BLK), 0, s, v5, A.b, partial);
  hipLaunchKernelGGL(k_sum_rows, dim3(ns_), dim3(BLK), 0, s, partial, nb, sums);
  THR_HIP(hipMemcpyAsync(bn, sums, sizeof(double) * ns_, hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  DslashLaunch L{};
  L.d = d;
  L.nrhs = ns_;
  L.prec = P64;
  L.imass = A.imass;
  L.am3 = A.am3;
  for (int k = 0; k < ns_; k++) L.am[k] = A.am[k];
  L.links = A.u64;
  L.mode = EPI_STORE;
  L.sc_stride = 1;
  L.done_stride = 1;
  for (int i = 0; i < n; i++) {
    // gather shift i of every source into one batch
    for (int k = 0; k < ns_; k++)
      THR_HIP(hipMemcpyAsync(xs + k * nsp, A.x + ((long)k * n + i) * nsp, sizeof(double2) * nsp,
                             hipMemcpyDeviceToDevice, s));
    DslashLaunch a = L;
    a.in = xs;
    a.out = t1;
    TRY_RC(launch_dslash(a, s));
    DslashLaunch b = L;
    b.dagger = 1;
    b.in = t1;
    b.out = t2;
    TRY_RC(launch_dslash(b, s));
    hipLaunchKernelGGL(k_true_res, dim3(nb, ns_), dim3(BLK), 0, s, v5, A.b, t2, xs, A.aden[i],
                       nsp, partial);
    hipLaunchKernelGGL(k_sum_rows, dim3(ns_), dim3(BLK), 0, s, partial, nb, sums);
    THR_HIP(hipMemcpyAsync(hsum, sums, sizeof(double) * ns_, hipMemcpyDeviceToHost, s));
    THR_HIP(hipStreamSynchronize(s));
    for (int k = 0; k < ns_; k++) relres[k * n + i] = std::sqrt(hsum[k] / bn[k]);
  }
  return THR_OK;
}

int msfast_correct(const MsfArgs &A, int max_rounds, double tol_true, double *relres,
                   int *corr_iters, hipStream_t s) {
  const Dims &d = A.d;
  const int n = A.nshift, ns_ = A.nsrc;
  const long v5 = d.v5(), nsp = 4 * v5;
  std::vector<double> rr((size_t)ns_ * n);
  int total = 0;
  for (int round = 0;; round++) {
    TRY_RC(msfast_true_residuals(A, rr.data(), s));
    // failing pairs q = k * n + i, grouped by shift so that each batched
    // correction holds pairs of similar conditioning (similar iteration counts)
    std::vector<int> fail;
    for (int i = 0; i < n; i++)
      for (int k = 0; k < ns_; k++)
        if (!(rr[k * n + i] <= tol_true)) fail.push_back(k * n + i);
    if (fail.empty() || round >= max_rounds) break;
    // buffers for one group of up to ns_ (source, shift) pairs
    double2 *xs = (double2 *)workspace(372, sizeof(double2) * nsp * ns_);
    double2 *t1 = (double2 *)workspace(370, sizeof(double2) * nsp * ns_);
    double2 *t2 = (double2 *)workspace(371, sizeof(double2) * nsp * ns_);
    double2 *bs = (double2 *)workspace(374, sizeof(double2) * nsp * ns_);
    double2 *rhs = (double2 *)workspace(375, sizeof(double2) * nsp * ns_);
    double2 *xc = (double2 *)workspace(376, sizeof(double2) * nsp * ns_);
    double *sig = (double *)workspace(377, sizeof(double) * ns_);
    THR_CHECK(xs && t1 && t2 && bs && rhs && xc && sig, THR_ERR_HIP, "workspace allocation failed");
    for (size_t g0 = 0; g0 < fail.size(); g0 += ns_) {
      const int g = (int)std::min<size_t>(ns_, fail.size() - g0);
      std::vector<double> sg(g), tl(g);
      std::vector<float> am(g);
      for (int j = 0; j < g; j++) {
        const int q = fail[g0 + j], k = q / n, i = q % n;
        sg[j] = A.aden[i];
        am[j] = A.am[k];
        // solve (M^+M + sigma_i) c = r_ki to |r_ki - (M^+M + sigma_i) c| <= tol |b_k| / 2
        tl[j] = 0.5 * tol_true / rr[q];
        THR_HIP(hipMemcpyAsync(xs + j * nsp, A.x + (long)q * nsp, sizeof(double2) * nsp,
                               hipMemcpyDeviceToDevice, s));
        THR_HIP(hipMemcpyAsync(bs + j * nsp, A.b + (long)k * nsp, sizeof(double2) * nsp,
                               hipMemcpyDeviceToDevice, s));
      }
      double *sgp = (double *)host_workspace(5, sizeof(double) * g);
      THR_CHECK(sgp, THR_ERR_HIP, "host workspace allocation failed");
      std::copy(sg.begin(), sg.end(), sgp);
      THR_HIP(hipMemcpyAsync(sig, sgp, sizeof(double) * g, hipMemcpyHostToDevice, s));
      DslashLaunch L{};
      L.d = d;
      L.nrhs = g;
      L.prec = P64;
      L.imass = A.imass;
      L.am3 = A.am3;
      for (int j = 0; j < g; j++) L.am[j] = am[j];
      L.links = A.u64;
      L.mode = EPI_STORE;
      L.sc_stride = 1;
      L.done_stride = 1;
      DslashLaunch a = L;
      a.in = xs;
      a.out = t1;
      TRY_RC(launch_dslash(a, s));
      DslashLaunch b = L;
      b.dagger = 1;
      b.in = t1;
      b.out = t2;
      TRY_RC(launch_dslash(b, s));
      hipLaunchKernelGGL(k_resvec, dim3(nblk(nsp), g), dim3(BLK), 0, s, nsp, bs, t2, xs, sig, rhs);
      THR_HIP(hipGetLastError());
      MsfArgs B = A;
      B.nsrc = g;
      B.nshift = 1;
      B.aden = sg.data();
      B.aden_src = sg.data();
      B.tol_src = tl.data();
      B.am = am.data();
      B.b = rhs;
      B.x = xc;
      B.th32 = 1.0;
      B.th16 = 0.0;
      B.start_prec = P32;
      B.note_reliable = false;
      MsfResult R{};
      TRY_RC(msfast_solve(B, &R, s));
      THR_CHECK(!R.nonfinite, THR_ERR_NONFINITE,
                "multishift correction: non-finite residual recurrence");
      total += R.iterations;
      for (int j = 0; j < g; j++) {
        hipLaunchKernelGGL(k_addvec, dim3(nblk(nsp)), dim3(BLK), 0, s, nsp,
                           A.x + (long)fail[g0 + j] * nsp, xc + j * nsp);
      }
      THR_HIP(hipGetLastError());
    }
  }
  if (relres)
    for (int q = 0; q < ns_ * n; q++) relres[q] = rr[q];
  if (corr_iters) *corr_iters = total;
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

}  // namespace thr
