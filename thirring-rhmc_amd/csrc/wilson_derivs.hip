// Wilson-kernel domain-wall operator (secondary generation kernel) and the
// Shamir fermion force.
//
// Wilson kernel, dirac.F90:226-299 (M) and 538-625 (M^dagger).  The reference
// builds it from 4-D Wilson applications on s-slices; with D_W that slice
// operator and the s-couplings collected into one field it reads
//   M  R = D_W (R + Q) + (R - Q),
//        Q(s) = P+ R(s-1) + P- R(s+1)       (interior s)
//             + c1 P+ R(Ls) at s=1, c2 P- R(1) at s=Ls
//   M^+R = T + R + P- cA (T - R)(s_A) + P+ cB (T - R)(s_B),  T = D_W^+ R
// so each application is one 4-D stencil pass plus two local passes.
//
// Force, derivs_module.F90:99-196: for every link (x, mu) the sum over s and
// spin of the Wilson and Dirac parts of d/dtheta Re<R, M^(+) X2>, with the
// float32 coefficient tzi_real = 2*real(anum).
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "kernels.h"

namespace thr {

namespace {

// Y = R + Q  (M) ; coefficients c1, c2 multiply the wall terms
__global__ __launch_bounds__(BLK) void k_wk_pre(const Dims d, int nrhs, const double2 *R,
                                                double2 *Y, double2 c1, double2 c2) {
  const long v5 = d.v5();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v5 * nrhs) return;
  const long rhs = i / v5, site = i % v5;
  const int s = (int)(site % d.ls);
  const long col = site - s;
  const double2 *r = R + rhs * 4 * v5;
  double2 *y = Y + rhs * 4 * v5;
  for (int a = 0; a < 4; a++) {
    cx<double> v = ldc(r, a * v5 + site);
    if (a < 2) {  // P+ from s-1
      if (s > 0) v = v + ldc(r, a * v5 + site - 1);
      if (s == 0) v = v + cmul(cx<double>{c1.x, c1.y}, ldc(r, a * v5 + col + d.ls - 1));
    } else {  // P- from s+1
      if (s + 1 < d.ls) v = v + ldc(r, a * v5 + site + 1);
      if (s == d.ls - 1) v = v + cmul(cx<double>{c2.x, c2.y}, ldc(r, a * v5 + col));
    }
    st(y, a * v5 + site, v);
  }
}

// Phi += 2R - Y
__global__ __launch_bounds__(BLK) void k_wk_post(long n, const double2 *R, const double2 *Y,
                                                 double2 *Phi) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const double2 r = R[i], y = Y[i], p = Phi[i];
  Phi[i] = make_double2(p.x + (2.0 * r.x - y.x), p.y + (2.0 * r.y - y.y));
}

// Phi = T + R + P- cA (T - R)(sA) + P+ cB (T - R)(sB)   (M^dagger)
__global__ __launch_bounds__(BLK) void k_wkd_post(const Dims d, int nrhs, const double2 *R,
                                                  const double2 *T, double2 *Phi, double2 z1,
                                                  double2 z2) {
  const long v5 = d.v5();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v5 * nrhs) return;
  const long rhs = i / v5, site = i % v5;
  const int s = (int)(site % d.ls);
  const long col = site - s;
  const double2 *r = R + rhs * 4 * v5;
  const double2 *tt = T + rhs * 4 * v5;
  double2 *ph = Phi + rhs * 4 * v5;
  for (int a = 0; a < 4; a++) {
    cx<double> v = ldc(tt, a * v5 + site) + ldc(r, a * v5 + site);
    if (a >= 2) {  // P-: from s-1, wall s=1 <- s=Ls with z1
      if (s > 0) v = v + (ldc(tt, a * v5 + site - 1) - ldc(r, a * v5 + site - 1));
      if (s == 0) {
        const long q = a * v5 + col + d.ls - 1;
        v = v + cmul(cx<double>{z1.x, z1.y}, ldc(tt, q) - ldc(r, q));
      }
    } else {  // P+: from s+1, wall s=Ls <- s=1 with z2
      if (s + 1 < d.ls) v = v + (ldc(tt, a * v5 + site + 1) - ldc(r, a * v5 + site + 1));
      if (s == d.ls - 1) {
        const long q = a * v5 + col;
        v = v + cmul(cx<double>{z2.x, z2.y}, ldc(tt, q) - ldc(r, q));
      }
    }
    st(ph, a * v5 + site, v);
  }
}

// gamval(mu,a)/akappa and gamin(mu,a)-1 for mu = 1..3 (gammamatrices.F90:27-75)
__constant__ double kGre[3][4] = {{0, 0, 0, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}};
__constant__ double kGim[3][4] = {{-1, -1, 1, 1}, {0, 0, 0, 0}, {-1, 1, 1, -1}};
__constant__ int kGin[3][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}};

// Shamir force, streaming form.  A block holds nl whole s-lines (columns
// (x, y), Ls consecutive threads each, s fastest: every load of a spin plane
// is one contiguous run) and marches them through a chunk of t.  At each step
// a thread forms the three link terms of its 5-D site, whose neighbours are
// the site itself and x+mu: the t neighbour loaded at step t is the centre of
// step t+1, so per step a thread loads R and X2 at (x+1), (y+1) and (t+1) only.
// The Ls terms of each link are summed in LDS in increasing s (a fixed order,
// so results are identical run to run) and added into dSdpi.
//   link term (x, mu) = w1 sum_a Im(R0_a^* X1_a - R1_a^* X0_a)
//                     + w2 sum_a Im(g_mu,a (R0_a^* X1_b + R1_a^* X0_b)),  b = gamin(mu, a)
// with (R0, X0) at x and (R1, X1) at x+mu, w1 = fl32(tzi/2), w2 = +-tzi
// (derivs_module.F90:131-187).
template <int MU>
__device__ inline double derivs_term(const cx<double> R0[4], const cx<double> X0[4],
                                     const cx<double> R1[4], const cx<double> X1[4], double w1,
                                     double w2) {
  // gamval(mu, a) / akappa (gammamatrices.F90:27-75) and gamin(mu, a) - 1
  constexpr double gre[3][4] = {{0, 0, 0, 0}, {-1, 1, 1, -1}, {0, 0, 0, 0}};
  constexpr double gim[3][4] = {{-1, -1, 1, 1}, {0, 0, 0, 0}, {-1, 1, 1, -1}};
  constexpr int gin[3][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}};
  double sw = 0.0, sd = 0.0;
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const int b = gin[MU][a];
    sw += cjmul(R0[a], X1[a]).i - cjmul(R1[a], X0[a]).i;
    const cx<double> g{0.5 * gre[MU][a], 0.5 * gim[MU][a]};
    sd += cmul(g, cjmul(R0[a], X1[b]) + cjmul(R1[a], X0[b])).i;
  }
  return w1 * sw + w2 * sd;
}

__device__ inline void ld_spinor(const double2 *__restrict__ p, unsigned v5, unsigned site,
                                 cx<double> v[4]) {
#pragma unroll
  for (int a = 0; a < 4; a++) v[a] = ldc(p, (long)a * v5 + site);
}

// LDS x exchange of the batched derivs: with 64-lane segments (32 < Ls <=
// 64) and Lx % 4 == 0 a block's 4 waves are 4 consecutive x of one (y, t),
// so the x+1 neighbour of waves 0-2 is the centre of the next wave: each wave
// stores its centre (R and X2, 8 complex numbers per lane, 32 KB per block)
// and reads the next wave's instead of loading it; wave 3 loads its own.
// 32^3 x 64, 25 poles: 1.74 against 1.82 ms; 16^3 x 48 even.  The same in
// the single derivs march (two barriers per t step, 156 against 128 VGPRs:
// 3 waves per SIMD instead of 4) was slower (84 against 78 us at 32^3 x 64).
struct DXBuf {
  double2 (*b)[8][64];
  __device__ inline void put(unsigned w, unsigned ln, const cx<double> R[4],
                             const cx<double> X[4]) const {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      b[w][a][ln] = make_double2(R[a].r, R[a].i);
      b[w][4 + a][ln] = make_double2(X[a].r, X[a].i);
    }
  }
  __device__ inline void get(unsigned w, unsigned ln, cx<double> R[4], cx<double> X[4]) const {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const double2 r = b[w][a][ln], x = b[w][4 + a][ln];
      R[a] = {r.x, r.y};
      X[a] = {x.x, x.y};
    }
  }
};

// Ls <= 64: each s-line takes a segment of W lanes of one wave (W the
// smallest power of two >= Ls, lanes s >= Ls idle), so the sum over s is a
// butterfly of wave shuffles inside the segment (a fixed order, no LDS and no
// block barrier); segment lane 0 adds the three link terms into dSdpi.  The
// t march and loads are those of k_derivs_march below.
__global__ __launch_bounds__(BLK, 3) void k_derivs_seg(const Dims d, const double2 *__restrict__ R,
                                                    const double2 *__restrict__ X2, float tzi,
                                                    int iflag, unsigned wseg, unsigned tchunk,
                                                    double *ds) {
  const unsigned lx = d.lx, ly = d.ly, lt = d.lt, ls = d.ls;
  const unsigned v2 = lx * ly, v3 = v2 * ls, v4 = v2 * lt, v5 = v4 * ls;
  const unsigned nl = BLK / wseg;           // lines per block
  const unsigned nb2 = (v2 + nl - 1) / nl;  // blocks per t-chunk
  const unsigned chunk = blockIdx.x / nb2, lb = blockIdx.x - chunk * nb2;
  const unsigned line = threadIdx.x / wseg, s = threadIdx.x - line * wseg;
  const unsigned col = lb * nl + line;  // 2-D site x + lx y
  const unsigned t0 = chunk * tchunk, t1 = min(t0 + tchunk, lt);
  const bool valid = s < ls && col < v2;
  const double w1 = (double)(tzi * 0.5f);  // tzi_real*real(akappa)
  const double w2 = (double)tzi * (iflag == 0 ? 1.0 : -1.0);
  unsigned x = 0, y = 0;
  if (col < v2) {
    x = col % lx;
    y = col / lx;
  }
  const unsigned xf = (x + 1 == lx) ? 0 : x + 1, yf = (y + 1 == ly) ? 0 : y + 1;
  const unsigned j = s + ls * col, jx = s + ls * (xf + lx * y), jy = s + ls * (x + lx * yf);
  cx<double> R0[4], X0[4];
  if (valid) {
    ld_spinor(R, v5, j + v3 * t0, R0);
    ld_spinor(X2, v5, j + v3 * t0, X0);
  }
  for (unsigned t = t0; t < t1; t++) {
    double v[3] = {0.0, 0.0, 0.0};
    if (valid) {
      // all six neighbour spinors of the step in flight together
      const unsigned tf = (t + 1 == lt) ? 0 : t + 1;
      cx<double> Rx[4], Xx[4], Ry[4], Xy[4], Rt[4], Xt[4];
      ld_spinor(R, v5, jx + v3 * t, Rx);
      ld_spinor(X2, v5, jx + v3 * t, Xx);
      ld_spinor(R, v5, jy + v3 * t, Ry);
      ld_spinor(X2, v5, jy + v3 * t, Xy);
      ld_spinor(R, v5, j + v3 * tf, Rt);
      ld_spinor(X2, v5, j + v3 * tf, Xt);
      v[0] = derivs_term<0>(R0, X0, Rx, Xx, w1, w2);
      v[1] = derivs_term<1>(R0, X0, Ry, Xy, w1, w2);
      v[2] = derivs_term<2>(R0, X0, Rt, Xt, w1, w2);
#pragma unroll
      for (int a = 0; a < 4; a++) {
        R0[a] = Rt[a];
        X0[a] = Xt[a];
      }
    }
    // every lane of the wave takes part in the shuffles (idle lanes add 0)
    for (unsigned off = wseg >> 1; off > 0; off >>= 1) {
#pragma unroll
      for (int mu = 0; mu < 3; mu++) v[mu] += __shfl_xor(v[mu], (int)off, 64);
    }
    if (s == 0 && col < v2) {
#pragma unroll
      for (int mu = 0; mu < 3; mu++) ds[(long)mu * v4 + col + v2 * t] += v[mu];
    }
  }
}

// k_derivs_seg over several (R, X2) pairs in one launch, one t step per
// thread: blockIdx.y is the pair, so the GPU sweeps the pairs in order with
// every neighbour line shared in L2 by the blocks of the same pair (a thread
// looping over the pairs instead let neighbouring waves drift onto different
// pairs and ran at 2.3 TB/s).  Each block writes its links' terms of its pair
// to part[pair][mu][t][y][x]; k_derivs_sum_pairs adds the pairs in order into
// dSdpi.  The force sums derivs over the poles of a rational approximation
// (qmrherm_module.F90:143-181); one launch over all of them reads each
// pair's fields once like separate calls but fills the GPU on lattices whose
// single call cannot (25 MB per pair at 16^3 x 48).
template <bool XL>
__global__ __launch_bounds__(BLK, 3) void k_derivs_seg_batch(const Dims d,
                                                          const DerivPair *__restrict__ pairs,
                                                          unsigned wseg, double *part) {
  const unsigned lx = d.lx, ly = d.ly, lt = d.lt, ls = d.ls;
  const unsigned v2 = lx * ly, v3 = v2 * ls, v4 = v2 * lt, v5 = v4 * ls;
  const unsigned nl = BLK / wseg;           // lines per block
  const unsigned nb2 = (v2 + nl - 1) / nl;  // blocks per t slice
  const unsigned t = blockIdx.x / nb2, lb = blockIdx.x - t * nb2;
  const unsigned line = threadIdx.x / wseg, s = threadIdx.x - line * wseg;
  const unsigned col = lb * nl + line;  // 2-D site x + lx y
  const bool valid = s < ls && col < v2 && t < lt;
  const DerivPair P = pairs[blockIdx.y];
  unsigned x = 0, y = 0;
  if (col < v2) {
    x = col % lx;
    y = col / lx;
  }
  const unsigned xf = (x + 1 == lx) ? 0 : x + 1, yf = (y + 1 == ly) ? 0 : y + 1;
  const unsigned tf = (t + 1 == lt) ? 0 : t + 1;
  const unsigned j = s + ls * col + v3 * t, jx = s + ls * (xf + lx * y) + v3 * t;
  const unsigned jy = s + ls * (x + lx * yf) + v3 * t, jt = s + ls * col + v3 * tf;
  double v[3] = {0.0, 0.0, 0.0};
  const unsigned w = threadIdx.x >> 6, ln = threadIdx.x & 63;
  cx<double> R0[4], X0[4], Rx[4], Xx[4], Ry[4], Xy[4], Rt[4], Xt[4];
  if (valid) {
    ld_spinor(P.r, v5, j, R0);
    ld_spinor(P.x2, v5, j, X0);
    if (!XL || w == 3) {
      ld_spinor(P.r, v5, jx, Rx);
      ld_spinor(P.x2, v5, jx, Xx);
    }
    ld_spinor(P.r, v5, jy, Ry);
    ld_spinor(P.x2, v5, jy, Xy);
    ld_spinor(P.r, v5, jt, Rt);
    ld_spinor(P.x2, v5, jt, Xt);
  }
  if constexpr (XL) {
    // x+1 of waves 0-2 from the next wave's centre (wseg = 64 and full
    // blocks: every wave takes the barrier)
    __shared__ double2 xraw[4][8][64];
    const DXBuf xb{xraw};
    if (valid) xb.put(w, ln, R0, X0);
    __syncthreads();
    if (valid && w < 3) xb.get(w + 1, ln, Rx, Xx);
  }
  if (valid) {
    v[0] = derivs_term<0>(R0, X0, Rx, Xx, P.w1, P.w2);
    v[1] = derivs_term<1>(R0, X0, Ry, Xy, P.w1, P.w2);
    v[2] = derivs_term<2>(R0, X0, Rt, Xt, P.w1, P.w2);
  }
  // every lane of the wave takes part in the shuffles (idle lanes add 0)
  for (unsigned off = wseg >> 1; off > 0; off >>= 1) {
#pragma unroll
    for (int mu = 0; mu < 3; mu++) v[mu] += __shfl_xor(v[mu], (int)off, 64);
  }
  if (s == 0 && col < v2 && t < lt) {
    double *o = part + (long)blockIdx.y * 3 * v4;
#pragma unroll
    for (int mu = 0; mu < 3; mu++) o[(long)mu * v4 + col + v2 * t] = v[mu];
  }
}

// dSdpi[e] += sum over the pairs, in pair order, of part[pair][e]
__global__ __launch_bounds__(BLK) void k_derivs_sum_pairs(long n, int npair,
                                                          const double *__restrict__ part,
                                                          double *ds) {
  const long e = (long)blockIdx.x * BLK + threadIdx.x;
  if (e >= n) return;
  double v = 0.0;
  for (int k = 0; k < npair; k++) v += part[(long)k * n + e];
  ds[e] += v;
}

__global__ __launch_bounds__(BLK) void k_derivs_march(const Dims d, const double2 *__restrict__ R,
                                                      const double2 *__restrict__ X2, float tzi,
                                                      int iflag, unsigned nl, unsigned tchunk,
                                                      double *ds) {
  __shared__ double terms[2][3][BLK];
  const unsigned lx = d.lx, ly = d.ly, lt = d.lt, ls = d.ls;
  const unsigned v2 = lx * ly, v3 = v2 * ls, v4 = v2 * lt, v5 = v4 * ls;
  const unsigned nb2 = (v2 + nl - 1) / nl;  // blocks per t-chunk
  const unsigned chunk = blockIdx.x / nb2, lb = blockIdx.x - chunk * nb2;
  const unsigned line = threadIdx.x / ls, s = threadIdx.x - line * ls;
  const unsigned col = lb * nl + line;  // 2-D site x + lx y
  const unsigned t0 = chunk * tchunk, t1 = min(t0 + tchunk, lt);
  const bool valid = line < nl && col < v2;
  const double w1 = (double)(tzi * 0.5f);  // tzi_real*real(akappa)
  const double w2 = (double)tzi * (iflag == 0 ? 1.0 : -1.0);
  unsigned x = 0, y = 0;
  if (valid) {
    x = col % lx;
    y = col / lx;
  }
  const unsigned xf = (x + 1 == lx) ? 0 : x + 1, yf = (y + 1 == ly) ? 0 : y + 1;
  const unsigned j = s + ls * col, jx = s + ls * (xf + lx * y), jy = s + ls * (x + lx * yf);
  cx<double> R0[4], X0[4];
  if (valid) {
    ld_spinor(R, v5, j + v3 * t0, R0);
    ld_spinor(X2, v5, j + v3 * t0, X0);
  }
  for (unsigned t = t0; t < t1; t++) {
    const unsigned pb = (t - t0) & 1;
    if (valid) {
      const unsigned tf = (t + 1 == lt) ? 0 : t + 1;
      cx<double> R1[4], X1[4], Rt[4], Xt[4];
      ld_spinor(R, v5, jx + v3 * t, R1);
      ld_spinor(X2, v5, jx + v3 * t, X1);
      terms[pb][0][threadIdx.x] = derivs_term<0>(R0, X0, R1, X1, w1, w2);
      ld_spinor(R, v5, jy + v3 * t, R1);
      ld_spinor(X2, v5, jy + v3 * t, X1);
      terms[pb][1][threadIdx.x] = derivs_term<1>(R0, X0, R1, X1, w1, w2);
      ld_spinor(R, v5, j + v3 * tf, Rt);
      ld_spinor(X2, v5, j + v3 * tf, Xt);
      terms[pb][2][threadIdx.x] = derivs_term<2>(R0, X0, Rt, Xt, w1, w2);
#pragma unroll
      for (int a = 0; a < 4; a++) {
        R0[a] = Rt[a];
        X0[a] = Xt[a];
      }
    }
    // one barrier per step: buffer pb is written again two steps later, after
    // the next step's barrier, by when its readers below are done
    __syncthreads();
    for (unsigned r = threadIdx.x; r < 3 * nl; r += blockDim.x) {  // 3 nl > blockDim when Ls < 3
      const unsigned mu = r / nl, ln = r - mu * nl, c = lb * nl + ln;
      if (c < v2) {
        double sum = 0.0;
        for (unsigned k = 0; k < ls; k++) sum += terms[pb][mu][ln * ls + k];
        ds[(long)mu * v4 + c + v2 * t] += sum;
      }
    }
  }
}

// any Ls: one thread per link, serial over s and spin (used when an s-line
// does not fit one block)
__global__ __launch_bounds__(BLK) void k_derivs(const Dims d, const double2 *__restrict__ R,
                                                const double2 *__restrict__ X2, float tzi,
                                                int iflag, double *ds) {
  const long v4 = d.v4(), v5 = d.v5();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v4 * 3) return;
  const int mu = (int)(i / v4);
  const long s4 = i % v4;
  const int x = (int)(s4 % d.lx), y = (int)((s4 / d.lx) % d.ly), t = (int)(s4 / ((long)d.lx * d.ly));
  int xf = x, yf = y, tf = t;
  if (mu == 0) xf = (x + 1 == d.lx) ? 0 : x + 1;
  if (mu == 1) yf = (y + 1 == d.ly) ? 0 : y + 1;
  if (mu == 2) tf = (t + 1 == d.lt) ? 0 : t + 1;
  const long f4 = xf + (long)d.lx * (yf + (long)d.ly * tf);
  const long c0 = s4 * d.ls, c1 = f4 * d.ls;
  const double w1 = (double)(tzi * 0.5f);  // tzi_real*real(akappa)
  const double w2 = (double)tzi * (iflag == 0 ? 1.0 : -1.0);
  double acc = 0.0;
  for (int a = 0; a < 4; a++) {
    const int b = kGin[mu][a];
    const cx<double> g{0.5 * kGre[mu][a], 0.5 * kGim[mu][a]};
    double sw = 0.0, sd = 0.0;
    for (int s = 0; s < d.ls; s++) {
      const cx<double> r0 = ldc(R, a * v5 + c0 + s), r1 = ldc(R, a * v5 + c1 + s);
      const cx<double> xa0 = ldc(X2, a * v5 + c0 + s), xa1 = ldc(X2, a * v5 + c1 + s);
      const cx<double> xb0 = ldc(X2, b * v5 + c0 + s), xb1 = ldc(X2, b * v5 + c1 + s);
      sw += cjmul(r0, xa1).i - cjmul(r1, xa0).i;
      sd += cmul(g, cjmul(r0, xb1) + cjmul(r1, xb0)).i;
    }
    acc = acc + w1 * sw;
    acc = acc + w2 * sd;
  }
  ds[i] += acc;
}

// Wilson-kernel extra force terms (derivs_module.F90:41-97, 198-410): the
// s-couplings of the Wilson-kernel operator, each a derivSlice() between a
// slice of R and a slice of X2 with spin masks and a complex multiplier.
struct SlicePair {
  int sl, sr;       // 0-based s of the R slice and the X2 slice
  int l0, l1;       // R spins kept [l0, l1)
  int r0, r1;       // X2 spins kept [r0, r1)
  double cr, ci;    // cmult
};

__global__ __launch_bounds__(BLK) void k_derivs_wilson_extra(const Dims d, const double2 *__restrict__ R,
                                                             const double2 *__restrict__ X2,
                                                             float tzi, int iflag,
                                                             const SlicePair *pairs, int npairs,
                                                             double *ds) {
  const long v4 = d.v4(), v5 = d.v5();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v4 * 3) return;
  const int mu = (int)(i / v4);
  const long s4 = i % v4;
  const int x = (int)(s4 % d.lx), y = (int)((s4 / d.lx) % d.ly), t = (int)(s4 / ((long)d.lx * d.ly));
  int xf = x, yf = y, tf = t;
  if (mu == 0) xf = (x + 1 == d.lx) ? 0 : x + 1;
  if (mu == 1) yf = (y + 1 == d.ly) ? 0 : y + 1;
  if (mu == 2) tf = (t + 1 == d.lt) ? 0 : t + 1;
  const long f4 = xf + (long)d.lx * (yf + (long)d.ly * tf);
  const double w1 = (double)(tzi * 0.5f);
  const double w2 = (double)(tzi * (iflag == 1 ? -1.0f : 1.0f));
  double acc = ds[i];
  for (int k = 0; k < npairs; k++) {
    const SlicePair P = pairs[k];
    const cx<double> cm{P.cr, P.ci};
    const long l0 = s4 * d.ls + P.sl, l1 = f4 * d.ls + P.sl;
    const long r0 = s4 * d.ls + P.sr, r1 = f4 * d.ls + P.sr;
    for (int a = 0; a < 4; a++) {
      const bool la = a >= P.l0 && a < P.l1;
      const cx<double> zero{0.0, 0.0};
      const cx<double> L0 = la ? ldc(R, a * v5 + l0) : zero, L1 = la ? ldc(R, a * v5 + l1) : zero;
      const bool ra = a >= P.r0 && a < P.r1;
      const cx<double> Ra0 = ra ? ldc(X2, a * v5 + r0) : zero, Ra1 = ra ? ldc(X2, a * v5 + r1) : zero;
      const double v = cmul(cm, cjmul(L0, Ra1)).i - cmul(cm, cjmul(L1, Ra0)).i;
      acc += w1 * v;
      const int b = kGin[mu][a];
      const bool rb = b >= P.r0 && b < P.r1;
      const cx<double> Rb0 = rb ? ldc(X2, b * v5 + r0) : zero, Rb1 = rb ? ldc(X2, b * v5 + r1) : zero;
      const cx<double> g{0.5 * kGre[mu][a], 0.5 * kGim[mu][a]};
      acc += w2 * cmul(cmul(cm, g), cjmul(L0, Rb1) + cjmul(L1, Rb0)).i;
    }
  }
  ds[i] = acc;
}

}  // namespace

int launch_derivs_wilson(const Dims &d, const double2 *R, const double2 *X2, double anum,
                         int iflag, float am, int imass, double *dsdpi, void *pair_buf,
                         hipStream_t s) {
  THR_CHECK(imass == 1 || imass == 3, THR_ERR_ARG, "Wilson kernel supports imass 1 and 3");
  int rc = launch_derivs(d, R, X2, anum, iflag, dsdpi, s);
  if (rc) return rc;
  const int L = d.ls;
  std::vector<SlicePair> pr;
  const double m = (double)am;
  // derivsWilsonMass (derivs_module.F90:198-249): cmplx(-am,0) | cmplx(0,-am), then
  // cmplx(-am,0) | cmplx(0,am)
  const double c1r = imass == 1 ? -m : 0.0, c1i = imass == 1 ? 0.0 : -m;
  const double c2r = imass == 1 ? -m : 0.0, c2i = imass == 1 ? 0.0 : m;
  if (iflag == 1) {
    pr.push_back({0, L - 1, 0, 4, 0, 2, c1r, c1i});   // DP: L all, R spins 1:2
    pr.push_back({L - 1, 0, 0, 4, 2, 4, c2r, c2i});   // DM: L all, R spins 3:4
  } else if (iflag == 0) {
    pr.push_back({0, L - 1, 2, 4, 0, 4, c1r, c1i});   // MD: L spins 3:4, R all
    pr.push_back({L - 1, 0, 0, 2, 0, 4, c2r, c2i});   // PD: L spins 1:2, R all
  }
  // s-couplings (derivs_module.F90:73-81)
  for (int il = 0; il < L - 1; il++) {
    if (iflag == 1) {
      pr.push_back({il, il + 1, 0, 4, 2, 4, 1.0, 0.0});  // DM
      pr.push_back({il + 1, il, 0, 4, 0, 2, 1.0, 0.0});  // DP
    } else if (iflag == 0) {
      pr.push_back({il, il + 1, 0, 2, 0, 4, 1.0, 0.0});  // PD
      pr.push_back({il + 1, il, 2, 4, 0, 4, 1.0, 0.0});  // MD
    }
  }
  if (pr.empty()) return THR_OK;
  THR_HIP(hipMemcpyAsync(pair_buf, pr.data(), sizeof(SlicePair) * pr.size(), hipMemcpyHostToDevice, s));
  const float tzi = 2.0f * (float)anum;
  const unsigned nb = (unsigned)((d.v4() * 3 + BLK - 1) / BLK);
  hipLaunchKernelGGL(k_derivs_wilson_extra, dim3(nb), dim3(BLK), 0, s, d, R, X2, tzi, iflag,
                     (const SlicePair *)pair_buf, (int)pr.size(), dsdpi);
  THR_HIP(hipGetLastError());
  // pr must stay alive until the copy is done
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

int launch_dslash_wilson(const Dims &d, int nrhs, int dagger, int imass, float am, float am3,
                         const double2 *in, double2 *out, const double2 *links, double2 *tmp,
                         hipStream_t s) {
  THR_CHECK(imass == 1 || imass == 3, THR_ERR_ARG, "Wilson kernel supports imass 1 and 3");
  const long n = d.v5() * nrhs;
  const unsigned nb = (unsigned)((n + BLK - 1) / BLK);
  DslashLaunch L{};
  L.d = d;
  L.nrhs = nrhs;
  L.prec = P64;
  L.dagger = dagger;
  L.imass = imass;
  L.am3 = am3;
  L.links = links;
  L.mode = EPI_STORE;
  L.wilson4d = 1;
  const double m = (double)am;
  if (!dagger) {
    // c1 = -am | -i am (dirac.F90:274-278), c2 = -am | +i am (287-291)
    const double2 c1 = imass == 1 ? make_double2(-m, 0.0) : make_double2(0.0, -m);
    const double2 c2 = imass == 1 ? make_double2(-m, 0.0) : make_double2(0.0, m);
    hipLaunchKernelGGL(k_wk_pre, dim3(nb), dim3(BLK), 0, s, d, nrhs, in, tmp, c1, c2);
    THR_HIP(hipGetLastError());
    L.in = tmp;
    L.out = out;
    int rc = launch_dslash(L, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_wk_post, dim3((unsigned)((n * 4 + BLK - 1) / BLK)), dim3(BLK), 0, s,
                       n * 4, in, tmp, out);
  } else {
    L.in = in;
    L.out = tmp;
    int rc = launch_dslash(L, s);
    if (rc) return rc;
    // zkappa: -am | -i am (dirac.F90:602-606), -am | +i am (615-619)
    const double2 z1 = imass == 1 ? make_double2(-m, 0.0) : make_double2(0.0, -m);
    const double2 z2 = imass == 1 ? make_double2(-m, 0.0) : make_double2(0.0, m);
    hipLaunchKernelGGL(k_wkd_post, dim3(nb), dim3(BLK), 0, s, d, nrhs, in, tmp, out, z1, z2);
  }
  THR_HIP(hipGetLastError());
  return THR_OK;
}

DerivPair derivs_pair(const double2 *R, const double2 *X2, double anum, int iflag) {
  const float tzi = 2.0f * (float)anum;  // real :: tzi_real = 2*real(anum)
  return DerivPair{R, X2, (double)(tzi * 0.5f), (double)tzi * (iflag == 0 ? 1.0 : -1.0)};
}

// the batched derivs' LDS x exchange applies with 64-lane segments (32 < Ls
// <= 64) when a block's 4 lines are 4 consecutive x of one y (Lx % 4 == 0);
// THR_DERIVS_LDSX=0 switches it off
static bool derivs_lds_x(const Dims &d, unsigned wseg) {
  if (const char *e = std::getenv("THR_DERIVS_LDSX"))
    if (std::atoi(e) == 0) return false;
  return wseg == 64 && d.lx % 4 == 0;
}

int launch_derivs_batch(const Dims &d, int npair, const DerivPair *pairs_host,
                        DerivPair *pairs_dev, double *part, double *dsdpi, hipStream_t s) {
  THR_CHECK(npair >= 1 && npair <= 65535, THR_ERR_ARG, "derivs batch: 1..65535 pairs");
  THR_CHECK(d.v5() * 4 < (1L << 31), THR_ERR_SIZE, "derivs: 4*V5 must be below 2^31");
  THR_CHECK(d.ls <= 64, THR_ERR_ARG, "derivs batch: Ls above 64");
  THR_CHECK(pairs_dev && part, THR_ERR_ARG, "derivs batch: NULL buffer");
  const long n = 3 * d.v4();
  THR_HIP(hipMemcpyAsync(pairs_dev, pairs_host, sizeof(DerivPair) * npair, hipMemcpyHostToDevice,
                         s));
  unsigned w = 1;
  while (w < (unsigned)d.ls) w <<= 1;
  const unsigned nl = BLK / w;
  const unsigned v2 = (unsigned)(d.lx * d.ly);
  const unsigned nb = ((v2 + nl - 1) / nl) * (unsigned)d.lt;
  if (derivs_lds_x(d, w))
    hipLaunchKernelGGL(k_derivs_seg_batch<true>, dim3(nb, npair), dim3(BLK), 0, s, d,
                       (const DerivPair *)pairs_dev, w, part);
  else
    hipLaunchKernelGGL(k_derivs_seg_batch<false>, dim3(nb, npair), dim3(BLK), 0, s, d,
                       (const DerivPair *)pairs_dev, w, part);
  hipLaunchKernelGGL(k_derivs_sum_pairs, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, s, n,
                     npair, (const double *)part, dsdpi);
  THR_HIP(hipGetLastError());
  // the host pair array must outlive the copy
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

int launch_derivs(const Dims &d, const double2 *R, const double2 *X2, double anum, int iflag,
                  double *dsdpi, hipStream_t s) {
  const float tzi = 2.0f * (float)anum;  // real :: tzi_real = 2*real(anum)
  THR_CHECK(d.v5() * 4 < (1L << 31), THR_ERR_SIZE, "derivs: 4*V5 must be below 2^31");
  // one t step per thread: marching longer t-chunks reuses the t neighbour
  // from registers but leaves fewer threads in flight, and measured slower
  // (32^3 x 64: 76-80 us at one step, 82-90 us at 2-16; 16^3 x 48: 12-13
  // against 14-44 us) (THR_DERIVS_TCHUNK overrides)
  unsigned tchunk = 1;
  if (const char *e = std::getenv("THR_DERIVS_TCHUNK"))
    tchunk = (unsigned)std::max(1, std::min(d.lt, std::atoi(e)));
  // resident blocks per CU capped by LDS padding (THR_DERIVS_OCC = k: k + 1
  // blocks): fewer resident s-lines per XCD keep the x+1 / y+1 neighbours in L2
  size_t pad = 0;
  if (const char *e = std::getenv("THR_DERIVS_OCC")) {
    const int k = std::max(0, std::min(3, std::atoi(e)));
    if (k > 0) pad = (size_t)(160 * 1024 / (k + 1)) & ~(size_t)511;
    if (pad > 65536) pad = 65536;
  }
  const unsigned nchunk = (d.lt + tchunk - 1) / tchunk;
  const unsigned v2 = (unsigned)d.lx * d.ly;
  if (d.ls <= 64) {
    unsigned w = 1;
    while (w < (unsigned)d.ls) w <<= 1;
    const unsigned nl = BLK / w;
    const unsigned nb = ((v2 + nl - 1) / nl) * nchunk;
    hipLaunchKernelGGL(k_derivs_seg, dim3(nb), dim3(BLK), pad, s, d, R, X2, tzi, iflag, w, tchunk,
                       dsdpi);
    THR_HIP(hipGetLastError());
    return THR_OK;
  }
  if (d.ls <= BLK) {
    // whole s-lines per block, summed in LDS
    const unsigned nl = BLK / d.ls;
    const unsigned nb = ((v2 + nl - 1) / nl) * nchunk;
    hipLaunchKernelGGL(k_derivs_march, dim3(nb), dim3(nl * d.ls), 0, s, d, R, X2, tzi, iflag, nl,
                       tchunk, dsdpi);
    THR_HIP(hipGetLastError());
    return THR_OK;
  }
  const unsigned nb = (unsigned)((d.v4() * 3 + BLK - 1) / BLK);
  hipLaunchKernelGGL(k_derivs, dim3(nb), dim3(BLK), 0, s, d, R, X2, tzi, iflag, dsdpi);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

}  // namespace thr
