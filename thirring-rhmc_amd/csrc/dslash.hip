// Shamir domain-wall Dirac operator M and M^dagger on gfx950.
//
// Restates dirac.F90:29-179 (dslash_shamir) and dirac.F90:303-494
// (dslashd_local_shamir + dslashd_shamir) as one fused stencil pass:
//   Phi(s,x,a) = D R(s,x,a)
//      + sum_mu [ -1/2 (U_mu(x) R(s,x+mu,a) + U_mu(x-mu)^* R(s,x-mu,a))
//                 +/- 1/2 g_mu(a) (U_mu(x) R(s,x+mu,b) - U_mu(x-mu)^* R(s,x-mu,b)) ]
//      - P-/P+ hops along s, + wall mass coupling s=1 <-> s=Ls (imass 1/3/5)
// with D = (3 - am3) + 1 in float32 and am in float32.
//
// One thread computes the four spins of one 5-D site.  Threads are laid out
// with s fastest, so every spin plane of a wave's loads is one contiguous run
// of the field (coalesced 16-byte lanes for FP64).  Links are complex scalars
// (U(1)), so link traffic is 1/(8 Ls) of the spinor traffic.  The fifth-
// dimension neighbours s+-1 are the adjacent lanes' centre loads and hit in
// L1; the spin projection and gamma algebra are register arithmetic (no
// MFMA: the operator is a sparse stencil, not a dense contraction).
#include "kernels.h"

namespace thr {

namespace {

template <int PREC> struct Acc {
  using C = typename Store<PREC>::C;
  using T = typename Store<PREC>::T;
  const T *__restrict__ p;
  const float *__restrict__ nrm;
  unsigned v5;
  __device__ inline void load4(unsigned site, cx<C> v[4]) const {
    if constexpr (PREC == P16) {
      const float sc = nrm[site];
#pragma unroll
      for (int a = 0; a < 4; a++) v[a] = ldh(p, a * v5 + site, sc);
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++) v[a] = ldc(p, a * v5 + site);
    }
  }
  __device__ inline cx<C> load1(unsigned site, unsigned a) const {
    if constexpr (PREC == P16) {
      return ldh(p, a * v5 + site, nrm[site]);
    } else {
      return ldc(p, a * v5 + site);
    }
  }
};

template <int PREC> struct Out {
  using C = typename Store<PREC>::C;
  using T = typename Store<PREC>::T;
  T *__restrict__ p;
  float *__restrict__ nrm;
  unsigned v5;
  __device__ inline void store4(unsigned site, const cx<C> v[4]) const {
    if constexpr (PREC == P16) {
      float m = 0.f;
#pragma unroll
      for (int a = 0; a < 4; a++) m = fmaxf(m, fmaxf(fabsf(v[a].r), fabsf(v[a].i)));
      const float inv = m > 0.f ? 1.0f / m : 0.f;
      nrm[site] = m;
#pragma unroll
      for (int a = 0; a < 4; a++) p[a * v5 + site] = __floats2half2_rn(v[a].r * inv, v[a].i * inv);
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++) st(p, a * v5 + site, v[a]);
    }
  }
};

template <typename C> __device__ inline cx<C> link(const void *links, long i) {
  if constexpr (sizeof(C) == 8) {
    const double2 v = ((const double2 *)links)[i];
    return {v.x, v.y};
  } else {
    const float2 v = ((const float2 *)links)[i];
    return {v.x, v.y};
  }
}

// o[a] += 1/2 gamval(mu,a) D[gamin(mu,a)]  (gammamatrices.F90:27-75, kappa = 1/2)
template <typename C> __device__ inline void add_gamma(int mu, cx<C> o[4], const cx<C> D[4]) {
  const C h = C(0.5);
  if (mu == 0) {  // gamma_1: (-i,-i,i,i) at (4,3,2,1)
    o[0] = o[0] - h * muli(D[3]);
    o[1] = o[1] - h * muli(D[2]);
    o[2] = o[2] + h * muli(D[1]);
    o[3] = o[3] + h * muli(D[0]);
  } else if (mu == 1) {  // gamma_2: (-1,1,1,-1) at (4,3,2,1)
    o[0] = o[0] - h * D[3];
    o[1] = o[1] + h * D[2];
    o[2] = o[2] + h * D[1];
    o[3] = o[3] - h * D[0];
  } else {  // gamma_3: (-i,i,i,-i) at (3,4,1,2)
    o[0] = o[0] - h * muli(D[2]);
    o[1] = o[1] + h * muli(D[3]);
    o[2] = o[2] + h * muli(D[0]);
    o[3] = o[3] - h * muli(D[1]);
  }
}

template <int PREC, bool DAG>
__global__ __launch_bounds__(BLK) void k_dslash(const DslashLaunch L) {
  using C = typename Store<PREC>::C;
  using T = typename Store<PREC>::T;
  const int rhs = blockIdx.y;
  if (L.done && L.done[(long)rhs * L.done_stride]) return;
  // 32-bit index arithmetic: 4*V5 < 2^31 is checked on the host
  const unsigned lx = L.d.lx, ly = L.d.ly, lt = L.d.lt, ls = L.d.ls;
  const unsigned v4 = lx * ly * lt, v5 = v4 * ls;
  const unsigned i = blockIdx.x * BLK + threadIdx.x;
  double red = 0.0;
  if (i < v5) {
    const unsigned s = i % ls;
    const unsigned site4 = i / ls;
    const unsigned x = site4 % lx;
    const unsigned yt = site4 / lx;
    const unsigned y = yt % ly;
    const unsigned t = yt / ly;
    const long roff = (long)rhs * 4 * v5;
    Acc<PREC> in{(const T *)L.in + roff, L.in_norm ? L.in_norm + (long)rhs * v5 : nullptr, v5};

    cx<C> o[4], c[4];
    in.load4(i, c);
    // D = (3.0 - am3) + 1.0 in default real (dirac.F90:45); the 4-D Wilson slice
    // operator DWilson(.., -am3) has diag = 3.0 + (-am3) (dirac.F90:191)
    const C diag = L.wilson4d ? (C)(3.0f + (-L.am3)) : (C)((3.0f - L.am3) + 1.0f);
#pragma unroll
    for (int a = 0; a < 4; a++) o[a] = diag * c[a];

#pragma unroll
    for (int mu = 0; mu < 3; mu++) {
      unsigned xf = x, yf = y, tf = t, xb = x, yb = y, tb = t;
      if (mu == 0) {
        xf = (x + 1 == lx) ? 0 : x + 1;
        xb = (x == 0) ? lx - 1 : x - 1;
      } else if (mu == 1) {
        yf = (y + 1 == ly) ? 0 : y + 1;
        yb = (y == 0) ? ly - 1 : y - 1;
      } else {
        tf = (t + 1 == lt) ? 0 : t + 1;
        tb = (t == 0) ? lt - 1 : t - 1;
      }
      const unsigned sf4 = xf + lx * (yf + ly * tf);
      const unsigned sb4 = xb + lx * (yb + ly * tb);
      const cx<C> uf = link<C>(L.links, mu * v4 + site4);
      const cx<C> ub = conj(link<C>(L.links, mu * v4 + sb4));
      cx<C> rf[4], rb[4];
      in.load4(sf4 * ls + s, rf);
      in.load4(sb4 * ls + s, rb);
      cx<C> W[4], D[4];
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const cx<C> A = cmul(uf, rf[a]);
        const cx<C> B = cmul(ub, rb[a]);
        W[a] = A + B;
        D[a] = DAG ? (B - A) : (A - B);
      }
#pragma unroll
      for (int a = 0; a < 4; a++) o[a] = o[a] - C(0.5) * W[a];
      add_gamma<C>(mu, o, D);
    }

    // hops along s (dirac.F90:80-87 for M, 317-324 for M^dagger): spins that
    // hop from s+1 are {3,4} for M and {1,2} for M^dagger; the others from s-1.
    const int up0 = DAG ? 0 : 2, dn0 = DAG ? 2 : 0;
    const int imass = L.wilson4d ? 0 : L.imass;
    if (!L.wilson4d && s + 1 < ls) {
      o[up0] = o[up0] - in.load1(i + 1, up0);
      o[up0 + 1] = o[up0 + 1] - in.load1(i + 1, up0 + 1);
    }
    if (!L.wilson4d && s > 0) {
      o[dn0] = o[dn0] - in.load1(i - 1, dn0);
      o[dn0 + 1] = o[dn0 + 1] - in.load1(i - 1, dn0 + 1);
    }
    // wall mass (dirac.F90:119-176 for M, 356-402 for M^dagger)
    const C m = (C)L.am[rhs];
    const unsigned col = site4 * ls;  // s = 0 of this (x,y,t)
    if (imass == 1 || imass == 3) {
      // top: s = Ls couples to s = 1 on the up0 spins; bottom: s = 1 couples to s = Ls
      if (s == ls - 1) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
          const cx<C> r0 = in.load1(col, up0 + k);
          o[up0 + k] = o[up0 + k] + (imass == 1 ? m * r0 : m * muli(r0));
        }
      }
      if (s == 0) {
#pragma unroll
        for (int k = 0; k < 2; k++) {
          const cx<C> rL = in.load1(col + ls - 1, dn0 + k);
          o[dn0 + k] = o[dn0 + k] + (imass == 1 ? m * rL : (-m) * muli(rL));
        }
      }
    } else if (imass == 5) {
      // same-s chirality flip at both walls: +i m (M), -i m (M^dagger)
      const C sg = DAG ? C(-1) : C(1);
      if (s == ls - 1) {
#pragma unroll
        for (int k = 0; k < 2; k++)
          o[up0 + k] = o[up0 + k] + (sg * m) * muli(c[dn0 + k]);
      }
      if (s == 0) {
#pragma unroll
        for (int k = 0; k < 2; k++)
          o[dn0 + k] = o[dn0 + k] + (sg * m) * muli(c[up0 + k]);
      }
    }

    // epilogue
    if (L.mode == EPI_CG) {
      const C a = (C)L.a_ptr[(long)rhs * L.sc_stride];
      T *r = (T *)L.r + roff;
      T *ro = L.r_out ? (T *)L.r_out + roff : r;
      Acc<PREC> rr{r, L.r_norm ? L.r_norm + (long)rhs * v5 : nullptr, v5};
      cx<C> rv[4];
      rr.load4(i, rv);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        rv[k] = rv[k] - a * o[k];
        red += (double)norm2(rv[k]);
      }
      Out<PREC> rout{ro, L.r_out_norm ? L.r_out_norm + (long)rhs * v5 : nullptr, v5};
      rout.store4(i, rv);
      if (L.b_ptr) {
        const C b = (C)L.b_ptr[(long)rhs * L.sc_stride];
        Acc<PREC> pp{(const T *)L.p + roff, nullptr, v5};
        Acc<PREC> xx{(const T *)L.x + roff, nullptr, v5};
        cx<C> pv[4], xv[4];
        pp.load4(i, pv);
        xx.load4(i, xv);
#pragma unroll
        for (int k = 0; k < 4; k++) xv[k] = xv[k] + b * pv[k];
        Out<PREC> xo{(T *)L.x + roff, nullptr, v5};
        xo.store4(i, xv);
      }
    } else {
      Out<PREC> out{(T *)L.out + roff, L.out_norm ? L.out_norm + (long)rhs * v5 : nullptr, v5};
      out.store4(i, o);
      if (L.mode == EPI_NORM) {
#pragma unroll
        for (int k = 0; k < 4; k++) red += (double)norm2(o[k]);
      } else if (L.mode == EPI_DOTW) {
        Acc<PREC> w{(const T *)L.w + roff, nullptr, v5};
        cx<C> wv[4];
        w.load4(i, wv);
#pragma unroll
        for (int k = 0; k < 4; k++) red += (double)(wv[k].r * o[k].r + wv[k].i * o[k].i);
      }
    }
  }
  if (L.mode != EPI_STORE) {
    const double tot = block_sum(red);
    if (threadIdx.x == 0) L.partial[(long)rhs * gridDim.x + blockIdx.x] = tot;
  }
}

}  // namespace

long dslash_blocks(const Dims &d) { return (d.v5() + BLK - 1) / BLK; }

int launch_dslash(const DslashLaunch &L, hipStream_t s) {
  THR_CHECK(L.nrhs >= 1 && L.nrhs <= MAXRHS, THR_ERR_ARG, "nrhs out of range");
  THR_CHECK(L.d.lx > 0 && L.d.ly > 0 && L.d.lt > 0 && L.d.ls > 0, THR_ERR_ARG, "bad dims");
  THR_CHECK(L.imass == 1 || L.imass == 3 || L.imass == 5, THR_ERR_ARG, "imass must be 1, 3 or 5");
  THR_CHECK(L.mode == EPI_STORE || L.partial, THR_ERR_ARG, "partial buffer required");
  THR_CHECK(L.prec != P16 || L.mode != EPI_CG || (L.r_norm && !L.b_ptr &&
                                                   (L.r_out ? L.r_out_norm != nullptr : true)),
            THR_ERR_ARG, "fp16 CG epilogue needs r scales and no x update");
  THR_CHECK(L.prec != P16 || L.mode != EPI_DOTW, THR_ERR_ARG, "fp16 dot epilogue unsupported");
  dim3 grid((unsigned)dslash_blocks(L.d), (unsigned)L.nrhs);
  dim3 block(BLK);
#define THR_LAUNCH(P, D) hipLaunchKernelGGL((k_dslash<P, D>), grid, block, 0, s, L)
  if (L.prec == P64) {
    if (L.dagger) THR_LAUNCH(P64, true); else THR_LAUNCH(P64, false);
  } else if (L.prec == P32) {
    if (L.dagger) THR_LAUNCH(P32, true); else THR_LAUNCH(P32, false);
  } else {
    if (L.dagger) THR_LAUNCH(P16, true); else THR_LAUNCH(P16, false);
  }
#undef THR_LAUNCH
  THR_HIP(hipGetLastError());
  return THR_OK;
}

// ---------------------------------------------------------------------------
// layout conversion
// ---------------------------------------------------------------------------
namespace {

__device__ inline long hidx(const Dims &d, int s, int x, int y, int t, int a) {
  const long S = d.ls + 2, X = d.lx + 2, Y = d.ly + 2, T = d.lt + 2;
  return s + S * (x + X * (y + Y * (t + T * (long)a)));
}

template <int PREC>
__global__ __launch_bounds__(BLK) void k_pack(const Dims d, const double2 *__restrict__ h,
                                              typename Store<PREC>::T *dst, float *nrm) {
  const long v5 = d.v5();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v5) return;
  const int s = (int)(i % d.ls);
  const long s4 = i / d.ls;
  const int x = (int)(s4 % d.lx), y = (int)((s4 / d.lx) % d.ly), t = (int)(s4 / ((long)d.lx * d.ly));
  using C = typename Store<PREC>::C;
  cx<C> v[4];
#pragma unroll
  for (int a = 0; a < 4; a++) {
    const double2 z = h[hidx(d, s + 1, x + 1, y + 1, t + 1, a)];
    v[a] = {(C)z.x, (C)z.y};
  }
  Out<PREC> o{dst, nrm, (unsigned)v5};
  o.store4((unsigned)i, v);
}

template <int PREC>
__global__ __launch_bounds__(BLK) void k_unpack(const Dims d, const typename Store<PREC>::T *src,
                                                const float *nrm, double2 *__restrict__ h,
                                                int fill_halo) {
  const long S = d.ls + 2, X = d.lx + 2, Y = d.ly + 2, T = d.lt + 2;
  const long nh = S * X * Y * T;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= nh) return;
  int s = (int)(i % S);
  long r = i / S;
  int x = (int)(r % X);
  r /= X;
  int y = (int)(r % Y);
  int t = (int)(r / Y);
  const bool halo = s == 0 || s == S - 1 || x == 0 || x == X - 1 || y == 0 || y == Y - 1 ||
                    t == 0 || t == T - 1;
  if (halo && !fill_halo) return;
  // periodic wrap to interior (0-based device coordinates)
  const int si = s == 0 ? d.ls - 1 : (s == S - 1 ? 0 : s - 1);
  const int xi = x == 0 ? d.lx - 1 : (x == X - 1 ? 0 : x - 1);
  const int yi = y == 0 ? d.ly - 1 : (y == Y - 1 ? 0 : y - 1);
  const int ti = t == 0 ? d.lt - 1 : (t == T - 1 ? 0 : t - 1);
  const long v5 = d.v5();
  const long site = si + (long)d.ls * (xi + (long)d.lx * (yi + (long)d.ly * ti));
  Acc<PREC> in{src, nrm, (unsigned)v5};
  cx<typename Store<PREC>::C> v[4];
  in.load4((unsigned)site, v);
#pragma unroll
  for (int a = 0; a < 4; a++) h[i + nh * a] = make_double2((double)v[a].r, (double)v[a].i);
}

__global__ __launch_bounds__(BLK) void k_pack_links(const Dims d, const double2 *__restrict__ h,
                                                    double2 *d64, float2 *d32) {
  const long v4 = d.v4();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v4 * 3) return;
  const int mu = (int)(i / v4);
  const long s4 = i % v4;
  const int x = (int)(s4 % d.lx), y = (int)((s4 / d.lx) % d.ly), t = (int)(s4 / ((long)d.lx * d.ly));
  const long X = d.lx + 2, Y = d.ly + 2, T = d.lt + 2;
  const double2 z = h[(x + 1) + X * ((y + 1) + Y * ((t + 1) + T * (long)mu))];
  if (d64) d64[i] = z;
  if (d32) d32[i] = make_float2((float)z.x, (float)z.y);
}

template <int SP, int DP>
__global__ __launch_bounds__(BLK) void k_convert(long nsite, long v5, const void *src,
                                                 const float *snrm, void *dst, float *dnrm) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= nsite) return;
  const long rhs = i / v5, site = i % v5;
  Acc<SP> in{(const typename Store<SP>::T *)src + rhs * 4 * v5, snrm ? snrm + rhs * v5 : nullptr, (unsigned)v5};
  using CS = typename Store<SP>::C;
  using CD = typename Store<DP>::C;
  cx<CS> v[4];
  in.load4((unsigned)site, v);
  cx<CD> w[4];
#pragma unroll
  for (int a = 0; a < 4; a++) w[a] = {(CD)v[a].r, (CD)v[a].i};
  Out<DP> o{(typename Store<DP>::T *)dst + rhs * 4 * v5, dnrm ? dnrm + rhs * v5 : nullptr, (unsigned)v5};
  o.store4((unsigned)site, w);
}

__global__ __launch_bounds__(BLK) void k_coef(const Dims d, const float *theta, int ibound,
                                              double2 *u) {
  const long v4 = d.v4();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= v4 * 3) return;
  const int mu = (int)(i / v4);
  const int t = (int)((i % v4) / ((long)d.lx * d.ly));
  double re = (double)1.0f, im = (double)theta[i];
  if (ibound == -1 && mu == 2 && t == d.lt - 1) {
    re = -re;
    im = -im;
  }
  u[i] = make_double2(re, im);
}

}  // namespace

int launch_pack_spinor(const Dims &d, const double2 *haloed, void *dst, float *dst_norm, int prec,
                       hipStream_t s) {
  const unsigned nb = (unsigned)((d.v5() + BLK - 1) / BLK);
  if (prec == P64)
    hipLaunchKernelGGL(k_pack<P64>, dim3(nb), dim3(BLK), 0, s, d, haloed, (double2 *)dst, nullptr);
  else if (prec == P32)
    hipLaunchKernelGGL(k_pack<P32>, dim3(nb), dim3(BLK), 0, s, d, haloed, (float2 *)dst, nullptr);
  else
    hipLaunchKernelGGL(k_pack<P16>, dim3(nb), dim3(BLK), 0, s, d, haloed, (__half2 *)dst, dst_norm);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int launch_unpack_spinor(const Dims &d, const void *src, const float *src_norm, int prec,
                         double2 *haloed, int fill_halo, hipStream_t s) {
  const long nh = (long)(d.ls + 2) * (d.lx + 2) * (d.ly + 2) * (d.lt + 2);
  const unsigned nb = (unsigned)((nh + BLK - 1) / BLK);
  if (prec == P64)
    hipLaunchKernelGGL(k_unpack<P64>, dim3(nb), dim3(BLK), 0, s, d, (const double2 *)src, nullptr,
                       haloed, fill_halo);
  else if (prec == P32)
    hipLaunchKernelGGL(k_unpack<P32>, dim3(nb), dim3(BLK), 0, s, d, (const float2 *)src, nullptr,
                       haloed, fill_halo);
  else
    hipLaunchKernelGGL(k_unpack<P16>, dim3(nb), dim3(BLK), 0, s, d, (const __half2 *)src, src_norm,
                       haloed, fill_halo);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int launch_pack_links(const Dims &d, const double2 *haloed, double2 *dst64, float2 *dst32,
                      hipStream_t s) {
  const unsigned nb = (unsigned)((d.v4() * 3 + BLK - 1) / BLK);
  hipLaunchKernelGGL(k_pack_links, dim3(nb), dim3(BLK), 0, s, d, haloed, dst64, dst32);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int launch_coef(const Dims &d, const float *theta, int ibound, double2 *links, hipStream_t s) {
  const unsigned nb = (unsigned)((d.v4() * 3 + BLK - 1) / BLK);
  hipLaunchKernelGGL(k_coef, dim3(nb), dim3(BLK), 0, s, d, theta, ibound, links);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

// precision conversion between device spinor batches (nrhs spinors of v5 sites)
int convert_spinors(const Dims &d, int nrhs, const void *src, const float *snrm, int sprec,
                    void *dst, float *dnrm, int dprec, hipStream_t s) {
  const long v5 = d.v5();
  const long nsite = v5 * nrhs;
  const unsigned nb = (unsigned)((nsite + BLK - 1) / BLK);
#define CONV(A, B) hipLaunchKernelGGL((k_convert<A, B>), dim3(nb), dim3(BLK), 0, s, nsite, v5, src, snrm, dst, dnrm)
  if (sprec == P64 && dprec == P32) CONV(P64, P32);
  else if (sprec == P64 && dprec == P16) CONV(P64, P16);
  else if (sprec == P32 && dprec == P64) CONV(P32, P64);
  else if (sprec == P32 && dprec == P16) CONV(P32, P16);
  else if (sprec == P16 && dprec == P32) CONV(P16, P32);
  else if (sprec == P16 && dprec == P64) CONV(P16, P64);
  else if (sprec == P64 && dprec == P64) CONV(P64, P64);
  else if (sprec == P32 && dprec == P32) CONV(P32, P32);
  else CONV(P16, P16);
#undef CONV
  THR_HIP(hipGetLastError());
  return THR_OK;
}

}  // namespace thr
