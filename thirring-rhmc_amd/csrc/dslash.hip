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
#include <algorithm>
#include <cstdlib>
#include <map>
#include <mutex>
#include <tuple>
#include <type_traits>
#include <vector>

#include "dslash_site.h"

namespace thr {

namespace {

// Spinor layouts: FP64 stores spin planes [spin][site]; FP32 and FP16 store
// the 4 spins of a site together [site][spin] (FP32: 32 bytes, two 128-bit
// accesses; FP16: 16 bytes, one, with a per-site float scale).  A site's FP32
// spins as spin planes would take four 64-bit accesses: twice the address
// work per byte of the FP64 planes, which kept the FP32 kernel bound by the
// texture-address units.
template <int PREC> struct Acc {
  using C = typename Store<PREC>::C;
  using T = typename Store<PREC>::T;
  const T *__restrict__ p;
  const float *__restrict__ nrm;
  unsigned v5;
  __device__ inline void load4(unsigned site, cx<C> v[4]) const {
    if constexpr (PREC == P16) {
      const float sc = nrm[site];
      const uint4 q = *reinterpret_cast<const uint4 *>(p + 4 * (long)site);
      const unsigned w[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
      for (int a = 0; a < 4; a++) {
        const float2 f = __half22float2(__builtin_bit_cast(__half2, w[a]));
        v[a] = {f.x * sc, f.y * sc};
      }
    } else if constexpr (PREC == P32) {
      const float4 q0 = *reinterpret_cast<const float4 *>(p + 4 * (long)site);
      const float4 q1 = *reinterpret_cast<const float4 *>(p + 4 * (long)site + 2);
      v[0] = {q0.x, q0.y};
      v[1] = {q0.z, q0.w};
      v[2] = {q1.x, q1.y};
      v[3] = {q1.z, q1.w};
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++) v[a] = ldc(p, a * v5 + site);
    }
  }
  __device__ inline cx<C> load1(unsigned site, unsigned a) const {
    if constexpr (PREC == P16) {
      return ldh(p, 4 * (long)site + a, nrm[site]);
    } else if constexpr (PREC == P32) {
      return ldc(p, 4 * (long)site + a);
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
      for (int a = 0; a < 4; a++) p[4 * (long)site + a] = __floats2half2_rn(v[a].r * inv, v[a].i * inv);
    } else if constexpr (PREC == P32) {
#pragma unroll
      for (int a = 0; a < 4; a++) st(p, 4 * (long)site + a, v[a]);
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++) st(p, a * v5 + site, v[a]);
    }
  }
};

// Output stores through buffer resources; NT marks them non-temporal (the
// output is not read again by this kernel, so it should not displace the
// input lines that neighbouring waves still need from L2).
// FP64: one resource per spin plane, stored with soffset 0 (FP32 and FP16:
// one resource over the site-major field, soffset 0): a store of more than
// 64 bits reads its data VGPRs after issue, and the compiler's wait-state
// check for a following VALU write to those VGPRs skips stores whose soffset
// is an SGPR (the data can then be overwritten before it is read: measured as
// corrupted imaginary parts in FP64 output when other kernels share the CU).
template <int PREC, bool NT> struct BOut {
  using C = typename Store<PREC>::C;
  __amdgpu_buffer_rsrc_t rs[4], rn;
  __device__ inline BOut(void *p, float *nrm, unsigned v5) {
    const unsigned plane = v5 * (unsigned)Store<PREC>::bytes;
    if constexpr (PREC != P64) {  // [site][spin]: one resource over the field
      rs[0] = __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)(4u * plane), 0x00020000);
      rs[1] = rs[2] = rs[3] = rs[0];
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++)
        rs[a] = __builtin_amdgcn_make_buffer_rsrc((char *)p + (size_t)a * plane, (short)0,
                                                  (int)plane, 0x00020000);
    }
    rn = __builtin_amdgcn_make_buffer_rsrc(nrm, (short)0, nrm ? (int)(v5 * 4u) : 0, 0x00020000);
  }
  __device__ inline void store4(unsigned site, const cx<C> v[4]) const {
    constexpr int aux = NT ? 2 : 0;  // gfx9.4+: bit 1 = non-temporal
    const unsigned off = site * (unsigned)Store<PREC>::bytes;
    if constexpr (PREC == P64) {
#pragma unroll
      for (int a = 0; a < 4; a++)
        __builtin_amdgcn_raw_buffer_store_b128(
            __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, make_double2(v[a].r, v[a].i)),
            rs[a], off, 0, aux);
    } else if constexpr (PREC == P32) {  // [site][spin]: two 128-bit stores
      const unsigned o32 = site * 32u;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const float4 q = make_float4(v[2 * h].r, v[2 * h].i, v[2 * h + 1].r, v[2 * h + 1].i);
        __builtin_amdgcn_raw_buffer_store_b128(
            __builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned, q), rs[0],
            o32 + 16u * h, 0, aux);
      }
    } else {
      float m = 0.f;
#pragma unroll
      for (int a = 0; a < 4; a++) m = fmaxf(m, fmaxf(fabsf(v[a].r), fabsf(v[a].i)));
      const float inv = m > 0.f ? 1.0f / m : 0.f;
      __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(m), rn, site * 4u, 0, aux);
      __attribute__((ext_vector_type(4))) unsigned q;
#pragma unroll
      for (int a = 0; a < 4; a++)
        q[a] = __builtin_bit_cast(unsigned, __floats2half2_rn(v[a].r * inv, v[a].i * inv));
      __builtin_amdgcn_raw_buffer_store_b128(q, rs[0], site * 16u, 0, aux);
    }
  }
};

// Buffer-resource loads: base and size live in SGPRs, each lane supplies a
// 32-bit byte offset and the spin plane is a scalar offset, so the four spin
// loads of a site share one VGPR address (global loads need a 64-bit VGPR
// pair each).  Out-of-range offsets return 0 instead of faulting.
constexpr int BUF_DWORD3 = 0x00020000;  // gfx9 raw buffer: 32-bit data format

template <int PREC> struct BAcc {
  using C = typename Store<PREC>::C;
  __amdgpu_buffer_rsrc_t rs, rn;
  unsigned plane;  // bytes per spin plane
  __device__ inline BAcc(const void *p, const float *nrm, unsigned v5) {
    plane = v5 * (unsigned)Store<PREC>::bytes;
    rs = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)(4u * plane),
                                           BUF_DWORD3);
    rn = __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(nrm), (short)0,
                                           nrm ? (int)(v5 * 4u) : 0, BUF_DWORD3);
  }
  __device__ inline cx<C> ld(unsigned site, unsigned a, float sc) const {
    const unsigned off = site * (unsigned)Store<PREC>::bytes;
    if constexpr (PREC == P64) {
      const double2 v =
          __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(rs, off, a * plane, 0));
      return {v.x, v.y};
    } else if constexpr (PREC == P32) {  // [site][spin]
      const float2 v = __builtin_bit_cast(
          float2, __builtin_amdgcn_raw_buffer_load_b64(rs, site * 32u + a * 8u, 0, 0));
      return {v.x, v.y};
    } else {  // [site][spin]
      const __half2 h =
          __builtin_bit_cast(__half2, __builtin_amdgcn_raw_buffer_load_b32(rs, site * 16u + a * 4u, 0, 0));
      const float2 f = __half22float2(h);
      return {f.x * sc, f.y * sc};
    }
  }
  __device__ inline float scale(unsigned site) const {
    if constexpr (PREC == P16)
      return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rn, site * 4u, 0, 0));
    else
      return 1.0f;
  }
  __device__ inline void load4(unsigned site, cx<C> v[4]) const {
    const float sc = scale(site);
    if constexpr (PREC == P32) {  // the site's 4 spins in two 128-bit loads
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const auto q = __builtin_amdgcn_raw_buffer_load_b128(rs, site * 32u + 16u * h, 0, 0);
        // the elements as values first (see the FP16 branch)
        const unsigned e0 = q[0], e1 = q[1], e2 = q[2], e3 = q[3];
        v[2 * h] = {__uint_as_float(e0), __uint_as_float(e1)};
        v[2 * h + 1] = {__uint_as_float(e2), __uint_as_float(e3)};
      }
    } else if constexpr (PREC == P16) {  // the site's 4 spins in one 128-bit load
      const auto q = __builtin_amdgcn_raw_buffer_load_b128(rs, site * 16u, 0, 0);
#pragma unroll
      for (int a = 0; a < 4; a++) {
        // take the element as a value first: clang's bit_cast of an
        // ext-vector element lvalue reads the vector's first element
        const unsigned e = q[a];
        const float2 f = __half22float2(__builtin_bit_cast(__half2, e));
        v[a] = {f.x * sc, f.y * sc};
      }
    } else {
#pragma unroll
      for (int a = 0; a < 4; a++) v[a] = ld(site, a, sc);
    }
  }
  __device__ inline cx<C> load1(unsigned site, unsigned a) const { return ld(site, a, scale(site)); }
};

template <int PREC> __device__ inline Site<PREC> load_site(const BAcc<PREC> &in, unsigned site) {
  Site<PREC> S;
  if constexpr (PREC == P16) {  // the site's 4 spins in one 128-bit load
    const auto q = __builtin_amdgcn_raw_buffer_load_b128(in.rs, site * 16u, 0, 0);
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const unsigned e = q[a];  // the element as a value first (see BAcc::load4)
      S.h[a] = e;
    }
    S.sc = in.scale(site);
  } else {
    in.load4(site, S.v);
  }
  return S;
}
template <int PREC>
__device__ inline Spin<PREC> load_spin(const BAcc<PREC> &in, unsigned site, unsigned a) {
  if constexpr (PREC == P16)
    return {__builtin_amdgcn_raw_buffer_load_b32(in.rs, site * 16u + a * 4u, 0, 0), in.scale(site)};
  else
    return {in.ld(site, a, 1.0f)};
}

template <typename C> struct BLinks {
  __amdgpu_buffer_rsrc_t r;
  __device__ inline BLinks(const void *p, unsigned n) {
    r = __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0,
                                          (int)(n * 2u * (unsigned)sizeof(C)), BUF_DWORD3);
  }
  __device__ inline cx<C> operator()(unsigned i) const {
    if constexpr (sizeof(C) == 8) {
      const double2 v =
          __builtin_bit_cast(double2, __builtin_amdgcn_raw_buffer_load_b128(r, i * 16u, 0, 0));
      return {v.x, v.y};
    } else {
      const float2 v =
          __builtin_bit_cast(float2, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8u, 0, 0));
      return {v.x, v.y};
    }
  }
};

// Links through the scalar cache: when a wave is exactly one s-line (Ls = 64)
// every lane of the wave needs the same links, so the index is made
// wave-uniform and the load is a scalar load (no vector memory instruction);
// the links are read-only for the whole kernel.
typedef const __attribute__((address_space(4))) double ConstF64;
typedef const __attribute__((address_space(4))) float ConstF32;
template <typename C> struct SLinks {
  const void *p;
  __device__ inline SLinks(const void *q, unsigned) : p(q) {}
  __device__ inline cx<C> operator()(unsigned i) const {
    const unsigned u = __builtin_amdgcn_readfirstlane(i);
    if constexpr (sizeof(C) == 8) {
      ConstF64 *q = (ConstF64 *)p;
      return {q[2 * u], q[2 * u + 1]};
    } else {
      ConstF32 *q = (ConstF32 *)p;
      return {q[2 * u], q[2 * u + 1]};
    }
  }
};

// Ls = 32 or 16: a wave holds NL = 2 or 4 s-lines of 64/NL lanes, each with
// its own links: NL scalar loads per link and a per-lane select.  Lines are
// valid or invalid as a whole and a later line never precedes an earlier
// one, so the first active lane belongs to the first line, and a later
// line's index is read only when its first lane is active.
template <typename C> __device__ inline cx<C> slink(const void *p, unsigned u) {
  if constexpr (sizeof(C) == 8) {
    ConstF64 *q = (ConstF64 *)p;
    return {q[2 * u], q[2 * u + 1]};
  } else {
    ConstF32 *q = (ConstF32 *)p;
    return {q[2 * u], q[2 * u + 1]};
  }
}
template <typename C, int NL> struct SLinksN {
  const void *p;
  __device__ inline SLinksN(const void *q, unsigned) : p(q) {}
  __device__ inline cx<C> operator()(unsigned i) const {
    const unsigned long long ex = __builtin_amdgcn_read_exec();
    const unsigned u0 = __builtin_amdgcn_readfirstlane(i);
    if constexpr (NL == 2) {
      const unsigned u1 = ((ex >> 32) & 1) ? (unsigned)__builtin_amdgcn_readlane((int)i, 32) : u0;
      const cx<C> a = slink<C>(p, u0), b = slink<C>(p, u1);
      return (threadIdx.x & 32) ? b : a;
    } else {
      const unsigned u1 = ((ex >> 16) & 1) ? (unsigned)__builtin_amdgcn_readlane((int)i, 16) : u0;
      const unsigned u2 = ((ex >> 32) & 1) ? (unsigned)__builtin_amdgcn_readlane((int)i, 32) : u0;
      const unsigned u3 = ((ex >> 48) & 1) ? (unsigned)__builtin_amdgcn_readlane((int)i, 48) : u0;
      const cx<C> a = slink<C>(p, u0), b = slink<C>(p, u1), c = slink<C>(p, u2),
                  e = slink<C>(p, u3);
      const unsigned line = (threadIdx.x >> 4) & 3;
      const cx<C> lo = (line & 1) ? b : a, hi = (line & 1) ? e : c;
      return (line & 2) ? hi : lo;
    }
  }
};

// value of lane k of the wave (a wave-uniform broadcast)
__device__ inline float lane_value(float v, int k) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ inline double lane_value(double v, int k) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), k);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <typename C> __device__ inline cx<C> lane_value(cx<C> v, int k) {
  return {lane_value(v.r, k), lane_value(v.i, k)};
}

template <typename C> __device__ inline cx<C> link(const void *links, long i) {
  if constexpr (sizeof(C) == 8) {
    const double2 v = ((const double2 *)links)[i];
    return {v.x, v.y};
  } else {
    const float2 v = ((const float2 *)links)[i];
    return {v.x, v.y};
  }
}

// value of v held by lane+1 (NEXT) or lane-1 (PREV) of the wave: DPP wave
// shifts, register-to-register; lanes at the wave edge read 0 (when Ls
// divides 64 they sit at s = Ls-1 / s = 0, where the s hop is absent;
// otherwise the kernel loads their neighbour)
template <bool NEXT> __device__ inline float lane_shift(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), NEXT ? 0x130 : 0x138,
                                                    0xF, 0xF, false));
}
template <bool NEXT> __device__ inline double lane_shift(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), NEXT ? 0x130 : 0x138,
                                             0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), NEXT ? 0x130 : 0x138, 0xF, 0xF,
                                             false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
template <bool NEXT, typename C> __device__ inline cx<C> lane_shift(cx<C> v) {
  return {lane_shift<NEXT>(v.r), lane_shift<NEXT>(v.i)};
}

// one spin of the site held by lane+1 (NEXT) or lane-1 of the wave
template <bool NEXT> __device__ inline unsigned lane_shift_u(unsigned v) {
  return (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, NEXT ? 0x130 : 0x138, 0xF, 0xF, false);
}
template <bool NEXT, int PREC> __device__ inline Spin<PREC> shift_spin(const Site<PREC> &S, int a) {
  if constexpr (PREC == P16)
    return {lane_shift_u<NEXT>(S.h[a]), lane_shift<NEXT>(S.sc)};
  else
    return {lane_shift<NEXT>(S.v[a])};
}
// one spin of the site held by lane k of the wave
template <int PREC> __device__ inline Spin<PREC> lane_spin(const Site<PREC> &S, int a, int k) {
  if constexpr (PREC == P16)
    return {(unsigned)__builtin_amdgcn_readlane((int)S.h[a], k), lane_value(S.sc, k)};
  else
    return {lane_value(S.v[a], k)};
}

// x-neighbour exchange through LDS (one buffer per step parity): FP64/FP32
// sites as re/im values, FP16 sites as their 4 half2 words and scale
template <int PREC, int NW = 4> struct XBuf {
  using C = typename Store<PREC>::C;
  C (*b)[NW][4][64][2];
  __device__ inline void put(unsigned pb, unsigned w, unsigned ln, const Site<PREC> &S) const {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      b[pb][w][a][ln][0] = S.v[a].r;
      b[pb][w][a][ln][1] = S.v[a].i;
    }
  }
  __device__ inline Site<PREC> get(unsigned pb, unsigned w, unsigned ln) const {
    Site<PREC> S;
#pragma unroll
    for (int a = 0; a < 4; a++) S.v[a] = {b[pb][w][a][ln][0], b[pb][w][a][ln][1]};
    return S;
  }
};
template <int NW> struct XBuf<P16, NW> {
  unsigned (*b)[NW][5][64];
  __device__ inline void put(unsigned pb, unsigned w, unsigned ln, const Site<P16> &S) const {
#pragma unroll
    for (int a = 0; a < 4; a++) b[pb][w][a][ln] = S.h[a];
    b[pb][w][4][ln] = __float_as_uint(S.sc);
  }
  __device__ inline Site<P16> get(unsigned pb, unsigned w, unsigned ln) const {
    Site<P16> S;
#pragma unroll
    for (int a = 0; a < 4; a++) S.h[a] = b[pb][w][a][ln];
    S.sc = __uint_as_float(b[pb][w][4][ln]);
    return S;
  }
};

// One thread owns the s-x-y site (s fastest) of one t-chunk and marches
// through it: R(t+1) loaded at step t is the centre at t+1 and the backward
// t neighbour at t+2, so the t hops come from registers.  x and y hops are
// L1/L2 hits (neighbouring lanes and waves of the same and nearby blocks).
// With SHUF (Ls divides 64) the s hops are DPP lane shifts of the centre.
// With UL = 1 (Ls = 64: a wave is one s-line) links are scalar loads and the
// wall couplings s = 0 <-> Ls-1 are lane reads of the centre; with XL the
// inner x neighbours come from the other waves of the block through LDS.
// UL = 2 (Ls = 32: two s-lines per wave) loads both lines' links as scalars.
// Hardware block b runs on XCD b % 8, and each XCD has its own L2, so the
// tiling decides which neighbours are L2 hits:
//   linear  blocks walk the s-x-y slab of a t-chunk in order;
//   remap   XCD k gets the k-th contiguous eighth of the linear order;
//   boxes   XCD k owns one box of a px x py x pt split of (x, y, t) (the split
//           with the smallest halo) and its blocks tile that box's s-x-y
//           cross-section, marching its t-range in chunks: only box faces
//           are fetched across XCDs.
// Partial sums go to partial[rhs * pstride + blockIdx.x]; block 0 zero-fills
// [gridDim.x, pstride) so consumers can always sum pstride entries.
template <int PREC, bool DAG, bool SHUF, int MODE, int UL, bool XL, bool NT, bool YL = false>
__global__ __launch_bounds__(YL ? 2 * BLK : BLK) void k_dslash(const DslashLaunch L,
                                                               const Tiling TL) {
  using C = typename Store<PREC>::C;
  using T = typename Store<PREC>::T;
  // YL (Ls = 64 with the LDS x exchange): a block is 8 waves, 4 consecutive
  // x by 2 consecutive y, wave w at x0 + (w & 3), y0 + (w >> 2)
  constexpr int NTH = YL ? 2 * BLK : BLK;
  static_assert(!YL || (UL == 1 && XL), "the y exchange extends the Ls = 64 x exchange");
  const int rhs = blockIdx.y;
  if (L.done && L.done[(long)rhs * L.done_stride]) return;
  // the previous pass's partial sums of this right-hand side (the finalize
  // kernels' loop and tree, so the value is theirs bit for bit)
  double fs = 0.0;
  if constexpr (MODE == EPI_CG || MODE == EPI_HREC) {
    if (L.fsum) {
      __shared__ double s_fs;
      // the first BLK threads sum as the finalize kernels do (the other
      // waves add exact zeros after their sums)
      const double v = block_sum_n<NTH>(
          threadIdx.x < BLK ? partial_thread(L.fsum + (long)rhs * L.fsum_stride, L.fsum_n) : 0.0);
      if (threadIdx.x == 0) {
        s_fs = v;
        if (blockIdx.x == 0) L.frec[(long)rhs * L.sc_stride] = v;
      }
      __syncthreads();
      fs = s_fs;
    }
  }
  // 32-bit index arithmetic: 4*V5 < 2^31 is checked on the host
  const unsigned lx = L.d.lx, ly = L.d.ly, lt = L.d.lt, ls = L.d.ls;
  const unsigned v2 = lx * ly, v3 = v2 * ls, v4 = v2 * lt, v5 = v4 * ls;
  unsigned j, s, x, y, t0, t1;
  bool valid;
  if (TL.mode == TILE_BOXES) {
    const unsigned k = blockIdx.x & 7, l = blockIdx.x >> 3;
    const unsigned rx = k % TL.px, ry = (k / TL.px) % TL.py, rt = k / (TL.px * TL.py);
    const unsigned chunk = l / TL.nb3;
    t0 = rt * TL.bt + chunk * TL.tchunk;
    t1 = min(t0 + TL.tchunk, rt * TL.bt + TL.bt);
    if constexpr (YL) {
      const unsigned tb = l - chunk * TL.nb3, w = threadIdx.x >> 6;
      valid = chunk * TL.tchunk < TL.bt;
      s = threadIdx.x & 63;
      x = rx * TL.bx + 4 * (tb % TL.ntx) + (w & 3);
      y = ry * TL.by + 2 * (tb / TL.ntx) + (w >> 2);
    } else {
      const unsigned jr = (l - chunk * TL.nb3) * BLK + threadIdx.x;
      valid = jr < ls * TL.bx * TL.by && chunk * TL.tchunk < TL.bt;
      s = jr % ls;
      const unsigned q = jr / ls;
      x = rx * TL.bx + q % TL.bx;
      y = ry * TL.by + q / TL.bx;
    }
    j = s + ls * (x + lx * y);
  } else {
    unsigned b = blockIdx.x;
    if (TL.mode == TILE_REMAP) {
      // blocks b, b+8, b+16, ... (one XCD) -> consecutive logical blocks
      const unsigned nbl = gridDim.x, per = nbl >> 3, full = per << 3;
      if (b < full) b = (b & 7) * per + (b >> 3);
    }
    const unsigned chunk = b / TL.nb3;
    t0 = chunk * TL.tchunk;
    t1 = min(t0 + TL.tchunk, lt);
    if constexpr (YL) {
      const unsigned tb = b - chunk * TL.nb3, w = threadIdx.x >> 6;
      s = threadIdx.x & 63;
      x = 4 * (tb % TL.ntx) + (w & 3);
      y = 2 * (tb / TL.ntx) + (w >> 2);
      j = s + ls * (x + lx * y);
      valid = t0 < lt;
    } else {
      j = (b - chunk * TL.nb3) * BLK + threadIdx.x;
      valid = j < v3 && t0 < lt;
      s = j % ls;
      const unsigned xy = j / ls;
      x = xy % lx;
      y = xy / lx;
    }
  }
  double red = 0.0;
  if (valid) {
    const unsigned xy = x + lx * y;  // 2-D site (link offset within a t-slice)
    const unsigned xf = (x + 1 == lx) ? 0 : x + 1, xb = (x == 0) ? lx - 1 : x - 1;
    const unsigned yf = (y + 1 == ly) ? 0 : y + 1, yb = (y == 0) ? ly - 1 : y - 1;
    // s-x-y offsets of the x and y neighbours, and the 2-D link offsets
    const unsigned jxf = s + ls * (xf + lx * y), jxb = s + ls * (xb + lx * y);
    const unsigned jyf = s + ls * (x + lx * yf), jyb = s + ls * (x + lx * yb);
    const unsigned lxb = xb + lx * y, lyb = x + lx * yb;
    const long roff = (long)rhs * 4 * v5;
    using In = BAcc<PREC>;
    using Lk = std::conditional_t<UL == 1, SLinks<C>,
                                  std::conditional_t<(UL > 1), SLinksN<C, (UL > 1 ? UL : 2)>,
                                                     BLinks<C>>>;
    const In in((const T *)L.in + roff, L.in_norm ? L.in_norm + (long)rhs * v5 : nullptr, v5);
    const Lk lk(L.links, 3 * v4);
    const C diag = L.wilson4d ? (C)(3.0f + (-L.am3)) : (C)((3.0f - L.am3) + 1.0f);
    const int up0 = DAG ? 0 : 2, dn0 = DAG ? 2 : 0;
    const int imass = L.wilson4d ? 0 : L.imass;
    const C m = (C)L.am[rhs];
    const bool shops = !L.wilson4d;
    // epilogue scalars (device-resident solver state)
    C ea = C(0), eb = C(0);
    if constexpr (MODE == EPI_CG) {
      if (L.fsum) {
        const double om = -L.fdelta[(long)rhs * L.sc_stride] / fs;  // omega = -delta / alpha
        ea = (C)(-om);
      } else {
        ea = (C)L.a_ptr[(long)rhs * L.sc_stride];
      }
      if (L.b_ptr) eb = (C)L.b_ptr[(long)rhs * L.sc_stride];
    } else if constexpr (MODE == EPI_HREC) {
      if (L.fsum)
        eb = (C)(fs / L.fdelta[(long)rhs * L.sc_stride]);  // gamma = lambda / delta
      else
        eb = (C)L.b_ptr[(long)rhs * L.sc_stride];
    }

    Site<PREC> prv, cur, nxt;
    const unsigned tb0 = (t0 == 0) ? lt - 1 : t0 - 1;
    prv = load_site(in, j + v3 * tb0);
    cx<C> ubt = conj(lk(2 * v4 + xy + v2 * tb0));
    cur = load_site(in, j + v3 * t0);
    for (unsigned t = t0; t < t1; t++) {
      const unsigned tf = (t + 1 == lt) ? 0 : t + 1;
      nxt = load_site(in, j + v3 * tf);
      const cx<C> uft = lk(2 * v4 + xy + v2 * t);
      const unsigned i = j + v3 * t;
      const unsigned l4 = v2 * t;
      cx<C> o[4];
      centre(o, diag, cur);
      Site<PREC> yf, yb;  // the y neighbours
      if constexpr (XL) {
        // the block's 4 waves are 4 consecutive x at one (y, t) (8 when
        // Ls = 32): the x neighbours inside the block are the centre
        // registers of the next/previous line, exchanged through LDS
        // (double-buffered by step parity, one barrier per step); the outer
        // lines load theirs
        constexpr int NW = YL ? 8 : 4;
        using XB = std::conditional_t<PREC == P16, unsigned[2][NW][5][64], C[2][NW][4][64][2]>;
        __shared__ XB xraw;
        const XBuf<PREC, NW> xbuf{xraw};
        const unsigned w = threadIdx.x >> 6, ln = threadIdx.x & 63, pb = (t - t0) & 1;
        xbuf.put(pb, w, ln, cur);
        __syncthreads();
        Site<PREC> rf, rb;
        if constexpr (UL > 1) {
          // Ls = 64 / UL: the block is 4 UL consecutive x, wave w holding
          // x = UL w + k in lanes [64 k / UL, 64 (k + 1) / UL); position p = UL w + k
          constexpr unsigned NL = UL > 1 ? UL : 2, W = 64 / NL;
          const unsigned p = NL * w + ln / W, sl = ln % W;
          if (p < 4 * NL - 1) {
            rf = xbuf.get(pb, (p + 1) / NL, ((p + 1) % NL) * W + sl);
          } else {
            rf = load_site(in, jxf + v3 * t);
          }
          if (p > 0) {
            rb = xbuf.get(pb, (p - 1) / NL, ((p - 1) % NL) * W + sl);
          } else {
            rb = load_site(in, jxb + v3 * t);
          }
        } else {
          // (with YL the rows are waves 0-3 and 4-7; the y neighbour of
          // either row in the other comes from LDS, the outer one is loaded)
          const unsigned wx = w & 3;
          if (wx < 3) {
            rf = xbuf.get(pb, w + 1, ln);
          } else {
            rf = load_site(in, jxf + v3 * t);
          }
          if (wx > 0) {
            rb = xbuf.get(pb, w - 1, ln);
          } else {
            rb = load_site(in, jxb + v3 * t);
          }
        }
        hop<0, DAG>(o, lk(xy + l4), rf, conj(lk(lxb + l4)), rb);
        if constexpr (YL) {
          // after the x hop (its neighbours are dead): the same LDS buffer
          // holds the other row's centres until the next step's barrier
          yf = w < 4 ? xbuf.get(pb, w + 4, ln) : load_site(in, jyf + v3 * t);
          yb = w >= 4 ? xbuf.get(pb, w - 4, ln) : load_site(in, jyb + v3 * t);
        }
      } else {
        const Site<PREC> rf = load_site(in, jxf + v3 * t), rb = load_site(in, jxb + v3 * t);
        hop<0, DAG>(o, lk(xy + l4), rf, conj(lk(lxb + l4)), rb);
      }
      if constexpr (!YL) {
        yf = load_site(in, jyf + v3 * t);
        yb = load_site(in, jyb + v3 * t);
      }
      hop<1, DAG>(o, lk(v4 + xy + l4), yf, conj(lk(v4 + lyb + l4)), yb);
      hop<2, DAG>(o, uft, nxt, ubt, prv);

      // hops along s (dirac.F90:80-87 for M, 317-324 for M^dagger): spins that
      // hop from s+1 are {3,4} for M and {1,2} for M^dagger; the others from s-1.
      if constexpr (SHUF) {
        Spin<PREC> u0 = shift_spin<true>(cur, up0), u1 = shift_spin<true>(cur, up0 + 1);
        Spin<PREC> d0 = shift_spin<false>(cur, dn0), d1 = shift_spin<false>(cur, dn0 + 1);
        if (UL == 0 && 64 % ls != 0) {
          // s-lines cross wave boundaries: the neighbour of the last lane
          // (s+1) and of the first lane (s-1) sit in the next/previous wave
          const unsigned ln = threadIdx.x & 63;
          if (ln == 63 && s + 1 < ls) {
            u0 = load_spin(in, i + 1, up0);
            u1 = load_spin(in, i + 1, up0 + 1);
          }
          if (ln == 0 && s > 0) {
            d0 = load_spin(in, i - 1, dn0);
            d1 = load_spin(in, i - 1, dn0 + 1);
          }
        }
        if (shops && s + 1 < ls) {
          sub_spin(o[up0], u0);
          sub_spin(o[up0 + 1], u1);
        }
        if (shops && s > 0) {
          sub_spin(o[dn0], d0);
          sub_spin(o[dn0 + 1], d1);
        }
      } else {
        if (shops && s + 1 < ls) {
          sub_spin(o[up0], load_spin(in, i + 1, up0));
          sub_spin(o[up0 + 1], load_spin(in, i + 1, up0 + 1));
        }
        if (shops && s > 0) {
          sub_spin(o[dn0], load_spin(in, i - 1, dn0));
          sub_spin(o[dn0 + 1], load_spin(in, i - 1, dn0 + 1));
        }
      }
      // wall mass (dirac.F90:119-176 for M, 356-402 for M^dagger)
      const unsigned col = i - s;  // s = 0 of this (x,y,t)
      if (imass == 1 || imass == 3) {
        // top: s = Ls couples to s = 1 on the up0 spins; bottom: s = 1 couples to s = Ls
        Spin<PREC> w0[2], wL[2];
        if constexpr (UL == 1) {
          // the wave is the s-line: s = 0 is lane 0, s = Ls-1 is lane 63
#pragma unroll
          for (int k = 0; k < 2; k++) {
            w0[k] = lane_spin(cur, up0 + k, 0);
            wL[k] = lane_spin(cur, dn0 + k, 63);
          }
        }
        if (s == ls - 1) {
#pragma unroll
          for (int k = 0; k < 2; k++) {
            const cx<C> r0 = spin_value(UL == 1 ? w0[k] : load_spin(in, col, up0 + k));
            o[up0 + k] = o[up0 + k] + (imass == 1 ? m * r0 : m * muli(r0));
          }
        }
        if (s == 0) {
#pragma unroll
          for (int k = 0; k < 2; k++) {
            const cx<C> rL = spin_value(UL == 1 ? wL[k] : load_spin(in, col + ls - 1, dn0 + k));
            o[dn0 + k] = o[dn0 + k] + (imass == 1 ? m * rL : (-m) * muli(rL));
          }
        }
      } else if (imass == 5) {
        // same-s chirality flip at both walls: +i m (M), -i m (M^dagger)
        const C sg = DAG ? C(-1) : C(1);
        if (s == ls - 1) {
#pragma unroll
          for (int k = 0; k < 2; k++)
            o[up0 + k] = o[up0 + k] + (sg * m) * muli(spin_value(spin_of(cur, dn0 + k)));
        }
        if (s == 0) {
#pragma unroll
          for (int k = 0; k < 2; k++)
            o[dn0 + k] = o[dn0 + k] + (sg * m) * muli(spin_value(spin_of(cur, up0 + k)));
        }
      }

      // epilogue
      if constexpr (MODE == EPI_CG) {
        const C a = ea;
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
        const BOut<PREC, NT> rout(ro, L.r_out_norm ? L.r_out_norm + (long)rhs * v5 : nullptr, v5);
        rout.store4(i, rv);
        if (L.b_ptr) {
          const C bb = eb;
          Acc<PREC> pp{(const T *)L.p + roff, nullptr, v5};
          Acc<PREC> xx{(const T *)L.x + roff, nullptr, v5};
          cx<C> pv[4], xv[4];
          pp.load4(i, pv);
          xx.load4(i, xv);
#pragma unroll
          for (int k = 0; k < 4; k++) xv[k] = xv[k] + bb * pv[k];
          Out<PREC> xo{(T *)L.x + roff, nullptr, v5};
          xo.store4(i, xv);
        }
      } else {
        const BOut<PREC, NT> out((T *)L.out + roff, L.out_norm ? L.out_norm + (long)rhs * v5 : nullptr,
                                 v5);
        if constexpr (MODE == EPI_HREC) {
          // out = M in + b out: this thread reads its own site before writing it
          const C bb = eb;
          Acc<PREC> ho{(const T *)L.out + roff, L.out_norm ? L.out_norm + (long)rhs * v5 : nullptr,
                       v5};
          cx<C> hv[4];
          ho.load4(i, hv);
#pragma unroll
          for (int k = 0; k < 4; k++) {
            o[k] = o[k] + bb * hv[k];
            red += (double)norm2(o[k]);
          }
        }
        out.store4(i, o);
        if constexpr (MODE == EPI_NORM) {
#pragma unroll
          for (int k = 0; k < 4; k++) red += (double)norm2(o[k]);
        } else if constexpr (MODE == EPI_DOTW) {
          Acc<PREC> w{(const T *)L.w + roff, nullptr, v5};
          cx<C> wv[4];
          w.load4(i, wv);
#pragma unroll
          for (int k = 0; k < 4; k++) red += (double)(wv[k].r * o[k].r + wv[k].i * o[k].i);
        }
      }
      prv = cur;
      cur = nxt;
      ubt = conj(uft);
    }
  }
  if constexpr (MODE != EPI_STORE) {
    const double tot = block_sum_n<NTH>(red);
    double *P = L.partial + (long)rhs * TL.pstride;
    if (threadIdx.x == 0) P[blockIdx.x] = tot;
    if (blockIdx.x == 0)
      for (unsigned k = gridDim.x + threadIdx.x; k < TL.pstride; k += NTH) P[k] = 0.0;
  }
}

}  // namespace

// bound on every Dslash grid: t-chunk 1 with ceil(Lx Ly Ls / BLK) blocks per
// t-slice, plus up to 8 partial blocks per slice from the XCD boxes; also
// covers V5 sites when used as a one-site-per-thread grid
long dslash_blocks(const Dims &d) {
  return (((long)d.lx * d.ly * d.ls + BLK - 1) / BLK + 8) * d.lt;
}

// px x py x pt = 8 split of (x, y, t) with the smallest halo fraction
// 2/(L/p) summed over split directions; false if no split divides evenly
static bool box_split(const Dims &d, unsigned &px, unsigned &py, unsigned &pt) {
  double best = 1e30;
  bool found = false;
  const unsigned ps[4] = {1, 2, 4, 8};
  for (unsigned a : ps)
    for (unsigned b : ps) {
      if (a * b > 8 || 8 % (a * b)) continue;
      const unsigned c = 8 / (a * b);
      if (d.lx % a || d.ly % b || d.lt % c) continue;
      const double cost = (a > 1 ? 2.0 * a / d.lx : 0.0) + (b > 1 ? 2.0 * b / d.ly : 0.0) +
                          (c > 1 ? 2.0 * c / d.lt : 0.0);
      if (cost < best - 1e-12) {
        best = cost;
        px = a;
        py = b;
        pt = c;
        found = true;
      }
    }
  return found;
}

// ---------------------------------------------------------------------------
// launch configuration per (dims, precision): t-chunk length, XCD remap or
// XCD boxes, uniform (scalar) links, lane-shift s hops.  The default is fixed (results
// identical run to run); dslash_autotune times the candidates on scratch
// fields and pins the fastest for the process.  THR_DSLASH_TCHUNK / _REMAP /
// _UNIFORM / _LANESHIFT / _BOXES / _LDSX / _NT override the default.
// ---------------------------------------------------------------------------
namespace {
// shapes are pinned per lattice, precision and batch class (one right-hand
// side, or several: a batch's larger grid favours other shapes)
struct TuneKey {
  int lx, ly, lt, ls, prec, multi;
  bool operator<(const TuneKey &o) const {
    return std::tie(lx, ly, lt, ls, prec, multi) <
           std::tie(o.lx, o.ly, o.lt, o.ls, o.prec, o.multi);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, DslashConfig> g_tuned;

DslashConfig default_config(const Dims &d) {
  DslashConfig c;
  // t-chunk 8 (the measured best at 32^3 x 64) unless the lattice is too
  // small for it to give ~512 blocks per right-hand side: a thread's t
  // steps run one after another, so a small lattice marched in long chunks
  // is a few blocks waiting on memory latency (8^3 x 8 at chunk 8: 2 blocks)
  {
    const long nb3 = ((long)d.lx * d.ly * d.ls + BLK - 1) / BLK;
    const long chunks = std::max(1L, (512 + nb3 - 1) / nb3);
    c.tchunk = (int)std::max(1L, std::min<long>(std::min(8, d.lt), d.lt / chunks));
  }
  // lane shifts pay when every wave holds whole s-lines; when lines cross
  // waves the edge-lane loads cost as much as they save (measured at Ls = 48)
  c.laneshift = (64 % d.ls) == 0 ? 1 : 0;
  c.uniform_links = (d.ls == 64 || d.ls == 32) ? 1 : 0;
  c.lds_x = c.uniform_links;
  if (const char *e = std::getenv("THR_DSLASH_TCHUNK")) c.tchunk = std::max(1, std::atoi(e));
  if (const char *e = std::getenv("THR_DSLASH_REMAP")) c.remap = std::atoi(e) ? 1 : 0;
  if (const char *e = std::getenv("THR_DSLASH_UNIFORM")) c.uniform_links = std::atoi(e) ? 1 : 0;
  if (const char *e = std::getenv("THR_DSLASH_LANESHIFT")) c.laneshift = std::atoi(e) ? 1 : 0;
  if (const char *e = std::getenv("THR_DSLASH_BOXES")) c.box = std::atoi(e) ? 1 : 0;
  if (const char *e = std::getenv("THR_DSLASH_LDSX")) c.lds_x = std::atoi(e) ? 1 : 0;
  if (const char *e = std::getenv("THR_DSLASH_LDSY")) c.lds_y = std::atoi(e) ? 1 : 0;
  if (const char *e = std::getenv("THR_DSLASH_NT")) c.nt_stores = std::atoi(e) ? 1 : 0;
  if (const char *e = std::getenv("THR_DSLASH_OCC")) c.occ_cap = std::max(0, std::min(3, std::atoi(e)));
  return c;
}
}  // namespace

DslashConfig dslash_config(const Dims &d, int prec, int nrhs) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  // a batch without a shape of its own takes the single-RHS one
  for (int m = nrhs > 1 ? 1 : 0; m >= 0; m--) {
    auto it = g_tuned.find(TuneKey{d.lx, d.ly, d.lt, d.ls, prec, m});
    if (it != g_tuned.end()) return it->second;
  }
  return default_config(d);
}

void dslash_set_config(const Dims &d, int prec, DslashConfig c, int which) {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  for (int m = 0; m < 2; m++)
    if (which < 0 || which == m) g_tuned[TuneKey{d.lx, d.ly, d.lt, d.ls, prec, m}] = c;
}

// block -> lattice mapping of a launch with configuration c; *nblocks
// receives the grid's x extent (blocks per right-hand side)
static Tiling make_tiling(const Dims &d, const DslashConfig &c, unsigned *nblocks) {
  Tiling TL{};
  TL.pstride = (unsigned)dslash_blocks(d);
  // 4 x 2-line blocks: the Ls = 64 LDS x exchange (scalar links, lane shifts)
  // on a slab or box of whole tiles
  const bool ylc = c.lds_y && c.lds_x && c.uniform_links && c.laneshift && d.ls == 64;
  unsigned px = 1, py = 1, pt = 8;
  if (c.box && box_split(d, px, py, pt)) {
    TL.mode = TILE_BOXES;
    TL.px = px;
    TL.py = py;
    TL.bx = d.lx / px;
    TL.by = d.ly / py;
    TL.bt = d.lt / pt;
    TL.tchunk = (unsigned)std::max(1, std::min<int>(c.tchunk, (int)TL.bt));
    TL.yl = ylc && TL.bx % 4 == 0 && TL.by % 2 == 0;
    TL.ntx = TL.bx / 4;
    TL.nb3 = TL.yl ? TL.ntx * (TL.by / 2) : ((unsigned)d.ls * TL.bx * TL.by + BLK - 1) / BLK;
    *nblocks = 8 * TL.nb3 * ((TL.bt + TL.tchunk - 1) / TL.tchunk);
  } else {
    TL.mode = c.remap ? TILE_REMAP : TILE_LINEAR;
    TL.tchunk = (unsigned)std::max(1, std::min(c.tchunk, d.lt));
    TL.yl = ylc && d.lx % 4 == 0 && d.ly % 2 == 0;
    TL.ntx = d.lx / 4;
    TL.nb3 = TL.yl ? TL.ntx * (d.ly / 2) : ((unsigned)d.lx * d.ly * d.ls + BLK - 1) / BLK;
    *nblocks = TL.nb3 * ((d.lt + TL.tchunk - 1) / TL.tchunk);
  }
  return TL;
}

long dslash_grid_blocks(const Dims &d, const DslashConfig &c) {
  unsigned nb = 0;
  (void)make_tiling(d, c, &nb);
  return nb;
}

// Dynamic LDS that caps a kernel's resident blocks per CU at cap + 1
// (occupancy cap of a launch shape): the padding is launch-time only, so
// results do not change.  The kernel's static LDS and the device's LDS per CU
// and per block are queried once.
static std::mutex g_pad_mu;
static std::map<const void *, size_t> g_static_lds;
static size_t lds_pad(const void *kern, int cap) {
  if (cap <= 0) return 0;
  static int lds_cu = -1, lds_blk = -1;
  std::lock_guard<std::mutex> lk(g_pad_mu);
  if (lds_cu < 0) {
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) !=
            hipSuccess ||
        hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) !=
            hipSuccess) {
      lds_cu = lds_blk = 0;
    }
  }
  if (lds_cu <= 0 || lds_blk <= 0) return 0;
  auto it = g_static_lds.find(kern);
  if (it == g_static_lds.end()) {
    hipFuncAttributes a{};
    const size_t st = hipFuncGetAttributes(&a, kern) == hipSuccess ? a.sharedSizeBytes : 0;
    it = g_static_lds.emplace(kern, st).first;
  }
  const size_t blocks = (size_t)cap + 1;
  size_t target = ((size_t)lds_cu / blocks) & ~(size_t)511;
  target = std::min(target, (size_t)lds_blk);
  // a cap that the per-block limit cannot enforce is not applied
  if ((blocks + 1) * target <= (size_t)lds_cu || target <= it->second) return 0;
  return target - it->second;
}

int launch_dslash_cfg(const DslashLaunch &L, DslashConfig c, hipStream_t s) {
  THR_CHECK(L.nrhs >= 1 && L.nrhs <= MAXRHS, THR_ERR_ARG, "nrhs out of range");
  THR_CHECK(L.d.lx > 0 && L.d.ly > 0 && L.d.lt > 0 && L.d.ls > 0, THR_ERR_ARG, "bad dims");
  THR_CHECK(L.imass == 1 || L.imass == 3 || L.imass == 5, THR_ERR_ARG, "imass must be 1, 3 or 5");
  THR_CHECK(L.mode == EPI_STORE || L.partial, THR_ERR_ARG, "partial buffer required");
  THR_CHECK(L.prec != P16 || L.mode != EPI_CG || (L.r_norm && !L.b_ptr &&
                                                   (L.r_out ? L.r_out_norm != nullptr : true)),
            THR_ERR_ARG, "fp16 CG epilogue needs r scales and no x update");
  THR_CHECK(L.prec != P16 || L.mode != EPI_DOTW, THR_ERR_ARG, "fp16 dot epilogue unsupported");
  THR_CHECK(L.mode != EPI_HREC || (L.b_ptr && (L.prec != P16 || L.out_norm)), THR_ERR_ARG,
            "recurrence epilogue needs its scalar (and fp16 scales)");
  THR_CHECK(4 * L.d.v5() < (1L << 31), THR_ERR_SIZE, "lattice too large for 32-bit indexing");
  unsigned nblocks = 0;
  const Tiling TL = make_tiling(L.d, c, &nblocks);
  THR_CHECK(nblocks <= TL.pstride, THR_ERR_ARG, "dslash grid exceeds partial stride");
  dim3 grid(nblocks, (unsigned)L.nrhs);
  dim3 block(TL.yl ? 2 * BLK : BLK);
  const bool shuf = c.laneshift != 0;
  // a wave is exactly one s-line when Ls = 64, two when Ls = 32, four when
  // Ls = 16 (blocks are whole waves and every tiling starts a wave at s = 0)
  const bool ul = c.uniform_links && shuf && L.d.ls == 64;
  const bool ul2 = c.uniform_links && shuf && L.d.ls == 32;
  const bool ul4 = c.uniform_links && shuf && L.d.ls == 16;
  // LDS x exchange: every block must be 4 full waves at 4 (Ls = 64) or 8
  // (Ls = 32) consecutive x of one (y, t) (all waves then take the same
  // barriers)
  const bool xl = ul && c.lds_x && L.d.lx % 4 == 0 &&
                  (TL.mode != TILE_BOXES || TL.bx % 4 == 0);
  const bool xl2 = ul2 && c.lds_x && L.d.lx % 8 == 0 &&
                   (TL.mode != TILE_BOXES || TL.bx % 8 == 0);
  const bool xl4 = ul4 && c.lds_x && L.d.lx % 16 == 0 &&
                   (TL.mode != TILE_BOXES || TL.bx % 16 == 0);
  const bool nt = c.nt_stores != 0;
  THR_CHECK(!TL.yl || xl, THR_ERR_STATE, "y exchange without the x exchange");
  THR_CHECK(4 * L.d.v5() * Store<P64>::bytes < (1L << 32), THR_ERR_SIZE,
            "spinor too large for 32-bit buffer offsets");
#define THR_GO(K)                                                                       \
  do {                                                                                  \
    const void *kf = (const void *)(K);                                                 \
    hipLaunchKernelGGL(K, grid, block, lds_pad(kf, c.occ_cap), s, L, TL);               \
  } while (0)
#define THR_LAUNCH_NT(P, D, S, M, NTV)                                 \
  do {                                                                 \
    if (TL.yl)                                                         \
      THR_GO((k_dslash<P, D, true, M, 1, true, NTV, true>));           \
    else if (xl)                                                       \
      THR_GO((k_dslash<P, D, true, M, 1, true, NTV>));                 \
    else if (ul)                                                       \
      THR_GO((k_dslash<P, D, true, M, 1, false, NTV>));                \
    else if (xl2)                                                      \
      THR_GO((k_dslash<P, D, true, M, 2, true, NTV>));                 \
    else if (ul2)                                                      \
      THR_GO((k_dslash<P, D, true, M, 2, false, NTV>));                \
    else if (xl4)                                                      \
      THR_GO((k_dslash<P, D, true, M, 4, true, NTV>));                 \
    else if (ul4)                                                      \
      THR_GO((k_dslash<P, D, true, M, 4, false, NTV>));                \
    else                                                               \
      THR_GO((k_dslash<P, D, S, M, 0, false, NTV>));                   \
  } while (0)
#define THR_LAUNCH(P, D, S, M)                 \
  do {                                         \
    if (nt)                                    \
      THR_LAUNCH_NT(P, D, S, M, true);         \
    else                                       \
      THR_LAUNCH_NT(P, D, S, M, false);        \
  } while (0)
#define THR_LAUNCH1(P, D, S)                        \
  switch (L.mode) {                                 \
    case EPI_STORE: THR_LAUNCH(P, D, S, EPI_STORE); break; \
    case EPI_NORM: THR_LAUNCH(P, D, S, EPI_NORM); break;   \
    case EPI_DOTW: THR_LAUNCH(P, D, S, EPI_DOTW); break;   \
    case EPI_HREC: THR_LAUNCH(P, D, S, EPI_HREC); break;   \
    default: THR_LAUNCH(P, D, S, EPI_CG); break;           \
  }
#define THR_LAUNCH2(P, D) \
  if (shuf) { THR_LAUNCH1(P, D, true); } else { THR_LAUNCH1(P, D, false); }
  if (L.prec == P64) {
    if (L.dagger) { THR_LAUNCH2(P64, true); } else { THR_LAUNCH2(P64, false); }
  } else if (L.prec == P32) {
    if (L.dagger) { THR_LAUNCH2(P32, true); } else { THR_LAUNCH2(P32, false); }
  } else {
    if (L.dagger) { THR_LAUNCH2(P16, true); } else { THR_LAUNCH2(P16, false); }
  }
#undef THR_LAUNCH2
#undef THR_LAUNCH1
#undef THR_LAUNCH
#undef THR_LAUNCH_NT
#undef THR_GO
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int launch_dslash(const DslashLaunch &L, hipStream_t s) {
  return launch_dslash_cfg(L, dslash_config(L.d, L.prec, L.nrhs), s);
}

int dslash_autotune(const Dims &d, int prec, const void *links, DslashConfig *best,
                    hipStream_t s, int nrhs) {
  THR_CHECK(links, THR_ERR_ARG, "links required");
  THR_CHECK(nrhs >= 1 && nrhs <= MAXRHS, THR_ERR_ARG, "nrhs out of range");
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  THR_HIP(hipStreamIsCapturing(s, &cs));
  THR_CHECK(cs == hipStreamCaptureStatusNone, THR_ERR_STATE, "cannot autotune while capturing");
  const size_t es = prec == P64 ? 16 : prec == P32 ? 8 : 4;
  const size_t nsp = (size_t)4 * d.v5() * nrhs;
  const size_t nsite = (size_t)d.v5() * nrhs;
  // rotate over enough input/output pairs to exceed the 256 MB Infinity
  // Cache, so every candidate is timed on cold fields as in a solver
  const size_t pair = 2 * (es * nsp + sizeof(float) * nsite);
  const int nbuf = (int)std::min<size_t>(8, std::max<size_t>(2, (768u << 20) / pair + 1));
  std::vector<void *> in(nbuf, nullptr), out(nbuf, nullptr);
  std::vector<float *> inn(nbuf, nullptr), outn(nbuf, nullptr);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = THR_OK;
  DslashConfig pick = default_config(d);
  auto cleanup = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    for (int k = 0; k < nbuf; k++) {
      (void)hipFree(in[k]);
      (void)hipFree(out[k]);
      (void)hipFree(inn[k]);
      (void)hipFree(outn[k]);
    }
  };
  bool ok = hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
  for (int k = 0; k < nbuf && ok; k++)
    ok = hipMalloc(&in[k], es * nsp) == hipSuccess && hipMalloc(&out[k], es * nsp) == hipSuccess &&
         hipMalloc((void **)&inn[k], sizeof(float) * nsite) == hipSuccess &&
         hipMalloc((void **)&outn[k], sizeof(float) * nsite) == hipSuccess;
  if (!ok) {
    cleanup();
    set_error("dslash_autotune: allocation failed");
    return THR_ERR_HIP;
  }
  for (int k = 0; k < nbuf; k++) {
    (void)hipMemsetAsync(in[k], 0, es * nsp, s);
    (void)hipMemsetAsync(inn[k], 0, sizeof(float) * nsite, s);
  }
  DslashLaunch L{};
  L.d = d;
  L.nrhs = nrhs;
  L.prec = prec;
  L.imass = 3;
  L.am3 = 1.0f;
  for (int k = 0; k < nrhs; k++) L.am[k] = 0.01f;
  L.links = links;
  L.mode = EPI_STORE;
  L.sc_stride = 1;
  L.done_stride = 1;
  int turn = 0;
  auto launch = [&](const DslashConfig &c) {
    const int k = turn++ % nbuf;
    L.in = in[k];
    L.in_norm = inn[k];
    L.out = out[k];
    L.out_norm = outn[k];
    L.dagger = turn & 1;
    return launch_dslash_cfg(L, c, s);
  };
  struct Cand {
    float ms;
    DslashConfig c;
  };
  std::vector<Cand> first;
  // time `reps` launches of shape c on the rotating scratch fields
  auto timed = [&](const DslashConfig &c, int reps, float *ms) {
    int r = THR_OK;
    (void)hipEventRecord(e0, s);
    for (int k = 0; k < reps && !r; k++) r = launch(c);
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) return (int)THR_ERR_HIP;
    if (!r && hipEventElapsedTime(ms, e0, e1) != hipSuccess) r = THR_ERR_HIP;
    return r;
  };
  // t-chunks: powers of two and the intermediate lengths that split Lt into
  // 3, 5, 6 or 12 chunks (so the grid fills the GPU in whole rounds of
  // resident blocks at more sizes)
  std::vector<int> tcs = {1, 2, 4, 8, 16};
  for (int k : {3, 5, 6, 12}) tcs.push_back((d.lt + k - 1) / k);
  std::sort(tcs.begin(), tcs.end());
  tcs.erase(std::unique(tcs.begin(), tcs.end()), tcs.end());
  for (int tc : tcs) {
    if (tc > d.lt || tc < 1) continue;
    for (int v = 0; v < 128 && !rc; v++) {
      if ((v & 2) && (!(v & 4) || (d.ls != 64 && d.ls != 32 && d.ls != 16)))
        continue;  // uniform links need Ls = 64, 32 or 16
      if ((v & 16) && !(v & 2)) continue;  // the LDS x exchange needs uniform links
      if ((v & 8) && (v & 1)) continue;            // boxes replace the remap
      // the y exchange extends the Ls = 64 x exchange (and must apply)
      if ((v & 64) && (!(v & 16) || d.ls != 64 || d.lx % 4 || d.ly % 2)) continue;
      DslashConfig c;
      c.tchunk = tc;
      c.remap = v & 1;
      c.uniform_links = (v >> 1) & 1;
      c.laneshift = (v >> 2) & 1;
      c.box = (v >> 3) & 1;
      c.lds_x = (v >> 4) & 1;
      c.nt_stores = (v >> 5) & 1;
      c.lds_y = (v >> 6) & 1;
      if (c.lds_y) {
        unsigned nbq = 0;
        if (!make_tiling(d, c, &nbq).yl) continue;  // boxes whose extents do not tile
      }
      for (int k = 0; k < 2 && !rc; k++) rc = launch(c);  // warm
      if (rc) break;
      float ms = 0.f;
      rc = timed(c, 2 * nbuf, &ms);
      if (!rc) first.push_back({ms, c});
    }
  }
  // the short first pass ranks the shapes; the best few are timed again in
  // interleaved rounds (clock and neighbour noise then hits them alike)
  if (!rc && !first.empty()) {
    std::sort(first.begin(), first.end(),
              [](const Cand &a, const Cand &b) { return a.ms < b.ms; });
    // the best six, each also with its resident blocks capped at 2, 3, 4
    // per CU (more resident waves hide latency but share each XCD's L2),
    // timed in ten interleaved rounds of at least 16 launches each
    std::vector<DslashConfig> fin;
    for (size_t k = 0; k < std::min<size_t>(6, first.size()); k++)
      for (int cap = 0; cap <= 3; cap++) {
        DslashConfig c = first[k].c;
        c.occ_cap = cap;
        fin.push_back(c);
      }
    const size_t nk = fin.size();
    std::vector<float> tot(nk, 0.f);
    const int reps = std::max(16, 2 * nbuf);
    for (int round = 0; round < 10 && !rc; round++)
      for (size_t k = 0; k < nk && !rc; k++) {
        float ms = 0.f;
        rc = timed(fin[k], reps, &ms);
        tot[k] += ms;
      }
    size_t kb = 0;
    for (size_t k = 1; k < nk; k++)
      if (tot[k] < tot[kb]) kb = k;
    pick = fin[kb];
  }
  cleanup();
  if (rc) return rc;
  dslash_set_config(d, prec, pick, nrhs > 1 ? 1 : 0);
  if (best) *best = pick;
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
