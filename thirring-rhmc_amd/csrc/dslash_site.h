// Per-site pieces of the domain-wall operator shared by the streaming
// Dslash kernels and the persistent solver: a 5-D site's four spins, the
// gamma-matrix tables (gammamatrices.F90:27-75) and the spin-projected 4-D
// hop of dirac.F90:49-117 / 330-355.
#pragma once
#include "kernels.h"

namespace thr {
namespace {

// A 5-D site's four spins as the kernel carries them: FP64 and FP32 as
// complex values; FP16 as the stored half2 words with the site's scale,
// expanded only inside the arithmetic, where v_fma_mix_f32 reads a half
// operand directly (no separate conversion or scaling instructions).
template <int PREC> struct Site {
  cx<typename Store<PREC>::C> v[4];
};
template <> struct Site<P16> {
  unsigned h[4];
  float sc;
};
// one spin of a site (s hops, wall terms)
template <int PREC> struct Spin {
  cx<typename Store<PREC>::C> v;
};
template <> struct Spin<P16> {
  unsigned h;
  float sc;
};

__device__ inline float lo16(unsigned w) { return __low2float(__builtin_bit_cast(__half2, w)); }
__device__ inline float hi16(unsigned w) { return __high2float(__builtin_bit_cast(__half2, w)); }
// sc * (half2 word), the conversion folded into the multiply (fma with a zero
// addend: hipcc then emits v_fma_mix_f32 with the half operand)
__device__ inline cx<float> expand(unsigned w, float sc) {
  return {fmaf(sc, lo16(w), 0.0f), fmaf(sc, hi16(w), 0.0f)};
}

template <int PREC> __device__ inline cx<typename Store<PREC>::C> spin_value(const Spin<PREC> &p) {
  if constexpr (PREC == P16)
    return expand(p.h, p.sc);
  else
    return p.v;
}
template <int PREC> __device__ inline Spin<PREC> spin_of(const Site<PREC> &S, int a) {
  if constexpr (PREC == P16)
    return {S.h[a], S.sc};
  else
    return {S.v[a]};
}
// o -= value of p
template <int PREC, typename C> __device__ inline void sub_spin(cx<C> &o, const Spin<PREC> &p) {
  if constexpr (PREC == P16) {
    o.r = fmaf(-p.sc, lo16(p.h), o.r);
    o.i = fmaf(-p.sc, hi16(p.h), o.i);
  } else {
    o = o - p.v;
  }
}

// gamma_mu (gammamatrices.F90:27-75, without kappa): (gamma_mu v)[a] =
// e(mu, a) v[p(mu, a)], p an involution pairing {0, 1} with {2, 3} and
// e(mu, a) e(mu, p(a)) = 1 (gamma_mu^2 = 1).  Phases: 0 -> +1, 1 -> -1,
// 2 -> +i, 3 -> -i.
__device__ constexpr int kGamP[3][4] = {{3, 2, 1, 0}, {3, 2, 1, 0}, {2, 3, 0, 1}};
__device__ constexpr int kGamE[3][4] = {{3, 3, 2, 2}, {1, 0, 0, 1}, {3, 2, 2, 3}};

template <int E, typename C> __device__ inline cx<C> phase(cx<C> v) {
  if constexpr (E == 0) return v;
  else if constexpr (E == 1) return {-v.r, -v.i};
  else if constexpr (E == 2) return {-v.i, v.r};
  else return {v.i, -v.r};
}

// One side of a 4-D hop: o -= 1/2 (1 + SIG gamma_mu) U v.  (1 + SIG gamma_mu)
// has rank 2: its upper components are c_a = v_a + SIG e_a v_p(a) (a = 0, 1)
// and its lower ones SIG e_b c_p(b) (b = 2, 3), so the link multiplies the two
// components of c only (Wilson spin projection).  For FP16 sites c is formed
// from the half words with the site scale folded in.
template <int MU, int SIG, int PREC, typename C>
__device__ inline void half_hop(cx<C> o[4], cx<C> u, const Site<PREC> &S) {
  constexpr int p0 = kGamP[MU][0], p1 = kGamP[MU][1];
  cx<C> c0, c1;
  if constexpr (PREC == P16) {
    const float sp = S.sc, sq = SIG > 0 ? S.sc : -S.sc;
    const cx<float> q0 = phase<kGamE[MU][0]>(cx<float>{lo16(S.h[p0]), hi16(S.h[p0])});
    const cx<float> q1 = phase<kGamE[MU][1]>(cx<float>{lo16(S.h[p1]), hi16(S.h[p1])});
    c0 = {fmaf(sp, lo16(S.h[0]), fmaf(sq, q0.r, 0.0f)), fmaf(sp, hi16(S.h[0]), fmaf(sq, q0.i, 0.0f))};
    c1 = {fmaf(sp, lo16(S.h[1]), fmaf(sq, q1.r, 0.0f)), fmaf(sp, hi16(S.h[1]), fmaf(sq, q1.i, 0.0f))};
  } else {
    const cx<C> q0 = phase<kGamE[MU][0]>(S.v[p0]), q1 = phase<kGamE[MU][1]>(S.v[p1]);
    c0 = SIG > 0 ? S.v[0] + q0 : S.v[0] - q0;
    c1 = SIG > 0 ? S.v[1] + q1 : S.v[1] - q1;
  }
  const cx<C> w0 = cmul(u, c0), w1 = cmul(u, c1);
  const C h = C(0.5), hs = SIG > 0 ? C(0.5) : C(-0.5);
  o[0] = o[0] - h * w0;
  o[1] = o[1] - h * w1;
  // lower components: b = p(a) for a = 0, 1
  o[p0] = o[p0] - hs * phase<kGamE[MU][p0]>(w0);
  o[p1] = o[p1] - hs * phase<kGamE[MU][p1]>(w1);
}

// 4-D hop in direction mu, dirac.F90:49-117 (M) and 330-355 (M^dagger):
//   M:   o -= 1/2 (1 - gamma_mu) U_mu(x) R(x+mu) + 1/2 (1 + gamma_mu) U_mu(x-mu)^* R(x-mu)
//   M^+: the projector signs swapped
// uf = U_mu(x), rf = R(x+mu), ub = U_mu(x-mu)^*, rb = R(x-mu)
template <int MU, bool DAG, int PREC, typename C>
__device__ inline void hop(cx<C> o[4], cx<C> uf, const Site<PREC> &rf, cx<C> ub,
                           const Site<PREC> &rb) {
  half_hop<MU, DAG ? 1 : -1>(o, uf, rf);
  half_hop<MU, DAG ? -1 : 1>(o, ub, rb);
}

// o = diag * S
template <int PREC, typename C>
__device__ inline void centre(cx<C> o[4], C diag, const Site<PREC> &S) {
  if constexpr (PREC == P16) {
    const float d = diag * S.sc;
#pragma unroll
    for (int a = 0; a < 4; a++) o[a] = expand(S.h[a], d);
  } else {
#pragma unroll
    for (int a = 0; a < 4; a++) o[a] = diag * S.v[a];
  }
}

}  // namespace
}  // namespace thr
