// Shared device/host definitions for the gfx950 domain-wall library.
//
// Device layout (halo-free, everything resident in HBM):
//   spinor   [rhs][spin 0..3][t][y][x][s]   complex, s fastest
//   links    [mu 0..2][t][y][x]             complex
//   fp16 spinors store half2 per complex plus one float scale per 5-D site
//   [rhs][t][y][x][s] shared by the 4 spins.
// This is the reference's column-major Phi(1:Ls,1:Lx,1:Ly,1:Lt,1:4) with the
// halos stripped: periodic and antiperiodic boundaries are index arithmetic
// on the device (the antiperiodic t sign lives in the links, as in coef).
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdio>
#include <string>

namespace thr {

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
void set_error(const std::string &msg);
const char *last_error();

#define THR_HIP(call)                                                                   \
  do {                                                                                  \
    hipError_t _e = (call);                                                             \
    if (_e != hipSuccess) {                                                             \
      ::thr::set_error(std::string(#call) + ": " + hipGetErrorString(_e));             \
      return THR_ERR_HIP;                                                               \
    }                                                                                   \
  } while (0)

#define TRY_RC(x)            \
  do {                       \
    int _rc = (x);           \
    if (_rc) return _rc;     \
  } while (0)

#define THR_CHECK(cond, code, msg) \
  do {                             \
    if (!(cond)) {                 \
      ::thr::set_error(msg);       \
      return (code);               \
    }                              \
  } while (0)

// ---------------------------------------------------------------------------
// precision tags
// ---------------------------------------------------------------------------
enum Prec { P64 = 0, P32 = 1, P16 = 2 };

struct Dims {
  int lx, ly, lt, ls;
  __host__ __device__ long v4() const { return (long)lx * ly * lt; }
  __host__ __device__ long v5() const { return v4() * ls; }
  __host__ __device__ long nspinor() const { return v5() * 4; }  // complex elements per spinor
};

// ---------------------------------------------------------------------------
// complex helpers (compute type C = double or float)
// ---------------------------------------------------------------------------
template <typename C> struct cx { C r, i; };

template <typename C> __host__ __device__ inline cx<C> mk(C r, C i) { return cx<C>{r, i}; }
template <typename C> __device__ inline cx<C> operator+(cx<C> a, cx<C> b) { return {a.r + b.r, a.i + b.i}; }
template <typename C> __device__ inline cx<C> operator-(cx<C> a, cx<C> b) { return {a.r - b.r, a.i - b.i}; }
template <typename C> __device__ inline cx<C> operator*(C s, cx<C> a) { return {s * a.r, s * a.i}; }
template <typename C> __device__ inline cx<C> cmul(cx<C> a, cx<C> b) {
  return {a.r * b.r - a.i * b.i, a.r * b.i + a.i * b.r};
}
// conj(a) * b
template <typename C> __device__ inline cx<C> cjmul(cx<C> a, cx<C> b) {
  return {a.r * b.r + a.i * b.i, a.r * b.i - a.i * b.r};
}
template <typename C> __device__ inline cx<C> conj(cx<C> a) { return {a.r, -a.i}; }
// i * a
template <typename C> __device__ inline cx<C> muli(cx<C> a) { return {-a.i, a.r}; }
template <typename C> __device__ inline C norm2(cx<C> a) { return a.r * a.r + a.i * a.i; }

// ---------------------------------------------------------------------------
// storage access: load/store a complex of compute type C from storage of Prec
// ---------------------------------------------------------------------------
template <int PREC> struct Store;

template <> struct Store<P64> {
  using C = double;
  using T = double2;
  static constexpr int bytes = 16;
};
template <> struct Store<P32> {
  using C = float;
  using T = float2;
  static constexpr int bytes = 8;
};
template <> struct Store<P16> {
  using C = float;
  using T = __half2;
  static constexpr int bytes = 4;
};

__device__ inline cx<double> ldc(const double2 *p, long i) {  // cached (stencil) load
  double2 v = p[i];
  return {v.x, v.y};
}
__device__ inline cx<float> ldc(const float2 *p, long i) {
  float2 v = p[i];
  return {v.x, v.y};
}
__device__ inline void st(double2 *p, long i, cx<double> v) { p[i] = make_double2(v.r, v.i); }
__device__ inline void st(float2 *p, long i, cx<float> v) { p[i] = make_float2(v.r, v.i); }

__device__ inline cx<float> ldh(const __half2 *p, long i, float scale) {
  float2 f = __half22float2(p[i]);
  return {f.x * scale, f.y * scale};
}

constexpr int BLK = 256;  // threads per block for every streaming kernel

// deterministic block reduction over NTH threads: wave shuffles, then a
// fixed-order LDS sum; the result is valid in thread 0
template <int NTH>
__device__ inline double block_sum_n(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __shared__ double sm[NTH / 64];
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NTH / 64; k++) t += sm[k];
  }
  __syncthreads();
  return t;
}
__device__ inline double block_sum(double v) { return block_sum_n<BLK>(v); }

// thread k (< BLK) of a sum of n per-block partial sums: p[k] + p[k + BLK] +
// ... in that order.  The loads are issued 8 at a time before they are added,
// so a long row costs a few memory latencies instead of one per element.
__device__ inline double partial_thread(const double *p, int n) {
  constexpr int U = 8;
  double v = 0.0;
  for (int k0 = threadIdx.x; k0 < n; k0 += U * BLK) {
    double a[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = k0 + u * BLK;
      a[u] = k < n ? p[k] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; u++)
      if (k0 + u * BLK < n) v += a[u];
  }
  return v;
}

// sum of n per-block partial sums by BLK threads (partial_thread, then
// block_sum): the result, valid in thread 0, depends only on the values and n
__device__ inline double partial_sum(const double *p, int n) {
  return block_sum(partial_thread(p, n));
}

}  // namespace thr

// status codes shared with the C ABI
#define THR_OK 0
#define THR_ERR_ARG 1
#define THR_ERR_HIP 2
#define THR_ERR_NOCONV 3
#define THR_ERR_SIZE 4
#define THR_ERR_STATE 5
#define THR_ERR_NONFINITE 6
