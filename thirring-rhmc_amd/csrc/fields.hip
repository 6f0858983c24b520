// Device-side field utilities for callers that keep everything resident:
// counter-based Gaussian fields (synthetic sources, theta) and per-vector
// norms.  No counterpart in the reference, whose noise comes from the host
// generators of random.f90/site_random.f90 (those are restated, bit for bit,
// in the Python site_random module where reference streams are needed).
//
// Generator: element pair k of a field with seed S uses the 64-bit mix
// h = mix64(S + (k + 1) * 0x9E3779B97F4A7C15) (the splitmix64 finaliser) and a
// second mix of h, i.e. two independent 64-bit words; their top 53 bits give
// uniforms u1 in (0, 1] and u2 in [0, 1), and Box-Muller turns them into two
// normals sigma sqrt(-2 ln u1) (cos, sin)(2 pi u2).  Element i depends only
// on (S, i), so fields are reproducible across launch shapes and streams.
#include <algorithm>
#include <cmath>
#include <cstdint>

#include "../../include/thirring_hip.h"
#include "solvers.h"

namespace thr {

namespace {

inline unsigned nblk(long n) { return (unsigned)((n + BLK - 1) / BLK); }

__device__ inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(BLK) void k_gauss(T *out, long n, double sigma, uint64_t seed) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;  // pair index
  if (2 * k >= n) return;
  const uint64_t h1 = mix64(seed + (uint64_t)(k + 1) * 0x9E3779B97F4A7C15ull);
  const uint64_t h2 = mix64(h1 ^ 0xD1B54A32D192ED03ull);
  const double u1 = ((double)(h1 >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
  const double u2 = (double)(h2 >> 11) * 0x1.0p-53;          // [0, 1)
  const double r = sigma * sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  out[2 * k] = (T)(r * cs);
  if (2 * k + 1 < n) out[2 * k + 1] = (T)(r * sn);
}

// partial sums of |x|^2 of nvec complex128 vectors of length n
__global__ __launch_bounds__(BLK) void k_vec_norm2(const double2 *x, long n, double *partial) {
  const int v = blockIdx.y;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * BLK + threadIdx.x; i < n; i += (long)gridDim.x * BLK) {
    const double2 a = x[(long)v * n + i];
    acc += a.x * a.x + a.y * a.y;
  }
  const double t = block_sum(acc);
  if (threadIdx.x == 0) partial[(long)v * gridDim.x + blockIdx.x] = t;
}

__global__ __launch_bounds__(BLK) void k_vec_norm_fin(const double *partial, int nb, double *out) {
  double acc = 0.0;
  for (int k = threadIdx.x; k < nb; k += BLK) acc += partial[(long)blockIdx.x * nb + k];
  const double t = block_sum(acc);
  if (threadIdx.x == 0) out[blockIdx.x] = sqrt(t);
}

}  // namespace

}  // namespace thr

using namespace thr;

extern "C" {

int thr_gaussian_dev(void *out, unsigned long long n, int prec, double sigma,
                     unsigned long long seed, void *stream) {
  THR_CHECK(out, THR_ERR_ARG, "NULL argument");
  THR_CHECK(prec == THR_F64 || prec == THR_F32, THR_ERR_ARG, "prec must be THR_F64 or THR_F32");
  THR_CHECK(n < (1ull << 36), THR_ERR_SIZE, "field too large (n < 2^36)");
  if (n == 0) return THR_OK;
  hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  const long pairs = (long)((n + 1) / 2);
  if (prec == THR_F64)
    hipLaunchKernelGGL(k_gauss<double>, dim3(nblk(pairs)), dim3(BLK), 0, s, (double *)out,
                       (long)n, sigma, (uint64_t)seed);
  else
    hipLaunchKernelGGL(k_gauss<float>, dim3(nblk(pairs)), dim3(BLK), 0, s, (float *)out, (long)n,
                       sigma, (uint64_t)seed);
  THR_HIP(hipGetLastError());
  if (!stream) THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

int thr_norms_dev(const void *x, int nvec, unsigned long long n, double *norms, void *stream) {
  THR_CHECK(x && norms, THR_ERR_ARG, "NULL argument");
  THR_CHECK(nvec >= 1 && nvec <= 65535, THR_ERR_ARG, "nvec out of range (1..65535)");
  hipStream_t s = stream ? (hipStream_t)stream : default_stream();
  const int nb = (int)std::min<long>(1024, std::max<long>(1, (long)nblk((long)n)));
  double *partial = (double *)workspace(600, sizeof(double) * (size_t)nb * nvec);
  double *dev = (double *)workspace(601, sizeof(double) * (size_t)nvec);
  double *host = (double *)host_workspace(600, sizeof(double) * (size_t)nvec);
  THR_CHECK(partial && dev && host, THR_ERR_HIP, "workspace allocation failed");
  hipLaunchKernelGGL(k_vec_norm2, dim3(nb, nvec), dim3(BLK), 0, s, (const double2 *)x, (long)n,
                     partial);
  hipLaunchKernelGGL(k_vec_norm_fin, dim3(nvec), dim3(BLK), 0, s, partial, nb, dev);
  THR_HIP(hipGetLastError());
  THR_HIP(hipMemcpyAsync(host, dev, sizeof(double) * nvec, hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  for (int k = 0; k < nvec; k++) norms[k] = host[k];
  return THR_OK;
}

}  // extern "C"
