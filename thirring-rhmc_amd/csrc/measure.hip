// Wall sources and wall contractions of the psibarpsi measurement
// (measure_shamir, measure_module.F90:170-423).
//
// A noise field is the reference's gauss0 output ps(0:Lx+1, 0:Ly+1, 0:Lt+1, 2)
// (real32, haloed; only the interior is read), the complex noise
// eta = cmplx(ps(..,1), ps(..,2)).  A wall source puts sign * eta on one spin
// at one s and zero everywhere else; a wall contraction is
// sum_x conj(eta(x)) X(s, x, spin) over the 4-D sites.  The Wilson-kernel
// measurement also projects the read slice: islice keeps two spins of X at
// that s and the contraction uses DWilson(islice, -am3) - islice.
#include "solvers.h"

namespace thr {

namespace {

inline unsigned nblk(long n) { return (unsigned)((n + BLK - 1) / BLK); }

// element (x, y, t, comp) of a haloed real32 4-D array, 0-based interior indices
__device__ inline float noise_at(const float *ps, const Dims &d, int x, int y, int t, int comp) {
  const long hx = d.lx + 2, hy = d.ly + 2, ht = d.lt + 2;
  return ps[(x + 1) + hx * ((y + 1) + hy * ((t + 1) + ht * comp))];
}

// out[spin][t][y][x][s] = sign * eta(x, y, t) at (spin, s_wall), 0 elsewhere
__global__ __launch_bounds__(BLK) void k_wall_source(const Dims d, const float *ps, int spin,
                                                     int s_wall, double sign, double2 *out) {
  const long v5 = d.v5();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= 4 * v5) return;
  const int a = (int)(i / v5);
  const long site = i - a * v5;
  const int s = (int)(site % d.ls);
  double2 v = make_double2(0.0, 0.0);
  if (a == spin && s == s_wall) {
    const long c = site / d.ls;
    const int x = (int)(c % d.lx), y = (int)((c / d.lx) % d.ly), t = (int)(c / ((long)d.lx * d.ly));
    // cmplx(ps1, ps2) is single precision; its promotion to complex(dp) is exact
    v = make_double2(sign * (double)noise_at(ps, d, x, y, t, 0),
                     sign * (double)noise_at(ps, d, x, y, t, 1));
  }
  out[i] = v;
}

// per-block partial sums of conj(eta) (X - Xs)(s_read, ., spin) (Xs optional):
// partial[2 b], partial[2 b + 1]
__global__ __launch_bounds__(BLK) void k_wall_dot(const Dims d, const float *ps, const double2 *X,
                                                  const double2 *Xs, int spin, int s_read,
                                                  double *partial) {
  const long v4 = d.v4(), v5 = d.v5();
  const long c = (long)blockIdx.x * BLK + threadIdx.x;
  double re = 0.0, im = 0.0;
  if (c < v4) {
    const int x = (int)(c % d.lx), y = (int)((c / d.lx) % d.ly), t = (int)(c / ((long)d.lx * d.ly));
    const double er = noise_at(ps, d, x, y, t, 0), ei = noise_at(ps, d, x, y, t, 1);
    const long e = spin * v5 + c * d.ls + s_read;
    double2 v = X[e];
    if (Xs) {
      const double2 w = Xs[e];
      v = make_double2(v.x - w.x, v.y - w.y);
    }
    re = er * v.x + ei * v.y;  // conj(eta) * v
    im = er * v.y - ei * v.x;
  }
  const double sr = block_sum(re);
  const double si = block_sum(im);
  if (threadIdx.x == 0) {
    partial[2 * blockIdx.x] = sr;
    partial[2 * blockIdx.x + 1] = si;
  }
}

// slice s_read of X, spins [lo, lo + 2) kept and the others zero, as a 4-D
// spinor [spin][t][y][x] (the islice of measure_module.F90:524-525, 597-598)
__global__ __launch_bounds__(BLK) void k_slice_project(const Dims d, const double2 *X, int s_read,
                                                       int lo, double2 *out) {
  const long v4 = d.v4(), v5 = d.v5();
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= 4 * v4) return;
  const int a = (int)(i / v4);
  const long c = i - a * v4;
  out[i] = (a >= lo && a < lo + 2) ? X[a * v5 + c * d.ls + s_read] : make_double2(0.0, 0.0);
}

}  // namespace

int launch_slice_project(const Dims &d, const double2 *X, int s_read, int lo, double2 *out,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_slice_project, dim3(nblk(4 * d.v4())), dim3(BLK), 0, s, d, X, s_read, lo,
                     out);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int launch_wall_source(const Dims &d, const float *ps, int spin, int s_wall, double sign,
                       double2 *out, hipStream_t s) {
  hipLaunchKernelGGL(k_wall_source, dim3(nblk(4 * d.v5())), dim3(BLK), 0, s, d, ps, spin, s_wall,
                     sign, out);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

long wall_dot_blocks(const Dims &d) { return nblk(d.v4()); }

int launch_wall_dot(const Dims &d, const float *ps, const double2 *X, const double2 *Xsub, int spin,
                    int s_read, double *partial, hipStream_t s) {
  hipLaunchKernelGGL(k_wall_dot, dim3(nblk(d.v4())), dim3(BLK), 0, s, d, ps, X, Xsub, spin, s_read,
                     partial);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

}  // namespace thr
