// C ABI entry points (declared in include/thirring_hip.h).
//
// Host-array entry points stage the reference's haloed complex128 arrays
// through device buffers: one H2D copy of the haloed array, a pack kernel
// that strips the halo into the device layout, the operator or solver, an
// unpack kernel that writes the interior and the periodic halo, one D2H copy.
#include <cstring>
#include <mutex>
#include <string>
#include <algorithm>
#include <vector>

#include "../../include/thirring_hip.h"
#include "msfast.h"
#include "rhmc.h"
#include "solvers.h"

namespace thr {

namespace {
thread_local std::string g_err;
int g_device = -1;
}  // namespace

void set_error(const std::string &msg) { g_err = msg; }
const char *last_error() { return g_err.c_str(); }
void release_workspace();

static hipStream_t pick(void *stream) {
  return stream ? (hipStream_t)stream : default_stream();
}

// entry points that only enqueue work are synchronous when called without a
// stream (their results are then ready for any copy on any stream)
static int done_if_default(void *stream, int rc) {
  if (rc || stream) return rc;
  THR_HIP(hipStreamSynchronize(default_stream()));
  return THR_OK;
}

static int check_dims(const int dims[4]) {
  THR_CHECK(dims, THR_ERR_ARG, "dims is NULL");
  for (int k = 0; k < 4; k++) THR_CHECK(dims[k] > 0, THR_ERR_ARG, "dims must be positive");
  const long v5 = (long)dims[0] * dims[1] * dims[2] * dims[3];
  THR_CHECK(v5 < (1L << 31) / 4, THR_ERR_SIZE, "lattice too large for 32-bit site indexing");
  return THR_OK;
}

static Dims mkdims(const int dims[4]) { return Dims{dims[0], dims[1], dims[2], dims[3]}; }

static long haloed5(const Dims &d) { return (long)(d.ls + 2) * (d.lx + 2) * (d.ly + 2) * (d.lt + 2) * 4; }
static long haloed4(const Dims &d) { return (long)(d.lx + 2) * (d.ly + 2) * (d.lt + 2) * 3; }

// workspace slots of the host staging path
enum {
  W_H0 = 100, W_H1, W_H2, W_L, W_LH, W_A, W_B, W_C, W_D, W_X1, W_F, W_T, W_P
};

static int stage_links(const Dims &d, const double *u, double2 **u64, hipStream_t s) {
  double2 *uh = (double2 *)workspace(W_LH, sizeof(double2) * haloed4(d));
  *u64 = (double2 *)workspace(W_L, sizeof(double2) * d.v4() * 3);
  THR_CHECK(uh && *u64, THR_ERR_HIP, "workspace allocation failed");
  THR_HIP(hipMemcpyAsync(uh, u, sizeof(double2) * haloed4(d), hipMemcpyHostToDevice, s));
  return launch_pack_links(d, uh, *u64, nullptr, s);
}

static int stage_in(const Dims &d, const double *h, int slot_h, double2 *dev, hipStream_t s) {
  double2 *hb = (double2 *)workspace(slot_h, sizeof(double2) * haloed5(d));
  THR_CHECK(hb, THR_ERR_HIP, "workspace allocation failed");
  THR_HIP(hipMemcpyAsync(hb, h, sizeof(double2) * haloed5(d), hipMemcpyHostToDevice, s));
  return launch_pack_spinor(d, hb, dev, nullptr, P64, s);
}

static int stage_out(const Dims &d, const double2 *dev, int slot_h, double *h, hipStream_t s) {
  double2 *hb = (double2 *)workspace(slot_h, sizeof(double2) * haloed5(d));
  THR_CHECK(hb, THR_ERR_HIP, "workspace allocation failed");
  int rc = launch_unpack_spinor(d, dev, nullptr, P64, hb, 1, s);
  if (rc) return rc;
  THR_HIP(hipMemcpyAsync(h, hb, sizeof(double2) * haloed5(d), hipMemcpyDeviceToHost, s));
  return THR_OK;
}

static OpDesc mkop(const Dims &d, int kernel, int imass, float am, float am3, const double2 *u64) {
  OpDesc op{};
  op.d = d;
  op.kernel = kernel;
  op.imass = imass;
  op.am = am;
  op.am3 = am3;
  op.u64 = u64;
  return op;
}

static int check_op(int kernel, int imass) {
  THR_CHECK(kernel == 0 || kernel == 1, THR_ERR_ARG, "kernel must be 0 (Shamir) or 1 (Wilson)");
  THR_CHECK(imass == 1 || imass == 3 || imass == 5, THR_ERR_ARG, "imass must be 1, 3 or 5");
  THR_CHECK(kernel == 0 || imass != 5, THR_ERR_ARG, "the Wilson kernel supports imass 1 and 3");
  return THR_OK;
}

#define TRY(x)              \
  do {                      \
    int _rc = (x);          \
    if (_rc) return _rc;    \
  } while (0)

static int host_op(double *phi, const double *r, const double *u, const int dims[4], float am,
                   int imass, float am3, int kernel, int dagger) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(phi && r && u, THR_ERR_ARG, "NULL array");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * d.nspinor());
  double2 *b = (double2 *)workspace(W_B, sizeof(double2) * d.nspinor());
  THR_CHECK(a && b, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, r, W_H0, a, s));
  TRY(apply_op(mkop(d, kernel, imass, am, am3, u64), dagger, a, b, s));
  TRY(stage_out(d, b, W_H1, phi, s));
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

}  // namespace thr

using namespace thr;

extern "C" {

int thr_init(int device) {
  int n = 0;
  THR_HIP(hipGetDeviceCount(&n));
  THR_CHECK(n > 0, THR_ERR_HIP, "no HIP device visible");
  THR_CHECK(device >= 0 && device < n, THR_ERR_ARG, "device index out of range");
  THR_HIP(hipSetDevice(device));
  g_device = device;
  THR_CHECK(default_stream(), THR_ERR_HIP, "stream creation failed");
  return THR_OK;
}

int thr_finalize(void) {
  THR_HIP(hipDeviceSynchronize());
  release_workspace();
  return THR_OK;
}

const char *thr_last_error(void) { return last_error(); }

int thr_context_create(void **ctx) {
  THR_CHECK(ctx, THR_ERR_ARG, "NULL argument");
  *ctx = (void *)context_create();
  return THR_OK;
}

int thr_context_destroy(void *ctx) {
  THR_CHECK(ctx, THR_ERR_ARG, "NULL context");
  context_destroy((Context *)ctx);
  return THR_OK;
}

int thr_context_bind(void *ctx) {
  context_bind((Context *)ctx);
  return THR_OK;
}

int thr_context_bytes(unsigned long long *bytes) {
  THR_CHECK(bytes, THR_ERR_ARG, "NULL argument");
  *bytes = (unsigned long long)workspace_bytes();
  return THR_OK;
}

int thr_context_trim(void) {
  // the context's own stream may still read the buffers: wait for it first
  THR_HIP(hipStreamSynchronize(default_stream()));
  release_workspace();
  return THR_OK;
}

int thr_context_set_reliable(double delta) {
  THR_CHECK(delta >= 0.0 && delta < 1.0, THR_ERR_ARG, "reliable-update factor out of range [0, 1)");
  context_set_reliable(delta);
  return THR_OK;
}

int thr_context_reliable_updates(int *count) {
  THR_CHECK(count, THR_ERR_ARG, "NULL argument");
  *count = context_last_reliable();
  return THR_OK;
}

static int links_to_host(const Dims &d, const double2 *u64, double *u, hipStream_t s);

int thr_coef(double *u, const float *theta, const int dims[4], int ibound) {
  TRY(check_dims(dims));
  THR_CHECK(u && theta, THR_ERR_ARG, "NULL array");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  float *th = (float *)workspace(W_F, sizeof(float) * d.v4() * 3);
  double2 *u64 = (double2 *)workspace(W_L, sizeof(double2) * d.v4() * 3);
  THR_CHECK(th && u64, THR_ERR_HIP, "workspace allocation failed");
  THR_HIP(hipMemcpyAsync(th, theta, sizeof(float) * d.v4() * 3, hipMemcpyHostToDevice, s));
  TRY(launch_coef(d, th, ibound, u64, s));
  return links_to_host(d, u64, u, s);
}

// device links [mu][t][y][x] -> the haloed host array u(0:Lx+1,0:Ly+1,0:Lt+1,3)
// with the periodic copy of the interior in the halo
static int links_to_host(const Dims &d, const double2 *u64, double *u, hipStream_t s) {
  std::vector<double2> h(d.v4() * 3);
  THR_HIP(hipMemcpyAsync(h.data(), u64, sizeof(double2) * h.size(), hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  const long X = d.lx + 2, Y = d.ly + 2, T = d.lt + 2;
  for (int mu = 0; mu < 3; mu++)
    for (int t = 0; t < T; t++)
      for (int y = 0; y < Y; y++)
        for (int x = 0; x < X; x++) {
          const int xi = x == 0 ? d.lx - 1 : (x == X - 1 ? 0 : x - 1);
          const int yi = y == 0 ? d.ly - 1 : (y == Y - 1 ? 0 : y - 1);
          const int ti = t == 0 ? d.lt - 1 : (t == T - 1 ? 0 : t - 1);
          const double2 v = h[xi + (long)d.lx * (yi + (long)d.ly * (ti + (long)d.lt * mu))];
          const long o = 2 * (x + X * (y + Y * (t + T * (long)mu)));
          u[o] = v.x;
          u[o + 1] = v.y;
        }
  return THR_OK;
}

int thr_dslash(double *phi, const double *r, const double *u, const int dims[4], float am,
               int imass, float am3, int kernel) {
  return host_op(phi, r, u, dims, am, imass, am3, kernel, 0);
}

int thr_dslashd(double *phi, const double *r, const double *u, const int dims[4], float am,
                int imass, float am3, int kernel) {
  return host_op(phi, r, u, dims, am, imass, am3, kernel, 1);
}

int thr_dirac_op_shifted(double *xout, const double *xin, const double *u, const int dims[4],
                         float am, int imass, float am3, double shift, int kernel) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(xout && xin && u, THR_ERR_ARG, "NULL array");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  const long n = d.nspinor();
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * n);
  double2 *b = (double2 *)workspace(W_B, sizeof(double2) * n);
  double2 *c = (double2 *)workspace(W_C, sizeof(double2) * n);
  THR_CHECK(a && b && c, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, xin, W_H0, a, s));
  const OpDesc op = mkop(d, kernel, imass, am, am3, u64);
  TRY(apply_op(op, 0, a, b, s));
  TRY(apply_op(op, 1, b, c, s));
  double2 *hb = (double2 *)workspace(W_H1, sizeof(double2) * haloed5(d));
  TRY(launch_unpack_spinor(d, c, nullptr, P64, hb, 1, s));
  THR_HIP(hipMemcpyAsync(xout, hb, sizeof(double2) * haloed5(d), hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  if (shift != 0.0) {
    const long nh = haloed5(d) * 2;
    for (long k = 0; k < nh; k++) xout[k] = xout[k] + shift * xin[k];
  }
  return THR_OK;
}

int thr_congrad(const double *phi, double *x, const double *u, const int dims[4], float res,
                float am, int imass, float am3, int niterc, int kernel, int *itercg,
                int *iterations) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(phi && x && u && itercg, THR_ERR_ARG, "NULL argument");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  const long n = d.nspinor();
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * n);
  double2 *b = (double2 *)workspace(W_B, sizeof(double2) * n);
  THR_CHECK(a && b, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, phi, W_H0, a, s));
  TRY(stage_in(d, x, W_H1, b, s));
  TRY(cg_reference(mkop(d, kernel, imass, am, am3, u64), a, b, res, niterc, itercg, iterations, s));
  TRY(stage_out(d, b, W_H1, x, s));
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

int thr_measure_shamir_noise(const float *ps, const float *pt, const double *u, const int dims[4],
                             float res, float am, int imass, float am3, int niterc, int kernel,
                             double pbp1[2], double pbp2[2], int *iters) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(ps && pt && u && pbp1 && pbp2 && iters, THR_ERR_ARG, "NULL argument");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  const long n = d.nspinor();
  const long nh4 = (long)(d.lx + 2) * (d.ly + 2) * (d.lt + 2) * 2;
  const long nbd = wall_dot_blocks(d);
  double2 *xi = (double2 *)workspace(W_A, sizeof(double2) * n);
  double2 *phi = (double2 *)workspace(W_B, sizeof(double2) * n);
  float *noise = (float *)workspace(W_P, sizeof(float) * 2 * nh4);
  double *part = (double *)workspace(W_T, sizeof(double) * 2 * nbd);
  double *hpart = (double *)host_workspace(W_T, sizeof(double) * 2 * nbd);
  THR_CHECK(xi && phi && noise && part && hpart, THR_ERR_HIP, "workspace allocation failed");
  THR_HIP(hipMemcpyAsync(noise, ps, sizeof(float) * nh4, hipMemcpyHostToDevice, s));
  THR_HIP(hipMemcpyAsync(noise + nh4, pt, sizeof(float) * nh4, hipMemcpyHostToDevice, s));
  // the source is dslashd_shamir's input whatever operator congrad uses
  // (measure_module.F90:245,325); congrad's guess is the source (vector::X)
  const OpDesc shamir = mkop(d, 0, imass, am, am3, u64);
  const OpDesc op = mkop(d, kernel, imass, am, am3, u64);
  const int wall0 = 0, wallL = d.ls - 1;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  int total = 0;
  for (int id = 0; id < 2; id++) {
    for (int side = 0; side < 2; side++) {
      // side 0: source on s = 1 from ps (psibarpsi1); side 1: on s = Ls from pt (psibarpsi2).
      // imass 1, 3: eta on spin id (side 0) / id + 2 (side 1), read at the other wall on the
      // same spin; imass 5: -eta on the chirality-flipped spin, read at the same wall
      // (measure_module.F90:224-234, 255-266, 284-293, 312-324, 334-348)
      const float *eta = noise + side * nh4;
      const int s_src = side == 0 ? wall0 : wallL;
      int spin_src, spin_read, s_read;
      double sign;
      if (imass != 5) {
        spin_src = spin_read = side == 0 ? id : id + 2;
        s_read = side == 0 ? wallL : wall0;
        sign = 1.0;
      } else {
        spin_src = side == 0 ? id : id + 2;
        spin_read = side == 0 ? id + 2 : id;
        s_read = s_src;
        sign = -1.0;
      }
      TRY(launch_wall_source(d, eta, spin_src, s_src, sign, xi, s));
      TRY(apply_op(shamir, 1, xi, phi, s));
      int itercg = 0;
      TRY(cg_reference(op, phi, xi, res, niterc, &itercg, nullptr, s));
      total += itercg;
      TRY(launch_wall_dot(d, eta, xi, nullptr, spin_read, s_read, part, s));
      THR_HIP(hipMemcpyAsync(hpart, part, sizeof(double) * 2 * nbd, hipMemcpyDeviceToHost, s));
      THR_HIP(hipStreamSynchronize(s));
      for (long b = 0; b < nbd; b++) {
        acc[side][0] += hpart[2 * b];
        acc[side][1] += hpart[2 * b + 1];
      }
    }
  }
  pbp1[0] = acc[0][0];
  pbp1[1] = acc[0][1];
  pbp2[0] = acc[1][0];
  pbp2[1] = acc[1][1];
  *iters = total;
  return THR_OK;
}

int thr_measure_wilson_noise(const float *ps, const float *pt, const double *u, const int dims[4],
                             float res, float am, int imass, float am3, int niterc, int kernel,
                             double pbp1[2], double pbp2[2], int *iters) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(imass == 1 || imass == 3, THR_ERR_ARG, "the Wilson-kernel measurement supports imass 1 and 3");
  THR_CHECK(ps && pt && u && pbp1 && pbp2 && iters, THR_ERR_ARG, "NULL argument");
  const Dims d = mkdims(dims);
  const Dims d4{d.lx, d.ly, d.lt, 1};
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  const long n = d.nspinor();
  const long nh4 = (long)(d.lx + 2) * (d.ly + 2) * (d.lt + 2) * 2;
  const long nbd = wall_dot_blocks(d);
  double2 *xi = (double2 *)workspace(W_A, sizeof(double2) * n);
  double2 *phi = (double2 *)workspace(W_B, sizeof(double2) * n);
  double2 *isl = (double2 *)workspace(W_C, sizeof(double2) * 4 * d.v4());
  double2 *osl = (double2 *)workspace(W_D, sizeof(double2) * 4 * d.v4());
  float *noise = (float *)workspace(W_P, sizeof(float) * 2 * nh4);
  double *part = (double *)workspace(W_T, sizeof(double) * 2 * nbd);
  double *hpart = (double *)host_workspace(W_T, sizeof(double) * 2 * nbd);
  THR_CHECK(xi && phi && isl && osl && noise && part && hpart, THR_ERR_HIP,
            "workspace allocation failed");
  THR_HIP(hipMemcpyAsync(noise, ps, sizeof(float) * nh4, hipMemcpyHostToDevice, s));
  THR_HIP(hipMemcpyAsync(noise + nh4, pt, sizeof(float) * nh4, hipMemcpyHostToDevice, s));
  // the source goes through dslashd_wilson whatever operator congrad uses
  // (measure_module.F90:509-533, 582-598)
  const OpDesc wilson = mkop(d, 1, imass, am, am3, u64);
  const OpDesc op = mkop(d, kernel, imass, am, am3, u64);
  // DWilson(islice, -am3): the 4-D Wilson slice operator, diag = 3 - am3 (dirac.F90:181-224)
  DslashLaunch W{};
  W.d = d4;
  W.nrhs = 1;
  W.prec = P64;
  W.imass = 3;
  W.am3 = am3;
  W.links = u64;
  W.mode = EPI_STORE;
  W.sc_stride = 1;
  W.done_stride = 1;
  W.wilson4d = 1;
  W.in = isl;
  W.out = osl;
  double acc[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
  int total = 0;
  for (int side = 0; side < 2; side++) {
    // side 0: sources on s = 1 from ps, read at s = Ls through P+ (spins 1, 2);
    // side 1: sources on s = Ls from pt, read at s = 1 through P- (spins 3, 4)
    const float *eta = noise + side * nh4;
    const int s_src = side == 0 ? 0 : d.ls - 1, s_read = side == 0 ? d.ls - 1 : 0;
    const int lo = side == 0 ? 0 : 2;
    for (int id = 0; id < 4; id++) {
      TRY(launch_wall_source(d, eta, id, s_src, 1.0, xi, s));
      TRY(apply_op(wilson, 1, xi, phi, s));
      int itercg = 0;
      TRY(cg_reference(op, phi, xi, res, niterc, &itercg, nullptr, s));
      total += itercg;
      TRY(launch_slice_project(d, xi, s_read, lo, isl, s));
      TRY(launch_dslash(W, s));
      TRY(launch_wall_dot(d4, eta, osl, isl, id, 0, part, s));
      THR_HIP(hipMemcpyAsync(hpart, part, sizeof(double) * 2 * nbd, hipMemcpyDeviceToHost, s));
      THR_HIP(hipStreamSynchronize(s));
      for (long b = 0; b < nbd; b++) {
        acc[side][0] += hpart[2 * b];
        acc[side][1] += hpart[2 * b + 1];
      }
    }
  }
  pbp1[0] = acc[0][0];
  pbp1[1] = acc[0][1];
  pbp2[0] = acc[1][0];
  pbp2[1] = acc[1][1];
  *iters = total;
  return THR_OK;
}

int thr_multishift(const double *u, const int dims[4], float am, int imass, float am3,
                   int ndiagq, const double *aden, const double *anum, double *output,
                   const double *input, float res, int maxcg, int use_sp, int kernel,
                   int *cg_return, int *cg_returns) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(u && aden && anum && output && input && cg_return, THR_ERR_ARG, "NULL argument");
  THR_CHECK(ndiagq >= 1 && ndiagq <= MAXSHIFT, THR_ERR_ARG, "ndiagq out of range");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  const long n = d.nspinor();
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * n);
  double2 *x1 = (double2 *)workspace(W_X1, sizeof(double2) * n * ndiagq);
  THR_CHECK(a && x1, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, input, W_H0, a, s));
  const OpDesc op = mkop(d, kernel, imass, am, am3, u64);
  if (use_sp)
    TRY(multishift_reference_sp(op, ndiagq, aden, anum, a, x1, res, maxcg, cg_return, cg_returns, s));
  else
    TRY(multishift_reference(op, ndiagq, aden, anum, a, x1, res, maxcg, cg_return, cg_returns, s));
  // the device shift layout equals the halo-free Fortran output(Ls,Lx,Ly,Lt,4,n)
  THR_HIP(hipMemcpyAsync(output, x1, sizeof(double2) * n * ndiagq, hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

int thr_derivs(const double *r, const double *x2, double anum, int iflag, const int dims[4],
               float am, int imass, float am3, int kernel, double *dsdpi) {
  (void)am3;
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(r && x2 && dsdpi, THR_ERR_ARG, "NULL argument");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  const long n = d.nspinor();
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * n);
  double2 *b = (double2 *)workspace(W_B, sizeof(double2) * n);
  double *f = (double *)workspace(W_T, sizeof(double) * d.v4() * 3);
  THR_CHECK(a && b && f, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, r, W_H0, a, s));
  TRY(stage_in(d, x2, W_H1, b, s));
  THR_HIP(hipMemcpyAsync(f, dsdpi, sizeof(double) * d.v4() * 3, hipMemcpyHostToDevice, s));
  TRY(derivs_dispatch(d, kernel, a, b, anum, iflag, am, imass, f, s));
  THR_HIP(hipMemcpyAsync(dsdpi, f, sizeof(double) * d.v4() * 3, hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

int thr_qmrherm(const double *phi, double *x, float res, int *itercg, float am, int imass,
                float am3, const double *anum, const double *aden, int ndiagq, int iflag,
                int use_sp, const double *u, const int dims[4], int max_qmr_iters, int kernel,
                double *x1h, double *phi0, double *dsdpi, int *cg_returns) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(phi && x && itercg && anum && aden && u && x1h, THR_ERR_ARG, "NULL argument");
  THR_CHECK((iflag != 1 && iflag != 3) || phi0, THR_ERR_ARG, "iflag 1/3 need the Phi0 register");
  THR_CHECK(iflag < 2 || dsdpi, THR_ERR_ARG, "iflag 2/3 need dSdpi");
  THR_CHECK(ndiagq >= 1 && ndiagq <= 25, THR_ERR_ARG,
            "qmrherm requires ndiagq <= ndiag = 25 (qmrherm_module.F90:66-73)");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  const long n = d.nspinor();
  const long nh = haloed5(d);
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * n);
  double2 *b = (double2 *)workspace(W_B, sizeof(double2) * n);
  double2 *x1 = (double2 *)workspace(W_X1, sizeof(double2) * n * ndiagq);
  double2 *p0 = (double2 *)workspace(W_C, sizeof(double2) * n * ndiagq);
  double *f = (double *)workspace(W_T, sizeof(double) * d.v4() * 3);
  THR_CHECK(a && b && x1 && p0 && f, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, phi, W_H0, a, s));
  if (iflag == 3)
    for (int k = 0; k < ndiagq; k++) TRY(stage_in(d, phi0 + 2 * nh * k, W_H2, p0 + n * k, s));
  if (iflag >= 2)
    THR_HIP(hipMemcpyAsync(f, dsdpi, sizeof(double) * d.v4() * 3, hipMemcpyHostToDevice, s));
  const OpDesc op = mkop(d, kernel, imass, am, am3, u64);
  TRY(qmrherm_device(op, a, b, res, itercg, anum, aden, ndiagq, iflag, use_sp, max_qmr_iters, x1,
                     p0, f, cg_returns, s));
  THR_HIP(hipMemcpyAsync(x1h, x1, sizeof(double2) * n * ndiagq, hipMemcpyDeviceToHost, s));
  if (iflag < 2) {
    TRY(stage_out(d, b, W_H1, x, s));  // interior plus periodic halo (96-108)
    if (iflag == 1)
      for (int k = 0; k < ndiagq; k++) {
        double2 *hb = (double2 *)workspace(W_H2, sizeof(double2) * nh);
        THR_CHECK(hb, THR_ERR_HIP, "workspace allocation failed");
        TRY(launch_unpack_spinor(d, p0 + n * k, nullptr, P64, hb, 1, s));
        THR_HIP(hipMemcpyAsync(phi0 + 2 * nh * k, hb, sizeof(double2) * nh, hipMemcpyDeviceToHost, s));
        THR_HIP(hipStreamSynchronize(s));
      }
  } else {
    // x = anum(0)*Phi over the whole haloed array, halo included (83)
    THR_HIP(hipStreamSynchronize(s));
    for (long k = 0; k < 2 * nh; k++) x[k] = anum[0] * phi[k];
    THR_HIP(hipMemcpyAsync(dsdpi, f, sizeof(double) * d.v4() * 3, hipMemcpyDeviceToHost, s));
  }
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

int thr_force(const double *phi, float res1, float am, int imass, float am3, float beta,
              int ibound, const float *theta, const double *u, const int dims[4],
              const double *anum4g, const double *aden4g, const double *bnum2g,
              const double *bden2g, const double *anum2g, const double *aden2g, int ndiagg,
              int use_sp, int max_qmr_iters, int kernel, double *dsdpi, int *itercg) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(phi && theta && u && anum4g && aden4g && bnum2g && bden2g && anum2g && aden2g && dsdpi,
            THR_ERR_ARG, "NULL argument");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * d.nspinor());
  float *th = (float *)workspace(W_F, sizeof(float) * d.v4() * 3);
  double *f = (double *)workspace(W_T, sizeof(double) * d.v4() * 3);
  THR_CHECK(a && th && f, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, phi, W_H0, a, s));
  THR_HIP(hipMemcpyAsync(th, theta, sizeof(float) * d.v4() * 3, hipMemcpyHostToDevice, s));
  ForceArgs A{};
  A.op = mkop(d, kernel, imass, am, am3, u64);
  A.phi = a;
  A.theta = th;
  A.res1 = res1;
  A.beta = beta;
  A.ibound = ibound;
  A.ndiagg = ndiagg;
  A.use_sp = use_sp;
  A.max_qmr_iters = max_qmr_iters;
  A.anum4g = anum4g;
  A.aden4g = aden4g;
  A.bnum2g = bnum2g;
  A.bden2g = bden2g;
  A.anum2g = anum2g;
  A.aden2g = aden2g;
  TRY(force_device(A, f, itercg, s));
  THR_HIP(hipMemcpyAsync(dsdpi, f, sizeof(double) * d.v4() * 3, hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  return THR_OK;
}

static EvolveArgs evolve_args(int iterl, float dt, float res, float am, int imass, float am3,
                              float beta, int ibound, const Dims &d, const double *anum4g,
                              const double *aden4g, const double *bnum2g, const double *bden2g,
                              const double *anum2g, const double *aden2g, int ndiagg, int use_sp,
                              int max_qmr_iters, int kernel, thr_uniform_fn ytest,
                              void *ytest_ctx) {
  EvolveArgs A{};
  A.f.op = mkop(d, kernel, imass, am, am3, nullptr);
  A.f.res1 = res;
  A.f.beta = beta;
  A.f.ibound = ibound;
  A.f.ndiagg = ndiagg;
  A.f.use_sp = use_sp;
  A.f.max_qmr_iters = max_qmr_iters;
  A.f.anum4g = anum4g;
  A.f.aden4g = aden4g;
  A.f.bnum2g = bnum2g;
  A.f.bden2g = bden2g;
  A.f.anum2g = anum2g;
  A.f.aden2g = aden2g;
  A.iterl = iterl;
  A.dt = dt;
  A.ytest = ytest;
  A.ytest_ctx = ytest_ctx;
  return A;
}

int thr_evolve_theta_pp(int iterl, float dt, float *pp, const double *phi, float res, float am,
                        int imass, float am3, float beta, int ibound, float *theta, double *u,
                        const int dims[4], const double *anum4g, const double *aden4g,
                        const double *bnum2g, const double *bden2g, const double *anum2g,
                        const double *aden2g, int ndiagg, int use_sp, int max_qmr_iters,
                        int kernel, thr_uniform_fn ytest, void *ytest_ctx, double *dsdpi,
                        int *itot) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(pp && phi && theta && u && anum4g && aden4g && bnum2g && bden2g && anum2g && aden2g &&
                ytest && dsdpi,
            THR_ERR_ARG, "NULL argument");
  THR_CHECK(iterl >= 1, THR_ERR_ARG, "iterl must be positive");
  const Dims d = mkdims(dims);
  const long n4 = d.v4() * 3;
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * d.nspinor());
  float *th = (float *)workspace(W_F, sizeof(float) * n4);
  double *f = (double *)workspace(W_T, sizeof(double) * n4);
  float *pd = (float *)workspace(W_P, sizeof(float) * n4);
  THR_CHECK(a && th && f && pd, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, phi, W_H0, a, s));
  THR_HIP(hipMemcpyAsync(th, theta, sizeof(float) * n4, hipMemcpyHostToDevice, s));
  THR_HIP(hipMemcpyAsync(pd, pp, sizeof(float) * n4, hipMemcpyHostToDevice, s));
  EvolveArgs A = evolve_args(iterl, dt, res, am, imass, am3, beta, ibound, d, anum4g, aden4g,
                             bnum2g, bden2g, anum2g, aden2g, ndiagg, use_sp, max_qmr_iters, kernel,
                             ytest, ytest_ctx);
  A.f.phi = a;
  int steps = 0;
  TRY(evolve_device(A, th, pd, u64, f, &steps, s));
  THR_HIP(hipMemcpyAsync(theta, th, sizeof(float) * n4, hipMemcpyDeviceToHost, s));
  THR_HIP(hipMemcpyAsync(pp, pd, sizeof(float) * n4, hipMemcpyDeviceToHost, s));
  THR_HIP(hipMemcpyAsync(dsdpi, f, sizeof(double) * n4, hipMemcpyDeviceToHost, s));
  TRY(links_to_host(d, u64, u, s));
  if (itot) *itot += steps;
  return THR_OK;
}

int thr_evolve_theta_pp_dev(int iterl, float dt, float *pp, const void *phi, float res, float am,
                            int imass, float am3, float beta, int ibound, float *theta,
                            void *links64, const int dims[4], const double *anum4g,
                            const double *aden4g, const double *bnum2g, const double *bden2g,
                            const double *anum2g, const double *aden2g, int ndiagg, int use_sp,
                            int max_qmr_iters, int kernel, thr_uniform_fn ytest, void *ytest_ctx,
                            double *dsdpi, int *itot, void *stream) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(pp && phi && theta && links64 && anum4g && aden4g && bnum2g && bden2g && anum2g &&
                aden2g && ytest && dsdpi,
            THR_ERR_ARG, "NULL argument");
  THR_CHECK(iterl >= 1, THR_ERR_ARG, "iterl must be positive");
  const Dims d = mkdims(dims);
  EvolveArgs A = evolve_args(iterl, dt, res, am, imass, am3, beta, ibound, d, anum4g, aden4g,
                             bnum2g, bden2g, anum2g, aden2g, ndiagg, use_sp, max_qmr_iters, kernel,
                             ytest, ytest_ctx);
  A.f.phi = (const double2 *)phi;
  int steps = 0;
  TRY(evolve_device(A, theta, pp, (double2 *)links64, dsdpi, &steps, pick(stream)));
  if (itot) *itot += steps;
  return THR_OK;
}

int thr_hamilton(const double *phi, float res2, float am, int imass, float am3, float beta,
                 const float *theta, const float *pp, const double *u, const int dims[4],
                 const double *anum4, const double *aden4, const double *bnum2,
                 const double *bden2, int ndiag, int use_sp, int max_qmr_iters, int kernel,
                 double *out, int *itercg) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(phi && theta && pp && u && anum4 && aden4 && bnum2 && bden2 && out, THR_ERR_ARG,
            "NULL argument");
  const Dims d = mkdims(dims);
  hipStream_t s = default_stream();
  double2 *u64;
  TRY(stage_links(d, u, &u64, s));
  double2 *a = (double2 *)workspace(W_A, sizeof(double2) * d.nspinor());
  THR_CHECK(a, THR_ERR_HIP, "workspace allocation failed");
  TRY(stage_in(d, phi, W_H0, a, s));
  HamArgs A{};
  A.op = mkop(d, kernel, imass, am, am3, u64);
  A.phi = a;
  A.theta = theta;
  A.pp = pp;
  A.res2 = res2;
  A.beta = beta;
  A.ndiag = ndiag;
  A.use_sp = use_sp;
  A.max_qmr_iters = max_qmr_iters;
  A.anum4 = anum4;
  A.aden4 = aden4;
  A.bnum2 = bnum2;
  A.bden2 = bden2;
  return hamilton_device(A, out, itercg, s);
}

// ---- device-resident ------------------------------------------------------

int thr_coef_dev(void *links64, const float *theta, const int dims[4], int ibound, void *stream) {
  TRY(check_dims(dims));
  THR_CHECK(links64 && theta, THR_ERR_ARG, "NULL argument");
  return done_if_default(stream,
                         launch_coef(mkdims(dims), theta, ibound, (double2 *)links64, pick(stream)));
}

int thr_links_to_f32_dev(void *links32, const void *links64, const int dims[4], void *stream) {
  TRY(check_dims(dims));
  THR_CHECK(links32 && links64, THR_ERR_ARG, "NULL argument");
  const Dims d = mkdims(dims);
  return done_if_default(
      stream, convert_d2f(d.v4() * 3, (const double2 *)links64, (float2 *)links32, pick(stream)));
}

int thr_device_copy(void *dst, const void *src, unsigned long long bytes, void *stream) {
  THR_CHECK(dst && src, THR_ERR_ARG, "NULL argument");
  return done_if_default(stream, stream_copy(src, dst, (size_t)bytes, pick(stream)));
}

int thr_dslash_dev(void *out, float *out_norm, const void *in, const float *in_norm,
                   const void *links, const int dims[4], int nrhs, const float *am, int imass,
                   float am3, int dagger, int prec, void *stream) {
  TRY(check_dims(dims));
  THR_CHECK(out && in && links && am, THR_ERR_ARG, "NULL argument");
  THR_CHECK(prec >= 0 && prec <= 2, THR_ERR_ARG, "prec must be THR_F64, THR_F32 or THR_F16");
  THR_CHECK(prec != 2 || (in_norm && out_norm), THR_ERR_ARG, "FP16 spinors need norm arrays");
  THR_CHECK(nrhs >= 1 && nrhs <= MAXRHS, THR_ERR_ARG, "nrhs out of range (1..32)");
  DslashLaunch L{};
  L.d = mkdims(dims);
  L.nrhs = nrhs;
  L.prec = prec;
  L.dagger = dagger ? 1 : 0;
  L.imass = imass;
  L.am3 = am3;
  for (int k = 0; k < nrhs; k++) L.am[k] = am[k];
  L.in = in;
  L.in_norm = in_norm;
  L.out = out;
  L.out_norm = out_norm;
  L.links = links;
  L.mode = EPI_STORE;
  L.sc_stride = 1;
  L.done_stride = 1;
  return done_if_default(stream, launch_dslash(L, pick(stream)));
}

static int shape_flags(const DslashConfig &c) {
  return (c.remap ? THR_SHAPE_XCD_REMAP : 0) | (c.uniform_links ? THR_SHAPE_UNIFORM_LINKS : 0) |
         (c.laneshift ? THR_SHAPE_LANESHIFT : 0) | (c.box ? THR_SHAPE_XCD_BOXES : 0) |
         (c.lds_x ? THR_SHAPE_LDS_X : 0) | (c.nt_stores ? THR_SHAPE_NT_STORES : 0) |
         ((c.occ_cap & 3) << THR_SHAPE_OCC_SHIFT) | (c.lds_y ? THR_SHAPE_LDS_Y : 0);
}

int thr_dslash_autotune(const void *links, const int dims[4], int prec, int *tchunk, int *flags,
                        void *stream) {
  TRY(check_dims(dims));
  THR_CHECK(links, THR_ERR_ARG, "NULL argument");
  THR_CHECK(prec >= 0 && prec <= 2, THR_ERR_ARG, "prec must be THR_F64, THR_F32 or THR_F16");
  DslashConfig c;
  TRY(dslash_autotune(mkdims(dims), prec, links, &c, pick(stream)));
  if (tchunk) *tchunk = c.tchunk;
  if (flags) *flags = shape_flags(c);
  return THR_OK;
}

int thr_dslash_autotune_nrhs(const void *links, const int dims[4], int prec, int nrhs,
                             int *tchunk, int *flags, void *stream) {
  TRY(check_dims(dims));
  THR_CHECK(links, THR_ERR_ARG, "NULL argument");
  THR_CHECK(prec >= 0 && prec <= 2, THR_ERR_ARG, "prec must be THR_F64, THR_F32 or THR_F16");
  THR_CHECK(nrhs >= 1 && nrhs <= MAXRHS, THR_ERR_ARG, "nrhs out of range (1..32)");
  DslashConfig c;
  TRY(dslash_autotune(mkdims(dims), prec, links, &c, pick(stream), nrhs));
  if (tchunk) *tchunk = c.tchunk;
  if (flags) *flags = shape_flags(c);
  return THR_OK;
}

int thr_dslash_set_config(const int dims[4], int prec, int tchunk, int flags) {
  TRY(check_dims(dims));
  THR_CHECK(prec >= 0 && prec <= 2, THR_ERR_ARG, "prec must be THR_F64, THR_F32 or THR_F16");
  THR_CHECK(tchunk >= 1 && tchunk <= dims[2], THR_ERR_ARG, "tchunk must be in 1..Lt");
  THR_CHECK((flags & ~511) == 0, THR_ERR_ARG, "unknown shape flags");
  DslashConfig c;
  c.tchunk = tchunk;
  c.remap = (flags & THR_SHAPE_XCD_REMAP) ? 1 : 0;
  c.uniform_links = (flags & THR_SHAPE_UNIFORM_LINKS) ? 1 : 0;
  c.laneshift = (flags & THR_SHAPE_LANESHIFT) ? 1 : 0;
  c.box = (flags & THR_SHAPE_XCD_BOXES) ? 1 : 0;
  c.lds_x = (flags & THR_SHAPE_LDS_X) ? 1 : 0;
  c.nt_stores = (flags & THR_SHAPE_NT_STORES) ? 1 : 0;
  c.occ_cap = (flags >> THR_SHAPE_OCC_SHIFT) & 3;
  c.lds_y = (flags & THR_SHAPE_LDS_Y) ? 1 : 0;
  dslash_set_config(mkdims(dims), prec, c);
  return THR_OK;
}

int thr_dslash_get_config(const int dims[4], int prec, int *tchunk, int *flags) {
  return thr_dslash_get_config_nrhs(dims, prec, 1, tchunk, flags);
}

int thr_dslash_get_config_nrhs(const int dims[4], int prec, int nrhs, int *tchunk, int *flags) {
  TRY(check_dims(dims));
  THR_CHECK(prec >= 0 && prec <= 2, THR_ERR_ARG, "prec must be THR_F64, THR_F32 or THR_F16");
  THR_CHECK(nrhs >= 1 && nrhs <= MAXRHS, THR_ERR_ARG, "nrhs out of range (1..32)");
  const DslashConfig c = dslash_config(mkdims(dims), prec, nrhs);
  if (tchunk) *tchunk = c.tchunk;
  if (flags) *flags = shape_flags(c);
  return THR_OK;
}

int thr_congrad_dev(const void *phi, void *x, const void *links64, const int dims[4], float res,
                    float am, int imass, float am3, int niterc, int kernel, int *itercg,
                    int *iterations, void *stream) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(phi && x && links64 && itercg, THR_ERR_ARG, "NULL argument");
  const Dims d = mkdims(dims);
  return cg_reference(mkop(d, kernel, imass, am, am3, (const double2 *)links64),
                      (const double2 *)phi, (double2 *)x, res, niterc, itercg, iterations,
                      pick(stream));
}

int thr_derivs_dev(const void *r, const void *x2, double anum, int iflag, const int dims[4],
                   float am, int imass, int kernel, void *dsdpi, void *stream) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(r && x2 && dsdpi, THR_ERR_ARG, "NULL argument");
  hipStream_t s = pick(stream);
  return done_if_default(stream, derivs_dispatch(mkdims(dims), kernel, (const double2 *)r,
                                                 (const double2 *)x2, anum, iflag, am, imass,
                                                 (double *)dsdpi, s));
}

int thr_derivs_batch_dev(const void *const *r, const void *const *x2, const double *anum,
                         const int *iflag, int npair, const int dims[4], void *dsdpi,
                         void *stream) {
  TRY(check_dims(dims));
  THR_CHECK(r && x2 && anum && iflag && dsdpi, THR_ERR_ARG, "NULL argument");
  THR_CHECK(npair >= 1 && npair <= 1024, THR_ERR_ARG, "npair out of range (1..1024)");
  const Dims d = mkdims(dims);
  THR_CHECK(d.ls <= 64, THR_ERR_ARG, "batched derivs needs Ls <= 64");
  std::vector<DerivPair> ph(npair);
  for (int k = 0; k < npair; k++) {
    THR_CHECK(r[k] && x2[k], THR_ERR_ARG, "NULL field pointer");
    THR_CHECK(iflag[k] == 0 || iflag[k] == 1, THR_ERR_ARG, "iflag must be 0 or 1");
    ph[k] = derivs_pair((const double2 *)r[k], (const double2 *)x2[k], anum[k], iflag[k]);
  }
  DerivPair *pd = (DerivPair *)workspace(456, sizeof(DerivPair) * npair);
  double *part = (double *)workspace(457, sizeof(double) * 3 * d.v4() * npair);
  THR_CHECK(pd && part, THR_ERR_HIP, "workspace allocation failed");
  hipStream_t s = pick(stream);
  // launch_derivs_batch waits for the pair copy, so ph may go out of scope after it
  return done_if_default(stream,
                         launch_derivs_batch(d, npair, ph.data(), pd, part, (double *)dsdpi, s));
}

int thr_multishift_dev(const void *links64, const int dims[4], float am, int imass, float am3,
                       int ndiagq, const double *aden, const double *anum, void *out,
                       const void *input, float res, int maxcg, int use_sp, int kernel,
                       int *cg_return, int *cg_returns, void *stream) {
  TRY(check_dims(dims));
  TRY(check_op(kernel, imass));
  THR_CHECK(links64 && aden && anum && out && input && cg_return, THR_ERR_ARG, "NULL argument");
  THR_CHECK(ndiagq >= 1 && ndiagq <= MAXSHIFT, THR_ERR_ARG, "ndiagq out of range");
  const OpDesc op = mkop(mkdims(dims), kernel, imass, am, am3, (const double2 *)links64);
  if (use_sp)
    return multishift_reference_sp(op, ndiagq, aden, anum, (const double2 *)input, (double2 *)out,
                                   res, maxcg, cg_return, cg_returns, pick(stream));
  return multishift_reference(op, ndiagq, aden, anum, (const double2 *)input, (double2 *)out, res,
                              maxcg, cg_return, cg_returns, pick(stream));
}

int thr_multishift_plain_dev(const void *links64, const int dims[4], float am, int imass,
                             float am3, int ndiagq, const double *aden, const void *input,
                             void *out, float tol, int maxcg, int *cg_return, int *cg_returns,
                             void *stream) {
  TRY(check_dims(dims));
  TRY(check_op(0, imass));
  THR_CHECK(links64 && aden && out && input && cg_return, THR_ERR_ARG, "NULL argument");
  THR_CHECK(ndiagq >= 1 && ndiagq <= MAXSHIFT, THR_ERR_ARG, "ndiagq out of range");
  std::vector<double> ones(ndiagq, 1.0);
  const OpDesc op = mkop(mkdims(dims), 0, imass, am, am3, (const double2 *)links64);
  return multishift_reference(op, ndiagq, aden, ones.data(), (const double2 *)input,
                              (double2 *)out, tol, maxcg, cg_return, cg_returns, pick(stream), 1);
}

int thr_multishift_fast_dev(const void *links64, const void *links32, const int dims[4],
                            int nsrc, const float *am, int imass, float am3, int nshift,
                            const double *aden, const void *b, void *x, double tol, int maxcg,
                            double th32, double th16, int max_corrections, int *iterations,
                            int *cg_return, int *corr_iterations, double *relres,
                            int *switches, void *stream) {
  TRY(check_dims(dims));
  TRY(check_op(0, imass));
  THR_CHECK(links64 && links32 && am && aden && b && x, THR_ERR_ARG, "NULL argument");
  THR_CHECK(max_corrections >= 0 && max_corrections <= 16, THR_ERR_ARG,
            "max_corrections out of range (0..16)");
  MsfArgs A{};
  A.d = mkdims(dims);
  A.nsrc = nsrc;
  A.nshift = nshift;
  A.aden = aden;
  A.am = am;
  A.imass = imass;
  A.am3 = am3;
  A.u64 = (const double2 *)links64;
  A.u32 = (const float2 *)links32;
  A.b = (const double2 *)b;
  A.x = (double2 *)x;
  // iterate each shift to 3/4 of the requested true residual, which leaves an
  // FP64 solve below tol without corrections (its true residual exceeds the
  // iterated one by rounding only)
  A.tol = 0.75 * tol;
  A.maxcg = maxcg;
  A.th32 = th32;
  A.th16 = th16;
  MsfResult R{};
  hipStream_t s = pick(stream);
  TRY(msfast_solve(A, &R, s));
  if (iterations) *iterations = R.iterations;
  if (cg_return)
    for (int k = 0; k < nsrc; k++) cg_return[k] = R.cg_return[k];
  if (switches)
    for (int k = 0; k < 3; k++) switches[k] = R.switches[k];
  if (corr_iterations) *corr_iterations = 0;
  THR_CHECK(!R.nonfinite, THR_ERR_NONFINITE,
            "multishift: non-finite dot product (NaN or Inf in a source or the links)");
  if (max_corrections > 0 || relres) {
    std::vector<double> rr((size_t)nsrc * nshift);
    if (max_corrections > 0)
      TRY(msfast_correct(A, max_corrections, tol, rr.data(), corr_iterations, s));
    else
      TRY(msfast_true_residuals(A, rr.data(), s));
    if (relres) std::copy(rr.begin(), rr.end(), relres);
    THR_HIP(hipStreamSynchronize(s));
    // the contract is the FP64 true residual: report any pair above tol
    // (NaN included) after the correction rounds, as the reference aborts
    // when its solver hits the iteration cap (dwf3d_lib.F90:860-881)
    int bad = 0;
    double worst = 0.0;
    for (double v : rr) {
      if (!(v <= tol)) bad++;
      if (!(v <= worst)) worst = v;
    }
    THR_CHECK(bad == 0, THR_ERR_NOCONV,
              "multishift: " + std::to_string(bad) + " (source, shift) pairs above the true-residual "
              "target (worst " + std::to_string(worst) + ")");
    return THR_OK;
  }
  THR_HIP(hipStreamSynchronize(s));
  THR_CHECK(R.unconverged == 0, THR_ERR_NOCONV,
            "multishift: " + std::to_string(R.unconverged) +
                " (source, shift) pairs still above the iterated target at maxcg");
  return THR_OK;
}

int thr_true_residuals_dev(const void *links64, const int dims[4], int nsrc, const float *am,
                           int imass, float am3, int nshift, const double *aden, const void *b,
                           const void *x, double *relres, void *stream) {
  TRY(check_dims(dims));
  TRY(check_op(0, imass));
  THR_CHECK(links64 && am && aden && b && x && relres, THR_ERR_ARG, "NULL argument");
  THR_CHECK(nsrc >= 1 && nsrc <= MAXRHS && nshift >= 1 && nshift <= MAXSHIFT, THR_ERR_ARG,
            "nsrc or nshift out of range");
  MsfArgs A{};
  A.d = mkdims(dims);
  A.nsrc = nsrc;
  A.nshift = nshift;
  A.aden = aden;
  A.am = am;
  A.imass = imass;
  A.am3 = am3;
  A.u64 = (const double2 *)links64;
  A.b = (const double2 *)b;
  A.x = (double2 *)x;
  return msfast_true_residuals(A, relres, pick(stream));
}

}  // extern "C"
