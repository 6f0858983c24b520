// Reference-faithful Krylov solvers with device-resident scalars.
//
// The reference runs every dot product through an MPI_AllReduce and a host
// round trip.  Here all scalars (alpha, beta, zeta recurrences, convergence
// window) live in a small device state struct written by single-block
// "finalize" kernels that sum the per-block partials in a fixed order, so
// results are deterministic and the host only polls a stop flag every few
// iterations.  Kernels of an iteration that starts after the stop flag was
// raised exit immediately, so the iteration counts and final vectors are
// exactly those of the reference's loop structure.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "solvers.h"

namespace thr {

// ---------------------------------------------------------------------------
// workspace: one set of device buffers and one stream per context; a thread
// works in the context it bound (thr_context_bind), else in the default one,
// so independent solves on different threads never share scratch memory
// ---------------------------------------------------------------------------
namespace {
struct Buf {
  void *p = nullptr;
  size_t n = 0;
};
}  // namespace

struct Context {
  std::map<int, Buf> ws;
  std::map<int, Buf> hws;  // pinned host buffers for small state transfers
  std::mutex mu;
  hipStream_t stream = nullptr;
  // hipGraph capture of solver blocks; off in created contexts, which run
  // next to other threads' allocations and copies (a stream capture is
  // invalidated by device-synchronising calls from any thread)
  bool graphs = true;
  // reliable updates of the mixed multishift's base residual: taken each time
  // the residual falls by this factor while the Krylov storage is FP32/FP16
  // (0: never); and how many the last solve in this context took
  double reliable = 0.1;
  int last_reliable = 0;
};

namespace {
Context g_default;
thread_local Context *t_ctx = nullptr;
Context &cur() { return t_ctx ? *t_ctx : g_default; }
}  // namespace

Context *context_create() {
  Context *c = new Context();
  c->graphs = false;
  return c;
}

bool context_graphs() { return cur().graphs; }
double context_reliable() { return cur().reliable; }
void context_set_reliable(double delta) { cur().reliable = delta; }
void context_note_reliable(int n) { cur().last_reliable = n; }
int context_last_reliable() { return cur().last_reliable; }

void context_destroy(Context *c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  for (auto &kv : c->ws)
    if (kv.second.p) (void)hipFree(kv.second.p);
  for (auto &kv : c->hws)
    if (kv.second.p) (void)hipHostFree(kv.second.p);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (t_ctx == c) t_ctx = nullptr;
  delete c;
}

void context_bind(Context *c) { t_ctx = c; }

void *workspace(int slot, size_t bytes) {
  Context &c = cur();
  std::lock_guard<std::mutex> lk(c.mu);
  Buf &b = c.ws[slot];
  if (b.n < bytes) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.n = 0;
    if (hipMalloc(&b.p, bytes) != hipSuccess) {
      set_error("hipMalloc of workspace failed");
      b.p = nullptr;
      return nullptr;
    }
    b.n = bytes;
  }
  return b.p;
}

void *host_workspace(int slot, size_t bytes) {
  Context &c = cur();
  std::lock_guard<std::mutex> lk(c.mu);
  Buf &b = c.hws[slot];
  if (b.n < bytes) {
    if (b.p) (void)hipHostFree(b.p);
    b.p = nullptr;
    b.n = 0;
    if (hipHostMalloc(&b.p, bytes, hipHostMallocDefault) != hipSuccess) {
      set_error("hipHostMalloc of host workspace failed");
      b.p = nullptr;
      return nullptr;
    }
    b.n = bytes;
  }
  return b.p;
}

size_t workspace_bytes() {
  Context &c = cur();
  std::lock_guard<std::mutex> lk(c.mu);
  size_t t = 0;
  for (auto &kv : c.ws) t += kv.second.n;
  return t;
}

double *partial_buffer(int nblocks) {
  return (double *)workspace(0, sizeof(double) * (size_t)nblocks);
}

hipStream_t default_stream() {
  Context &c = cur();
  // non-blocking: no implicit ordering with the legacy null stream, which
  // would break graph capture while another thread copies on it
  if (!c.stream) (void)hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking);
  return c.stream;
}

void release_workspace() {
  Context &c = cur();
  std::lock_guard<std::mutex> lk(c.mu);
  for (auto &kv : c.ws)
    if (kv.second.p) (void)hipFree(kv.second.p);
  c.ws.clear();
}

// ---------------------------------------------------------------------------
// streaming kernels (one thread per 5-D site, 4 spins; same grid as the Dslash
// so that partial-sum buffers line up)
// ---------------------------------------------------------------------------
namespace {

__device__ inline double fin_sum(const double *partial, int n) { return partial_sum(partial, n); }

template <int P> using TT = typename Store<P>::T;
template <int P> using CC = typename Store<P>::C;

template <int P> __device__ inline cx<CC<P>> L1(const TT<P> *p, long i) { return ldc(p, i); }

// r += a*s ; partial |r|^2
template <int P>
__global__ __launch_bounds__(BLK) void k_axpy_norm(long v5, TT<P> *r, const TT<P> *s,
                                                   const double *a, const int *stop,
                                                   double *partial) {
  if (stop && *stop) return;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  const CC<P> av = (CC<P>)(*a);
  double red = 0.0;
  if (i < v5) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      cx<CC<P>> rv = L1<P>(r, k * v5 + i) + av * L1<P>(s, k * v5 + i);
      st(r, k * v5 + i, rv);
      red += (double)norm2(rv);
    }
  }
  const double t = block_sum(red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// p = r + b*p
template <int P>
__global__ __launch_bounds__(BLK) void k_xpay(long n, const TT<P> *r, TT<P> *p, const double *b,
                                              const int *stop) {
  if (stop && *stop) return;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const CC<P> bv = (CC<P>)(*b);
  st(p, i, L1<P>(r, i) + bv * L1<P>(p, i));
}

// partial |x|^2 (per-element products in the storage precision, summed in FP64)
template <int P>
__global__ __launch_bounds__(BLK) void k_norm(long v5, const TT<P> *x, double *partial) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  double red = 0.0;
  if (i < v5) {
#pragma unroll
    for (int k = 0; k < 4; k++) red += (double)norm2(L1<P>(x, k * v5 + i));
  }
  const double t = block_sum(red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// epilogues for operators that cannot fuse them (Wilson kernel)
__global__ __launch_bounds__(BLK) void k_epi(long v5, int mode, const double2 *o, const double2 *w,
                                             double2 *r, double2 *x, const double2 *p,
                                             const double *a, const double *b, const int *stop,
                                             double *partial) {
  if (stop && *stop) return;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  double red = 0.0;
  if (i < v5) {
    for (int k = 0; k < 4; k++) {
      const long e = k * v5 + i;
      const cx<double> ov = ldc(o, e);
      if (mode == EPI_NORM) {
        red += norm2(ov);
      } else if (mode == EPI_DOTW) {
        const cx<double> wv = ldc(w, e);
        red += wv.r * ov.r + wv.i * ov.i;
      } else if (mode == EPI_CG) {
        const cx<double> rv = ldc(r, e) - (*a) * ov;
        st(r, e, rv);
        red += norm2(rv);
        if (b) st(x, e, ldc(x, e) + (*b) * ldc(p, e));
      }
    }
  }
  const double t = block_sum(red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// ---------------------------------------------------------------------------
// CG state (measure_module.F90:41-120)
// ---------------------------------------------------------------------------
struct CgState {
  double alpha, alpha_x, alphan, betacgd, betacgn, betacg;
  float resid;
  int niterc, nx, itercg, stop, iterations;
};

__global__ void k_cg_init(CgState *st, float resid, int niterc) {
  st->alpha = 1.0;
  st->alpha_x = 0.0;
  st->alphan = 0.0;
  st->betacgd = 1.0;
  st->betacgn = 0.0;
  st->betacg = 0.0;
  st->resid = resid;
  st->niterc = niterc;
  st->nx = 0;
  st->itercg = 0;
  st->stop = niterc < 1 ? 1 : 0;
  st->iterations = niterc < 1 ? 1 : 0;
}

__global__ __launch_bounds__(BLK) void k_cg_alpha(CgState *st, const double *partial, int nb) {
  if (st->stop) return;
  const double alphad = fin_sum(partial, nb);
  if (threadIdx.x == 0) {
    st->nx += 1;
    st->itercg += 1;
    if (st->nx != 1) {
      st->alpha = st->alphan / alphad;
      st->alpha_x = st->alpha;
    } else {
      st->alpha_x = 0.0;
    }
    // a non-finite |M p|^2 (NaN or Inf in the source, the guess or the links)
    // can never meet the stopping rule: end the solve here instead of at niterc
    if (!isfinite(alphad) || !isfinite(st->alpha)) {
      st->stop = 3;
      st->iterations = st->nx;
    }
  }
}

__global__ __launch_bounds__(BLK) void k_cg_beta(CgState *st, const double *partial, int nb) {
  if (st->stop) return;
  const double bn = fin_sum(partial, nb);
  if (threadIdx.x == 0) {
    st->betacgn = bn;
    st->betacg = bn / st->betacgd;
    st->betacgd = bn;
    st->alphan = bn;
    if (st->nx == 1) st->betacg = 0.0;
    if (!isfinite(bn)) {
      st->stop = 3;
      st->iterations = st->nx;
    } else if (bn < (double)st->resid) {
      st->stop = 1;
      st->iterations = st->nx;
    } else if (st->nx >= st->niterc) {
      // p = r + beta p still runs; the Fortran loop then ends with nx = niterc + 1
      st->iterations = st->niterc + 1;
      st->stop = 2;
    }
  }
}

// the p update of the final capped iteration must still run: stop == 2 lets
// it (stop 1: converged, 3: non-finite scalars)
__global__ __launch_bounds__(BLK) void k_cg_p(long n, const double2 *r, double2 *p,
                                              const CgState *S) {
  if (S->stop == 1 || S->stop == 3) return;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  st(p, i, ldc(r, i) + S->betacg * ldc(p, i));
}

// ---------------------------------------------------------------------------
// multishift state (multishift_module.F90:71-282)
// ---------------------------------------------------------------------------
struct MsState {
  double delta, alpha, lambda, omega, omega_save, gammag;
  double zeta_i[MAXSHIFT], zeta_ii[MAXSHIFT], zeta_iii[MAXSHIFT], omegas[MAXSHIFT],
      gammas[MAXSHIFT], corr[MAXSHIFT], aden[MAXSHIFT];
  int flags[MAXSHIFT], cg_returns[MAXSHIFT];
  int n, mini, maxi, umin, umax, cg_return, maxcg, stop, done, plain;
  int nonfinite;  // a dot product came out NaN/Inf: the solve stopped there
  float res;
  double resd;  // threshold used by the convergence test
};

__global__ __launch_bounds__(BLK) void k_ms_init(MsState *st, const double *partial, int nb) {
  const double d = fin_sum(partial, nb);
  if (threadIdx.x == 0) {
    st->delta = d;
    st->omega = 1.0;
    st->alpha = 0.0;
    st->gammag = 0.0;
    for (int i = 0; i < st->n; i++) {
      st->flags[i] = 1;
      st->zeta_i[i] = 1.0;
      st->zeta_ii[i] = 1.0;
      st->gammas[i] = 0.0;
      st->cg_returns[i] = 0;
    }
    st->cg_return = 0;
    st->maxi = st->n;
    st->mini = 1;
    st->umin = 1;
    st->umax = 0;
    st->nonfinite = isfinite(d) ? 0 : 1;
    st->stop = (st->maxcg <= 0 || st->n <= 0 || st->nonfinite) ? 1 : 0;
    st->done = st->stop;
    // reference rule: sqrt(delta zeta^2) corr_i < res; plain rule (corr_i = 1):
    // shifted residual below res |b|
    st->resd = st->plain ? (double)st->res * sqrt(d) : (double)st->res;
  }
}

__global__ void k_ms_alpha(MsState *st, const double *partial, int nb) {
  if (st->stop) {
    if (threadIdx.x == 0) st->done = 1;
    return;
  }
  const double al = fin_sum(partial, nb);
  if (threadIdx.x == 0) {
    st->cg_return += 1;
    st->alpha = al;
    st->omega_save = st->omega;
    st->omega = -st->delta / al;
    if (!isfinite(st->omega)) {
      // nothing downstream of a non-finite omega is meaningful: stop before
      // any vector of this iteration is updated
      st->nonfinite = 1;
      st->stop = 1;
      st->done = 1;
      return;
    }
    const double om = st->omega, os = st->omega_save, gg = st->gammag;
    for (int i = st->mini - 1; i < st->maxi; i++) {
      st->zeta_iii[i] = (st->zeta_i[i] * st->zeta_ii[i] * os) /
                        (om * gg * (st->zeta_i[i] - st->zeta_ii[i]) +
                         st->zeta_i[i] * os * (1.0 - st->aden[i] * om));
      st->omegas[i] = om * st->zeta_iii[i] / st->zeta_ii[i];
    }
  }
}

__global__ void k_ms_lambda(MsState *st, const double *partial, int nb) {
  if (st->stop) return;
  const double lam = fin_sum(partial, nb);
  if (threadIdx.x == 0) {
    if (!isfinite(lam)) {
      st->nonfinite = 1;
      st->stop = 1;
      st->done = 1;
      return;
    }
    st->lambda = lam;
    st->gammag = lam / st->delta;
    for (int i = st->mini - 1; i < st->maxi; i++) {
      st->gammas[i] = st->gammag * st->zeta_iii[i] * st->omegas[i] / (st->zeta_ii[i] * st->omega);
      st->cg_returns[i] = st->cg_return;
    }
    st->umin = st->mini;
    st->umax = st->maxi;
    int nmax = 0, nmin = st->n + 1;
    for (int i = 0; i < st->n; i++) {
      if (st->flags[i]) {
        if (sqrt(st->delta * st->zeta_ii[i] * st->zeta_ii[i]) * st->corr[i] < st->resd) {
          st->flags[i] = 0;
        } else {
          nmax = max(nmax, i + 1);
          nmin = min(nmin, i + 1);
        }
      }
    }
    st->maxi = nmax;
    st->mini = nmin;
    for (int i = nmin - 1; i < nmax; i++) {
      st->zeta_i[i] = st->zeta_ii[i];
      st->zeta_ii[i] = st->zeta_iii[i];
    }
    st->delta = lam;
    if (!(st->maxi > st->mini - 1 && st->cg_return < st->maxcg)) st->stop = 1;
  }
}

// p = r + gammag p ; x_k -= omegas_k p_k ; p_k = gammas_k p_k + zeta_iii_k r
// (multishift_module.F90:145, 155-176), all active shifts in one pass over r
template <int P>
__global__ __launch_bounds__(BLK) void k_ms_update(long n, const TT<P> *r, TT<P> *p, TT<P> *xs,
                                                   TT<P> *ps, const MsState *S) {
  if (S->done) return;
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  using C = CC<P>;
  const cx<C> rv = L1<P>(r, i);
  st(p, i, rv + (C)S->gammag * L1<P>(p, i));
  for (int k = S->umin - 1; k < S->umax; k++) {
    TT<P> *x = xs + (long)k * n;
    TT<P> *q = ps + (long)k * n;
    const cx<C> pk = L1<P>(q, i);
    st(x, i, L1<P>(x, i) - (C)S->omegas[k] * pk);
    st(q, i, (C)S->gammas[k] * pk + (C)S->zeta_iii[k] * rv);
  }
}

template <int P>
__global__ __launch_bounds__(BLK) void k_bcast_copy(long n, int nshift, const TT<P> *src, TT<P> *dst) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i >= n) return;
  const auto v = src[i];
  for (int k = 0; k < nshift; k++) dst[(long)k * n + i] = v;
}

// element-wise FP64 -> FP32 complex conversion (link fields)
__global__ __launch_bounds__(BLK) void k_d2f(long n, const double2 *a, float2 *b) {
  const long i = (long)blockIdx.x * BLK + threadIdx.x;
  if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

inline unsigned nblk(long n) { return (unsigned)((n + BLK - 1) / BLK); }

// streaming copy, 16 bytes per lane, non-temporal loads and stores, a
// grid-stride loop over a few blocks per CU: the fastest plain read+write
// stream measured on gfx950 (the HBM ceiling the Dslash is compared with)
typedef __attribute__((ext_vector_type(4))) unsigned U4;
__global__ __launch_bounds__(BLK) void k_stream_copy(const U4 *__restrict__ a, U4 *__restrict__ b,
                                                     long n) {
  const long stride = (long)gridDim.x * BLK;
  for (long i = (long)blockIdx.x * BLK + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

}  // namespace

int stream_copy(const void *src, void *dst, size_t bytes, hipStream_t s) {
  THR_CHECK(bytes % 16 == 0, THR_ERR_ARG, "byte count must be a multiple of 16");
  THR_CHECK(((uintptr_t)src | (uintptr_t)dst) % 16 == 0, THR_ERR_ARG, "pointers must be 16-byte aligned");
  if (bytes == 0) return THR_OK;
  int dev = 0, cus = 256;
  THR_HIP(hipGetDevice(&dev));
  THR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const long n = (long)(bytes / 16);
  const long want = (long)cus * 4;
  const unsigned grid = (unsigned)std::max(1L, std::min(want, (long)nblk(n)));
  hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(BLK), 0, s, (const U4 *)src, (U4 *)dst, n);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int convert_d2f(long n, const double2 *a, float2 *b, hipStream_t s) {
  hipLaunchKernelGGL(k_d2f, dim3(nblk(n)), dim3(BLK), 0, s, n, a, b);
  THR_HIP(hipGetLastError());
  return THR_OK;
}

int apply_op(const OpDesc &op, int dagger, const double2 *in, double2 *out, hipStream_t s) {
  if (op.kernel == 1) {
    double2 *tmp = (double2 *)workspace(1, sizeof(double2) * op.d.nspinor());
    THR_CHECK(tmp, THR_ERR_HIP, "workspace allocation failed");
    return launch_dslash_wilson(op.d, 1, dagger, op.imass, op.am, op.am3, in, out, op.u64, tmp, s);
  }
  DslashLaunch L{};
  L.d = op.d;
  L.nrhs = 1;
  L.prec = P64;
  L.dagger = dagger;
  L.imass = op.imass;
  L.am3 = op.am3;
  L.am[0] = op.am;
  L.in = in;
  L.out = out;
  L.links = op.u64;
  L.mode = EPI_STORE;
  L.sc_stride = 1;
  L.done_stride = 1;
  return launch_dslash(L, s);
}

// M or M^+ with a fused (Shamir) or separate (Wilson) epilogue, FP64
static int apply_op_epi(const OpDesc &op, int dagger, const double2 *in, double2 *out, int mode,
                        const double2 *w, double2 *r, double2 *x, const double2 *p,
                        const double *a, const double *b, double *partial, const int *stop,
                        hipStream_t s) {
  if (op.kernel == 1) {
    // skip when stopped: the Wilson path is unfused, so gate the whole sequence
    double2 *o = out ? out : (double2 *)workspace(2, sizeof(double2) * op.d.nspinor());
    double2 *tmp = (double2 *)workspace(1, sizeof(double2) * op.d.nspinor());
    THR_CHECK(o && tmp, THR_ERR_HIP, "workspace allocation failed");
    int rc = launch_dslash_wilson(op.d, 1, dagger, op.imass, op.am, op.am3, in, o, op.u64, tmp, s);
    if (rc) return rc;
    if (mode != EPI_STORE) {
      hipLaunchKernelGGL(k_epi, dim3(nblk(op.d.v5())), dim3(BLK), 0, s, op.d.v5(), mode, o, w, r,
                         x, p, a, b, stop, partial);
      THR_HIP(hipGetLastError());
    }
    return THR_OK;
  }
  DslashLaunch L{};
  L.d = op.d;
  L.nrhs = 1;
  L.prec = P64;
  L.dagger = dagger;
  L.imass = op.imass;
  L.am3 = op.am3;
  L.am[0] = op.am;
  L.in = in;
  L.out = out;
  L.links = op.u64;
  L.mode = mode;
  L.w = w;
  L.r = r;
  L.x = x;
  L.p = p;
  L.a_ptr = a;
  L.b_ptr = b;
  L.sc_stride = 1;
  L.partial = partial;
  L.done = stop;
  L.done_stride = 1;
  return launch_dslash(L, s);
}

// poll a device int every `batch` iterations; batches grow to bound host syncs.
// host_flag: where the flag is read back (NULL: a pinned int allocated here)
template <typename F>
static int run_batched(F &&one_iteration, const int *dev_flag, hipStream_t s, long max_iters,
                       int *host_flag = nullptr) {
  int *h = host_flag;
  if (!h) THR_HIP(hipHostMalloc((void **)&h, sizeof(int), hipHostMallocDefault));
  *h = 0;
  int batch = 4;
  long done_iters = 0;
  int rc = THR_OK;
  for (;;) {
    for (int k = 0; k < batch; k++) {
      rc = one_iteration();
      if (rc) break;
    }
    if (rc) break;
    done_iters += batch;
    if (hipMemcpyAsync(h, dev_flag, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess) {
      set_error("stream synchronisation failed in solver loop");
      rc = THR_ERR_HIP;
      break;
    }
    if (*h) break;
    if (done_iters > max_iters + 8) {
      set_error("solver did not raise its stop flag");
      rc = THR_ERR_STATE;
      break;
    }
    if (batch < 64) batch *= 2;
  }
  if (!host_flag) (void)hipHostFree(h);
  return rc;
}

size_t cg_state_bytes() { return sizeof(CgState); }

int cg_reference(const OpDesc &op, const double2 *phi, double2 *x, float res, int niterc,
                 int *itercg, int *iterations, hipStream_t s) {
  const Dims &d = op.d;
  const long n = d.nspinor();
  CgBuffers B{};
  B.x1 = (double2 *)workspace(10, sizeof(double2) * n);
  B.r = (double2 *)workspace(11, sizeof(double2) * n);
  B.p = (double2 *)workspace(12, sizeof(double2) * n);
  B.state = workspace(13, sizeof(CgState));
  B.partial = partial_buffer((int)dslash_blocks(d));
  THR_CHECK(B.x1 && B.r && B.p && B.state && B.partial, THR_ERR_HIP,
            "workspace allocation failed");
  return cg_reference_bufs(op, phi, x, res, niterc, itercg, iterations, s, B, nullptr);
}

int cg_reference_bufs(const OpDesc &op, const double2 *phi, double2 *x, float res, int niterc,
                      int *itercg, int *iterations, hipStream_t s, const CgBuffers &B,
                      int *host_flag) {
  const Dims &d = op.d;
  const long n = d.nspinor();
  const int nb = (int)dslash_blocks(d);
  double2 *x1 = B.x1, *r = B.r, *p = B.p;
  CgState *st = (CgState *)B.state;
  double *partial = B.partial;
  THR_CHECK(x1 && r && p && st && partial, THR_ERR_ARG, "congrad: missing buffers");
  THR_CHECK(op.kernel == 0 || !host_flag, THR_ERR_ARG,
            "congrad on caller buffers runs the Shamir kernel only");
  // resid = 4*ksize*ksize*ksizet*kthird*res*res in default real (measure_module.F90:41)
  const long nsite4 = 4L * d.lx * d.ly * d.lt * d.ls;
  const float resid = ((float)nsite4 * res) * res;
  THR_HIP(hipMemcpyAsync(p, x, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
  THR_HIP(hipMemcpyAsync(r, phi, sizeof(double2) * n, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_cg_init, dim3(1), dim3(1), 0, s, st, resid, niterc);
  const int *stop = &st->stop;
  auto iter = [&]() -> int {
    int rc = apply_op_epi(op, 0, p, x1, EPI_NORM, nullptr, nullptr, nullptr, nullptr, nullptr,
                          nullptr, partial, stop, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(BLK), 0, s, st, partial, nb);
    rc = apply_op_epi(op, 1, x1, nullptr, EPI_CG, nullptr, r, x, p, &st->alpha, &st->alpha_x,
                      partial, stop, s);
    if (rc) return rc;
    hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(BLK), 0, s, st, partial, nb);
    hipLaunchKernelGGL(k_cg_p, dim3(nblk(n)), dim3(BLK), 0, s, n, r, p, st);
    THR_HIP(hipGetLastError());
    return THR_OK;
  };
  int rc = run_batched(iter, stop, s, niterc, host_flag);
  if (rc) return rc;
  CgState h;
  THR_HIP(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  *itercg = h.itercg;
  if (iterations) *iterations = h.iterations;
  THR_CHECK(h.stop != 3, THR_ERR_NONFINITE,
            "congrad: non-finite dot product (NaN or Inf in the source, guess or links)");
  return THR_OK;
}

static int ms_setup(MsState *st, int n, const double *aden, const double *anum, float res,
                    int maxcg, hipStream_t s, int plain = 0) {
  THR_CHECK(n >= 1 && n <= MAXSHIFT, THR_ERR_ARG, "number of shifts out of range");
  MsState h;
  std::memset(&h, 0, sizeof(h));
  h.n = n;
  h.res = res;
  h.maxcg = maxcg;
  // correction = |anum/aden| / geometric mean (multishift_module.F90:87-88)
  double lsum = 0.0;
  for (int i = 0; i < n; i++) {
    h.corr[i] = std::fabs(anum[i] / aden[i]);
    lsum += std::log(h.corr[i]);
  }
  const double gm = std::exp(lsum / n);
  for (int i = 0; i < n; i++) {
    h.corr[i] = plain ? 1.0 : h.corr[i] / gm;
    h.aden[i] = aden[i];
  }
  h.plain = plain;
  THR_HIP(hipMemcpyAsync(st, &h, sizeof(h), hipMemcpyHostToDevice, s));
  THR_HIP(hipStreamSynchronize(s));  // h is a stack object
  return THR_OK;
}

int multishift_reference(const OpDesc &op, int n, const double *aden, const double *anum,
                         const double2 *input, double2 *out, float res, int maxcg,
                         int *cg_return, int *cg_returns, hipStream_t s, int plain) {
  const Dims &d = op.d;
  const long ns = d.nspinor();
  const int nb = (int)dslash_blocks(d);
  double2 *r = (double2 *)workspace(20, sizeof(double2) * ns);
  double2 *h = (double2 *)workspace(21, sizeof(double2) * ns);
  double2 *sv = (double2 *)workspace(22, sizeof(double2) * ns);
  double2 *p = (double2 *)workspace(23, sizeof(double2) * ns);
  double2 *ps = (double2 *)workspace(24, sizeof(double2) * ns * n);
  MsState *st = (MsState *)workspace(25, sizeof(MsState));
  double *partial = partial_buffer(nb);
  THR_CHECK(r && h && sv && p && ps && st && partial, THR_ERR_HIP, "workspace allocation failed");
  int rc = ms_setup(st, n, aden, anum, res, maxcg, s, plain);
  if (rc) return rc;
  THR_HIP(hipMemcpyAsync(r, input, sizeof(double2) * ns, hipMemcpyDeviceToDevice, s));
  THR_HIP(hipMemcpyAsync(p, input, sizeof(double2) * ns, hipMemcpyDeviceToDevice, s));
  THR_HIP(hipMemsetAsync(out, 0, sizeof(double2) * ns * n, s));
  hipLaunchKernelGGL(k_bcast_copy<P64>, dim3(nblk(ns)), dim3(BLK), 0, s, ns, n, input, ps);
  hipLaunchKernelGGL(k_norm<P64>, dim3(nb), dim3(BLK), 0, s, d.v5(), r, partial);
  hipLaunchKernelGGL(k_ms_init, dim3(1), dim3(BLK), 0, s, st, partial, nb);
  THR_HIP(hipGetLastError());
  const int *stop = &st->stop;
  auto iter = [&]() -> int {
    int rc2 = apply_op_epi(op, 0, p, h, EPI_STORE, nullptr, nullptr, nullptr, nullptr, nullptr,
                           nullptr, nullptr, stop, s);
    if (rc2) return rc2;
    rc2 = apply_op_epi(op, 1, h, sv, EPI_DOTW, p, nullptr, nullptr, nullptr, nullptr, nullptr,
                       partial, stop, s);
    if (rc2) return rc2;
    hipLaunchKernelGGL(k_ms_alpha, dim3(1), dim3(BLK), 0, s, st, partial, nb);
    hipLaunchKernelGGL(k_axpy_norm<P64>, dim3(nb), dim3(BLK), 0, s, d.v5(), r, sv, &st->omega,
                       stop, partial);
    hipLaunchKernelGGL(k_ms_lambda, dim3(1), dim3(BLK), 0, s, st, partial, nb);
    hipLaunchKernelGGL(k_ms_update<P64>, dim3(nblk(ns)), dim3(BLK), 0, s, ns, r, p, out, ps, st);
    THR_HIP(hipGetLastError());
    return THR_OK;
  };
  rc = run_batched(iter, stop, s, maxcg);
  if (rc) return rc;
  MsState hs;
  THR_HIP(hipMemcpyAsync(&hs, st, sizeof(MsState), hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  *cg_return = hs.cg_return;
  if (cg_returns)
    for (int i = 0; i < n; i++) cg_returns[i] = hs.cg_returns[i];
  THR_CHECK(!hs.nonfinite, THR_ERR_NONFINITE,
            "multishift: non-finite dot product (NaN or Inf in the source or links)");
  return THR_OK;
}

int multishift_reference_sp(const OpDesc &op, int n, const double *aden, const double *anum,
                            const double2 *input, double2 *out, float res, int maxcg,
                            int *cg_return, int *cg_returns, hipStream_t s) {
  THR_CHECK(op.kernel == 0, THR_ERR_ARG, "the FP32 multishift is Shamir-only (qmrherm_module.F90:54-63)");
  const Dims &d = op.d;
  const long ns = d.nspinor();
  const int nb = (int)dslash_blocks(d);
  float2 *r = (float2 *)workspace(30, sizeof(float2) * ns);
  float2 *h = (float2 *)workspace(31, sizeof(float2) * ns);
  float2 *sv = (float2 *)workspace(32, sizeof(float2) * ns);
  float2 *p = (float2 *)workspace(33, sizeof(float2) * ns);
  float2 *ps = (float2 *)workspace(34, sizeof(float2) * ns * n);
  float2 *xs = (float2 *)workspace(35, sizeof(float2) * ns * n);
  float2 *u32 = (float2 *)workspace(36, sizeof(float2) * d.v4() * 3);
  MsState *st = (MsState *)workspace(37, sizeof(MsState));
  double *partial = partial_buffer(nb);
  THR_CHECK(r && h && sv && p && ps && xs && u32 && st && partial, THR_ERR_HIP,
            "workspace allocation failed");
  int rc = ms_setup(st, n, aden, anum, res, maxcg, s);
  if (rc) return rc;
  // u = cmplx(udp), input = cmplx(inputdp) (multishift_module.F90:351-352)
  hipLaunchKernelGGL(k_d2f, dim3(nblk(d.v4() * 3)), dim3(BLK), 0, s, d.v4() * 3, op.u64, u32);
  // FP32 spinors keep a site's spins together: the conversion reorders too
  TRY_RC(convert_spinors(d, 1, input, nullptr, P64, r, nullptr, P32, s));
  THR_HIP(hipMemcpyAsync(p, r, sizeof(float2) * ns, hipMemcpyDeviceToDevice, s));
  THR_HIP(hipMemsetAsync(xs, 0, sizeof(float2) * ns * n, s));
  hipLaunchKernelGGL(k_bcast_copy<P32>, dim3(nblk(ns)), dim3(BLK), 0, s, ns, n, r, ps);
  hipLaunchKernelGGL(k_norm<P32>, dim3(nb), dim3(BLK), 0, s, d.v5(), r, partial);
  hipLaunchKernelGGL(k_ms_init, dim3(1), dim3(BLK), 0, s, st, partial, nb);
  THR_HIP(hipGetLastError());
  const int *stop = &st->stop;
  DslashLaunch L{};
  L.d = d;
  L.nrhs = 1;
  L.prec = P32;
  L.imass = op.imass;
  L.am3 = op.am3;
  L.am[0] = op.am;
  L.links = u32;
  L.sc_stride = 1;
  L.done = stop;
  L.done_stride = 1;
  L.partial = partial;
  auto iter = [&]() -> int {
    DslashLaunch a = L;
    a.dagger = 0;
    a.in = p;
    a.out = h;
    a.mode = EPI_STORE;
    int rc2 = launch_dslash(a, s);
    if (rc2) return rc2;
    DslashLaunch b = L;
    b.dagger = 1;
    b.in = h;
    b.out = sv;
    b.mode = EPI_DOTW;
    b.w = p;
    rc2 = launch_dslash(b, s);
    if (rc2) return rc2;
    hipLaunchKernelGGL(k_ms_alpha, dim3(1), dim3(BLK), 0, s, st, partial, nb);
    hipLaunchKernelGGL(k_axpy_norm<P32>, dim3(nb), dim3(BLK), 0, s, d.v5(), r, sv, &st->omega,
                       stop, partial);
    hipLaunchKernelGGL(k_ms_lambda, dim3(1), dim3(BLK), 0, s, st, partial, nb);
    hipLaunchKernelGGL(k_ms_update<P32>, dim3(nblk(ns)), dim3(BLK), 0, s, ns, r, p, xs, ps, st);
    THR_HIP(hipGetLastError());
    return THR_OK;
  };
  rc = run_batched(iter, stop, s, maxcg);
  if (rc) return rc;
  TRY_RC(convert_spinors(d, n, xs, nullptr, P32, out, nullptr, P64, s));
  MsState hs;
  THR_HIP(hipMemcpyAsync(&hs, st, sizeof(MsState), hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  *cg_return = hs.cg_return;
  if (cg_returns)
    for (int i = 0; i < n; i++) cg_returns[i] = hs.cg_returns[i];
  THR_CHECK(!hs.nonfinite, THR_ERR_NONFINITE,
            "multishift: non-finite dot product (NaN or Inf in the source or links)");
  return THR_OK;
}

}  // namespace thr
