// Persistent batched multishift CG for launch-bound lattices.
//
// The same recurrences as the blocked solver in msfast.hip (and so as
// multishift_module.F90:103-282), run by one cooperative kernel for the whole
// solve: per iteration an M pass (h = M r + gamma h, |h|^2), a grid barrier,
// the alpha recurrences, an M^dagger pass (r += omega M^+ h, |r|^2), a grid
// barrier, the lambda recurrences and the shift updates x_i -= omega_i p_i,
// p_i = gamma_i p_i + zeta_i r.  On a lattice of a few 10^4 sites per source
// the streaming solver spends its iteration in kernel dispatch (four launches
// of a few microseconds each); here an iteration costs two grid barriers,
// which on this GPU cost about as much (see msfast_persistent_fits), so the
// solver is an option (THR_PERSIST=1), not the default.
//
// Blocks are assigned to sources (blocks [src * bs, (src + 1) * bs) sweep
// source src), so every block reduces into its own source's partial sums and
// every block of a source evaluates that source's scalar recurrences from the
// same partials in the same order: identical values everywhere, no broadcast.
// Storage is FP64 throughout (on such lattices the solve is latency-bound, so
// lower-precision storage would buy nothing).
#include <algorithm>
#include <cstdlib>

#include "dslash_site.h"
#include "msfast.h"

namespace thr {

namespace {

constexpr int PRES = 2 + 2 * MAXSHIFT;  // per-source results: see k_msf_persistent

struct PArgs {
  Dims d;
  int nsrc, n, imass, maxcg, bs;  // bs: blocks per source
  float am3;
  float am[MAXRHS];
  const double2 *u;   // links [mu][t][y][x]
  const double2 *b;   // sources [nsrc][4 V5]
  double2 *x, *p;     // shift solutions and directions [nsrc][n][4 V5]
  double2 *r, *h;     // residual and h = M p [nsrc][4 V5]
  double *part;       // partial sums: [0][nsrc][bs] |h|^2, [1][nsrc][bs] |r|^2
  int *done;          // [nsrc] source stopped
  const double *sig;  // [nsrc][MAXSHIFT] shifts
  const double *tol;  // [nsrc] per-shift |r_i| <= tol |b|
  int *res;           // [nsrc][PRES]: cg_return, nonfinite, cg_returns[MAXSHIFT], active[MAXSHIFT]
};

__device__ inline Site<P64> site_at(const double2 *in, unsigned v5, unsigned site) {
  Site<P64> S;
#pragma unroll
  for (int a = 0; a < 4; a++) S.v[a] = ldc(in, (long)a * v5 + site);
  return S;
}

// o = (M or M^dagger) in at 5-D site i = s + Ls (x + Lx (y + Ly t)), FP64:
// the operator of the streaming Dslash kernels (dirac.F90:29-179, 303-494)
// written per site
template <bool DAG>
__device__ inline void op_site(const double2 *in, const double2 *u, const Dims &d, unsigned i,
                               double m, int imass, double diag, cx<double> o[4]) {
  const unsigned ls = d.ls, lx = d.lx, ly = d.ly, lt = d.lt;
  const unsigned v2 = lx * ly, v3 = ls * v2, v4 = v2 * lt, v5 = v3 * lt;
  const unsigned s = i % ls, c = i / ls;
  const unsigned x = c % lx, y = (c / lx) % ly, t = c / v2;
  const unsigned xf = (x + 1 == lx) ? 0 : x + 1, xb = (x == 0) ? lx - 1 : x - 1;
  const unsigned yf = (y + 1 == ly) ? 0 : y + 1, yb = (y == 0) ? ly - 1 : y - 1;
  const unsigned tf = (t + 1 == lt) ? 0 : t + 1, tb = (t == 0) ? lt - 1 : t - 1;
  const unsigned xy = x + lx * y, l4 = v2 * t;
  auto at = [&](unsigned xx, unsigned yy, unsigned tt) { return s + ls * (xx + lx * (yy + ly * tt)); };
  auto lk = [&](unsigned k) { return ldc(u, (long)k); };
  const Site<P64> cur = site_at(in, v5, i);
  centre(o, diag, cur);
  hop<0, DAG>(o, lk(xy + l4), site_at(in, v5, at(xf, y, t)), conj(lk(xb + lx * y + l4)),
              site_at(in, v5, at(xb, y, t)));
  hop<1, DAG>(o, lk(v4 + xy + l4), site_at(in, v5, at(x, yf, t)), conj(lk(v4 + x + lx * yb + l4)),
              site_at(in, v5, at(x, yb, t)));
  hop<2, DAG>(o, lk(2 * v4 + xy + l4), site_at(in, v5, at(x, y, tf)),
              conj(lk(2 * v4 + xy + v2 * tb)), site_at(in, v5, at(x, y, tb)));
  // s hops: spins {3,4} from s+1 and {1,2} from s-1 for M, the reverse for M^+
  const int up0 = DAG ? 0 : 2, dn0 = DAG ? 2 : 0;
  if (s + 1 < ls) {
    o[up0] = o[up0] - ldc(in, (long)up0 * v5 + i + 1);
    o[up0 + 1] = o[up0 + 1] - ldc(in, (long)(up0 + 1) * v5 + i + 1);
  }
  if (s > 0) {
    o[dn0] = o[dn0] - ldc(in, (long)dn0 * v5 + i - 1);
    o[dn0 + 1] = o[dn0 + 1] - ldc(in, (long)(dn0 + 1) * v5 + i - 1);
  }
  // wall mass (dirac.F90:119-176, 356-402)
  const unsigned col = i - s;
  if (imass == 1 || imass == 3) {
    if (s == ls - 1) {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        const cx<double> r0 = ldc(in, (long)(up0 + k) * v5 + col);
        o[up0 + k] = o[up0 + k] + (imass == 1 ? m * r0 : m * muli(r0));
      }
    }
    if (s == 0) {
#pragma unroll
      for (int k = 0; k < 2; k++) {
        const cx<double> rl = ldc(in, (long)(dn0 + k) * v5 + col + ls - 1);
        o[dn0 + k] = o[dn0 + k] + (imass == 1 ? m * rl : (-m) * muli(rl));
      }
    }
  } else if (imass == 5) {
    const double sg = DAG ? -1.0 : 1.0;
    if (s == ls - 1) {
#pragma unroll
      for (int k = 0; k < 2; k++) o[up0 + k] = o[up0 + k] + (sg * m) * muli(cur.v[dn0 + k]);
    }
    if (s == 0) {
#pragma unroll
      for (int k = 0; k < 2; k++) o[dn0 + k] = o[dn0 + k] + (sg * m) * muli(cur.v[up0 + k]);
    }
  }
}

// Grid barrier with the per-source reductions folded in.  Each block stores
// the barrier's generation into its own arrival slot (release, device scope:
// its partial sums and field stores before it are visible to whoever
// acquires the slot); block 0 acquires every slot, sums each source's
// partials in index order into `sums`, folds the sources' stop flags into
// `all_done` and releases the generation; every block acquires the release.
// No counter is shared, so arrivals do not serialise on one address (the
// cooperative-groups grid barrier serialises one atomic per block, ~0.15 us
// each here).  Every wait is bounded: a wait that exceeds SPIN_LIMIT polls, or
// sees another block's timeout, raises `error` and returns false, and the
// kernel then leaves its loop, so the grid always drains.
struct GridBar {
  unsigned *arrive;   // [nblocks] last generation each block arrived at
  unsigned *release;  // [1] last generation released by block 0
  int *error;         // [1] a wait timed out
  double *sums;       // [nsrc] the per-source sums of the last barrier
  int *all_done;      // [1] every source stopped (written at each barrier)
};
constexpr unsigned SPIN_LIMIT = 1u << 20;  // polls of ~1 us: about a second

__device__ inline bool bar_failed(const GridBar &B) {
  return __hip_atomic_load(B.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
}

__device__ bool grid_barrier(const GridBar &B, unsigned gen, int nblocks, int nsrc, int bs,
                             const double *part, const int *done) {
  __shared__ int s_ok;
  __syncthreads();  // the block's partial sums and field stores are issued
  if (threadIdx.x == 0)
    __hip_atomic_store(&B.arrive[blockIdx.x], gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  if (blockIdx.x == 0) {
    int ok = 1;
    for (int j = threadIdx.x; ok && j < nblocks; j += BLK) {
      unsigned spins = 0;
      while (__hip_atomic_load(&B.arrive[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != gen) {
        if (++spins > SPIN_LIMIT || bar_failed(B)) {
          ok = 0;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    ok = __syncthreads_and(ok);
    if (ok) {
      const int q = threadIdx.x;
      if (q < nsrc) {
        double v = 0.0;
        for (int k = 0; k < bs; k++) v += part[(long)q * bs + k];
        B.sums[q] = v;
      }
      if (q == 0) {
        int all = 1;
        for (int k = 0; k < nsrc; k++) all &= done[k] ? 1 : 0;
        *B.all_done = all;
      }
      __syncthreads();
      if (q == 0) __hip_atomic_store(B.release, gen, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (threadIdx.x == 0) {
      __hip_atomic_store(B.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (threadIdx.x == 0) {
    int ok = 1;
    unsigned spins = 0;
    while (__hip_atomic_load(B.release, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != gen) {
      if (++spins > SPIN_LIMIT || bar_failed(B)) {
        ok = 0;
        __hip_atomic_store(B.error, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    s_ok = ok;
  }
  __syncthreads();
  const int ok = s_ok;
  __syncthreads();  // s_ok is read by every thread before the next barrier writes it
  return ok != 0;
}

__global__ __launch_bounds__(BLK) void k_msf_persistent(PArgs A, GridBar B) {
  const int src = blockIdx.x / A.bs, lb = blockIdx.x % A.bs;
  const int nblocks = gridDim.x;
  const unsigned v5 = (unsigned)A.d.v5();
  const long ns = 4L * v5;
  const int n = A.n;
  const double2 *bv = A.b + src * ns;
  double2 *r = A.r + src * ns, *h = A.h + src * ns;
  double2 *X = A.x + (long)src * n * ns, *P = A.p + (long)src * n * ns;
  const double m = (double)A.am[src];
  const double diag = (double)((3.0f - A.am3) + 1.0f);
  double *pa = A.part, *pl = A.part + (long)A.nsrc * A.bs;  // [nsrc][bs] each
  const long pslot = (long)src * A.bs + lb;
  const unsigned first = (unsigned)lb * BLK + threadIdx.x, stride = (unsigned)A.bs * BLK;
  __shared__ double s_om[MAXSHIFT], s_gi[MAXSHIFT], s_zt[MAXSHIFT];
  __shared__ int s_mine[MAXSHIFT];
  __shared__ int s_nmin, s_nmax;
  unsigned gen = 0;

  // r = p_i = b, x_i = 0, |b|^2
  double red = 0.0;
  for (unsigned site = first; site < v5; site += stride) {
#pragma unroll
    for (int a = 0; a < 4; a++) {
      const long e = (long)a * v5 + site;
      const double2 v = bv[e];
      r[e] = v;
      red += v.x * v.x + v.y * v.y;
      for (int i = 0; i < n; i++) {
        P[i * ns + e] = v;
        X[i * ns + e] = make_double2(0.0, 0.0);
      }
    }
  }
  red = block_sum(red);
  if (threadIdx.x == 0) pl[pslot] = red;
  bool alive = grid_barrier(B, ++gen, nblocks, A.nsrc, A.bs, pl, A.done);
  const double b2 = alive ? B.sums[src] : 0.0;

  // the scalar recurrences: block-uniform values, and shift sh = threadIdx.x
  const int sh = threadIdx.x;
  const double sig = sh < n ? A.sig[src * MAXSHIFT + sh] : 0.0;
  double z_i = 1.0, z_ii = 1.0, z_iii = 1.0, om_i = 0.0;
  int act = (sh < n && b2 != 0.0) ? 1 : 0, cgr_i = 0;
  const double tol = A.tol[src], tol2 = tol * tol * b2;
  double delta = b2, omega = 1.0, gammag = 0.0;
  int mini = 1, maxi = n, cgret = 0;
  int nonfinite = isfinite(b2) ? 0 : 1;
  int stop = (A.maxcg <= 0 || b2 == 0.0 || nonfinite) ? 1 : 0;

  for (int k = 0; alive; k++) {
    // h = M r + gamma h (h = M p: p = r + gamma p is never formed) ; |h|^2
    red = 0.0;
    if (!stop) {
      for (unsigned site = first; site < v5; site += stride) {
        cx<double> o[4];
        op_site<false>(r, A.u, A.d, site, m, A.imass, diag, o);
#pragma unroll
        for (int a = 0; a < 4; a++) {
          const long e = (long)a * v5 + site;
          if (k > 0) o[a] = o[a] + gammag * ldc(h, e);
          st(h, e, o[a]);
          red += norm2(o[a]);
        }
      }
    }
    red = block_sum(red);
    if (threadIdx.x == 0) {
      pa[pslot] = red;
      if (lb == 0) A.done[src] = stop;
    }
    if (!(alive = grid_barrier(B, ++gen, nblocks, A.nsrc, A.bs, pa, A.done))) break;
    // every block reads the same flag, released by the barrier: the exit is
    // uniform over the grid, so every block takes the same barriers
    if (*B.all_done) break;

    // alpha, omega, zeta''' and omega_i (multishift_module.F90:118-135)
    if (!stop) {
      const double al = B.sums[src];
      const double os = omega, gg = gammag, om = -delta / al;
      if (!isfinite(om)) {
        stop = 1;
        nonfinite = 1;
      } else {
        if (sh < n && sh >= mini - 1 && sh < maxi && act) {
          z_iii = (z_i * z_ii * os) / (om * gg * (z_i - z_ii) + z_i * os * (1.0 - sig * om));
          om_i = om * z_iii / z_ii;
        }
        cgret += 1;
        omega = om;
      }
    }
    // r += omega M^+ h ; |r|^2
    red = 0.0;
    if (!stop) {
      for (unsigned site = first; site < v5; site += stride) {
        cx<double> o[4];
        op_site<true>(h, A.u, A.d, site, m, A.imass, diag, o);
#pragma unroll
        for (int a = 0; a < 4; a++) {
          const long e = (long)a * v5 + site;
          const cx<double> rv = ldc(r, e) + omega * o[a];
          st(r, e, rv);
          red += norm2(rv);
        }
      }
    }
    red = block_sum(red);
    if (threadIdx.x == 0) pl[pslot] = red;
    if (!(alive = grid_barrier(B, ++gen, nblocks, A.nsrc, A.bs, pl, A.done))) break;

    // lambda, gamma, gamma_i, convergence; x_i and p_i (136-176, 230-275)
    if (!stop) {
      const double L = B.sums[src];
      if (!isfinite(L)) {
        stop = 1;
        nonfinite = 1;
      } else {
        const double gam = L / delta;
        const bool mine = sh < n && sh >= mini - 1 && sh < maxi && act;
        if (threadIdx.x == 0) {
          s_nmax = 0;
          s_nmin = n + 1;
        }
        __syncthreads();
        if (mine) {
          s_om[sh] = om_i;
          s_gi[sh] = gam * z_iii * om_i / (z_ii * omega);
          s_zt[sh] = z_iii;
          cgr_i = cgret;
          if (z_iii * z_iii * L <= tol2) {
            act = 0;
          } else {
            atomicMax(&s_nmax, sh + 1);
            atomicMin(&s_nmin, sh + 1);
          }
        }
        if (sh < n) s_mine[sh] = mine ? 1 : 0;
        __syncthreads();
        const int nmin = s_nmin, nmax = s_nmax;
        if (sh < n && sh >= nmin - 1 && sh < nmax && act) {
          z_i = z_ii;
          z_ii = z_iii;
        }
        for (unsigned site = first; site < v5; site += stride) {
          cx<double> rv[4];
#pragma unroll
          for (int a = 0; a < 4; a++) rv[a] = ldc(r, (long)a * v5 + site);
          for (int i = 0; i < n; i++) {
            if (!s_mine[i]) continue;
            const double oi = s_om[i], gi = s_gi[i], zt = s_zt[i];
#pragma unroll
            for (int a = 0; a < 4; a++) {
              const long e = i * ns + (long)a * v5 + site;
              const cx<double> pv = ldc(P, e);
              st(X, e, ldc(X, e) - oi * pv);
              st(P, e, gi * pv + zt * rv[a]);
            }
          }
        }
        gammag = gam;
        mini = nmin;
        maxi = nmax;
        delta = L;
        if (nmax == 0 || cgret >= A.maxcg) stop = 1;
        __syncthreads();  // s_mine and the coefficients are rewritten next iteration
      }
    }
  }
  if (lb == 0) {
    int *o = A.res + src * PRES;
    if (threadIdx.x == 0) {
      o[0] = cgret;
      o[1] = nonfinite;
    }
    if (sh < MAXSHIFT) {
      o[2 + sh] = sh < n ? cgr_i : 0;
      o[2 + MAXSHIFT + sh] = sh < n ? act : 0;
    }
  }
}

}  // namespace

bool msfast_persistent_fits(const MsfArgs &A) {
  // opt-in (THR_PERSIST=1).  Measured on MI355X, 4 sources of (M^+M) x = b:
  // at 8^3 x Ls=8 an iteration takes 22.1 us here against 22.9 us for the
  // streaming kernels (four launches), and at 12^2 x 24 x Ls=16 100 us at
  // best against 39 us.  A grid barrier across the 8 XCDs has to write back
  // and invalidate each XCD's L2 as a kernel boundary does, so it costs about
  // what a launch costs, and the per-site operator here is slower than the
  // streaming Dslash kernels.
  const char *e = std::getenv("THR_PERSIST");
  return e && e[0] == '1';
}

int msfast_persistent(const MsfArgs &A, MsfResult *res, hipStream_t s) {
  const Dims &d = A.d;
  const int n = A.nshift, ns_ = A.nsrc;
  const long v5 = d.v5(), nsp = 4 * v5;
  int dev = 0, cus = 0, per_cu = 0, coop = 0;
  THR_HIP(hipGetDevice(&dev));
  THR_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  THR_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  THR_CHECK(coop, THR_ERR_STATE, "cooperative launch not supported on this device");
  THR_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_msf_persistent, BLK, 0));
  THR_CHECK(per_cu >= 1, THR_ERR_STATE, "persistent solver kernel does not fit a CU");
  const long resident = (long)per_cu * cus;
  THR_CHECK(resident >= ns_, THR_ERR_STATE, "persistent solver: too few resident blocks");
  long want = (v5 + BLK - 1) / BLK;
  // THR_PERSIST_BLOCKS caps the blocks per source (each block then sweeps
  // several site ranges), for tests of the strided sweep
  if (const char *cap = std::getenv("THR_PERSIST_BLOCKS")) want = std::max(1L, std::min(want, atol(cap)));
  // at most 128 blocks in all: the barrier's cost grows with the block count
  // (measured best: 64 blocks at 8^3 x 8, 128 at 12^2 x 24 x 16, 4 sources)
  const int bs = (int)std::max(1L, std::min({want, resident / ns_, std::max(1L, 128L / ns_)}));
  // every block must be resident at once for the grid barriers
  THR_CHECK((long)bs * ns_ <= resident, THR_ERR_STATE, "persistent grid exceeds residency");

  double2 *r = (double2 *)workspace(390, sizeof(double2) * nsp * ns_);
  double2 *h = (double2 *)workspace(391, sizeof(double2) * nsp * ns_);
  double2 *p = (double2 *)workspace(363, sizeof(double2) * nsp * ns_ * n);
  double *part = (double *)workspace(393, sizeof(double) * 2 * ns_ * bs);
  int *ib = (int *)workspace(394, sizeof(int) * (size_t)ns_ * (1 + PRES));
  double *db = (double *)workspace(395, sizeof(double) * (size_t)ns_ * (MAXSHIFT + 1));
  const int nblocks = bs * ns_;
  // barrier words: arrival slots [nblocks], release, error, all_done
  unsigned *bw = (unsigned *)workspace(396, sizeof(unsigned) * ((size_t)nblocks + 3));
  double *sums = (double *)workspace(397, sizeof(double) * (size_t)ns_);
  THR_CHECK(r && h && p && part && ib && db && bw && sums, THR_ERR_HIP,
            "workspace allocation failed");
  THR_HIP(hipMemsetAsync(bw, 0, sizeof(unsigned) * ((size_t)nblocks + 3), s));
  double *hd = (double *)host_workspace(8, sizeof(double) * (size_t)ns_ * (MAXSHIFT + 1));
  int *hi = (int *)host_workspace(9, sizeof(int) * (size_t)ns_ * PRES);
  THR_CHECK(hd && hi, THR_ERR_HIP, "host workspace allocation failed");
  for (int k = 0; k < ns_; k++) {
    for (int i = 0; i < MAXSHIFT; i++)
      hd[k * MAXSHIFT + i] = i < n ? (A.aden_src ? A.aden_src[k * n + i] : A.aden[i]) : 0.0;
    hd[ns_ * MAXSHIFT + k] = A.tol_src ? A.tol_src[k] : A.tol;
  }
  THR_HIP(hipMemcpyAsync(db, hd, sizeof(double) * ns_ * (MAXSHIFT + 1), hipMemcpyHostToDevice, s));

  PArgs P{};
  P.d = d;
  P.nsrc = ns_;
  P.n = n;
  P.imass = A.imass;
  P.maxcg = A.maxcg;
  P.bs = bs;
  P.am3 = A.am3;
  for (int k = 0; k < ns_; k++) P.am[k] = A.am[k];
  P.u = A.u64;
  P.b = A.b;
  P.x = A.x;
  P.p = p;
  P.r = r;
  P.h = h;
  P.part = part;
  P.done = ib;
  P.res = ib + ns_;
  P.sig = db;
  P.tol = db + (long)ns_ * MAXSHIFT;
  GridBar B{};
  B.arrive = bw;
  B.release = bw + nblocks;
  B.error = (int *)(bw + nblocks + 1);
  B.all_done = (int *)(bw + nblocks + 2);
  B.sums = sums;
  // a cooperative launch: the runtime starts it only when the whole grid can
  // be resident at once, which the barriers need
  void *args[] = {&P, &B};
  THR_HIP(hipLaunchCooperativeKernel((const void *)k_msf_persistent, dim3((unsigned)nblocks),
                                     dim3(BLK), args, 0, s));
  THR_HIP(hipMemcpyAsync(hi, ib + ns_, sizeof(int) * ns_ * PRES, hipMemcpyDeviceToHost, s));
  int *herr = (int *)host_workspace(10, sizeof(int));
  THR_CHECK(herr, THR_ERR_HIP, "host workspace allocation failed");
  THR_HIP(hipMemcpyAsync(herr, B.error, sizeof(int), hipMemcpyDeviceToHost, s));
  THR_HIP(hipStreamSynchronize(s));
  THR_CHECK(*herr == 0, THR_ERR_STATE, "persistent solver: a grid barrier wait timed out");
  if (A.note_reliable) context_note_reliable(0);
  if (res) {
    res->iterations = 0;
    res->blocks = 0;
    res->nonfinite = 0;
    res->unconverged = 0;
    res->reliable = 0;
    res->switches[0] = 0;
    res->switches[1] = res->switches[2] = -1;
    for (int k = 0; k < ns_; k++) {
      const int *o = hi + k * PRES;
      res->cg_return[k] = o[0];
      res->iterations = std::max(res->iterations, o[0]);
      res->nonfinite += o[1] ? 1 : 0;
      for (int i = 0; i < n; i++) {
        res->cg_returns[k * MAXSHIFT + i] = o[2 + i];
        res->unconverged += o[2 + MAXSHIFT + i] ? 1 : 0;
      }
    }
  }
  return THR_OK;
}

}  // namespace thr
