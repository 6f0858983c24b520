// Kernel launch interfaces (device code lives in dslash.hip / blas.hip).
#pragma once
#include "common.h"

namespace thr {

constexpr int MAXRHS = 32;

// epilogue modes of the fused Dslash kernel
enum EpiMode {
  EPI_STORE = 0,     // out = M r
  EPI_NORM = 1,      // out = M r ; partial += |out|^2
  EPI_DOTW = 2,      // out = M r ; partial += Re <w, out>
  EPI_CG = 3,        // (CG dagger step) r -= a*out ; x += b*p ; partial += |r|^2 ; out not stored
  EPI_HREC = 4,      // out = M r + b*out (in place) ; partial += |out|^2  (h = M p by recurrence)
};

struct DslashLaunch {
  Dims d;
  int nrhs;
  int prec;        // Prec of in/out (and of w/r/x/p in the fused modes)
  int dagger;
  int imass;
  float am3;
  float am[MAXRHS];  // per right-hand side (float32, as the reference's `real :: am`)
  const void *in;
  const float *in_norm;
  void *out;
  float *out_norm;
  const void *links;  // double2 for P64, float2 otherwise
  // epilogue
  int mode;
  const void *w;       // EPI_DOTW
  void *r, *x;         // EPI_CG: r read (and written unless r_out is set)
  const void *p;       // EPI_CG
  void *r_out;         // EPI_CG: optional separate destination of the updated r
  const float *r_norm; // EPI_CG with FP16 storage: per-site scales of r / r_out
  float *r_out_norm;
  const double *a_ptr; // EPI_CG: r -= a_ptr[rhs*sc_stride] * out (device scalar)
  const double *b_ptr; // EPI_CG: x += b_ptr[rhs*sc_stride] * p  (null: no x update);
                       // EPI_HREC: out = M in + b_ptr[rhs*sc_stride] * out
  int sc_stride;       // stride (in doubles) between the per-rhs scalars
  // Scalars derived in the kernel instead of read from a_ptr / b_ptr (the
  // batched multishift's iterations without finalize launches): every block
  // sums its right-hand side's partial sums of the previous pass,
  // s = sum_{k < fsum_n} fsum[rhs * fsum_stride + k] (the finalize kernels'
  // order), and takes the recorded scalars at fdelta[rhs * sc_stride]:
  //   EPI_CG:   a = fdelta / s            (omega = -delta / alpha, a = -omega)
  //   EPI_HREC: b = s / fdelta            (gamma = lambda / delta)
  // Block 0 of each right-hand side stores s at frec[rhs * sc_stride].
  const double *fsum;  // null: read a_ptr / b_ptr
  int fsum_n, fsum_stride;
  const double *fdelta;
  double *frec;
  double *partial;     // [nrhs][nblocks]
  const int *done;     // optional early-exit flag, done[rhs*done_stride] (device)
  int done_stride;
  int wilson4d;        // 1: 4-D Wilson slice operator only (diag 3 - am3, no s hops, no mass)
};

// launch shape of the Dslash kernel: t-chunk marched by each thread and
// whether blocks are remapped so each XCD owns a contiguous lattice range
// (the defaults are the fastest shape measured at 32^3 x 64; options whose
// preconditions a lattice does not meet fall back by themselves)
struct DslashConfig {
  int tchunk = 8;
  int remap = 1;
  int uniform_links = 1;  // scalar link loads: Ls = 64 (also lane-read wall couplings),
                          // Ls = 32 / 16 (two / four lines per wave; off by default at 16)
  int laneshift = 1;      // s hops as DPP lane shifts (wave-edge lanes load theirs when
                          // Ls does not divide 64; default on only when it does)
  int box = 1;            // XCD boxes: each XCD owns one box of an (x, y, t) split
  int lds_x = 1;          // with uniform links: inner x neighbours through LDS
  int lds_y = 0;          // Ls = 64 with the LDS x exchange: blocks of 4 x by 2 y lines
                          // (512 threads), the inner y neighbours also through LDS
  int nt_stores = 0;      // non-temporal output stores
  int occ_cap = 0;        // 0: no cap; k = 1..3: at most k + 1 blocks per CU (k + 1 waves per
                          // SIMD for 4-wave blocks), by padding each block's LDS
};

// block -> lattice mapping of one Dslash launch (see dslash.hip)
enum TileMode { TILE_LINEAR = 0, TILE_REMAP = 1, TILE_BOXES = 2 };
struct Tiling {
  int mode;
  unsigned tchunk, nb3, pstride;
  unsigned px, py;      // boxes: split of x and y (t split = 8 / (px py))
  unsigned bx, by, bt;  // boxes: box extents
  unsigned yl, ntx;     // 4 x 2-line blocks (lds_y); tiles per row of the slab or box
};

int launch_dslash(const DslashLaunch &L, hipStream_t s);
int launch_dslash_cfg(const DslashLaunch &L, DslashConfig c, hipStream_t s);
// partial-sum entries per right-hand side written by every Dslash epilogue
// (the row stride of the partial buffer; entries past a launch's grid are 0)
long dslash_blocks(const Dims &d);
// blocks per right-hand side of a launch with configuration c: the partial
// entries that launch fills with its sums
long dslash_grid_blocks(const Dims &d, const DslashConfig &c);
// the pinned shape for (d, prec) and a launch of nrhs right-hand sides: one
// shape per batch class (1, or more); a batch without its own takes the
// single-RHS shape, and without any the default
DslashConfig dslash_config(const Dims &d, int prec, int nrhs = 1);
// which: 0 single-RHS class, 1 batches, -1 both
void dslash_set_config(const Dims &d, int prec, DslashConfig c, int which = -1);
// times the launch shapes on scratch fields of nrhs right-hand sides and pins
// the fastest for (d, prec) in nrhs's batch class
int dslash_autotune(const Dims &d, int prec, const void *links, DslashConfig *best,
                    hipStream_t s, int nrhs = 1);

// Wilson-kernel operator on device (FP64): slice-composed, imass 1 or 3
int launch_dslash_wilson(const Dims &d, int nrhs, int dagger, int imass, float am, float am3,
                         const double2 *in, double2 *out, const double2 *links, double2 *tmp,
                         hipStream_t s);

// layout conversion between the reference's haloed complex128 arrays and the
// device layout; prec selects the device storage
int launch_pack_spinor(const Dims &d, const double2 *haloed, void *dst, float *dst_norm, int prec,
                       hipStream_t s);
int launch_unpack_spinor(const Dims &d, const void *src, const float *src_norm, int prec,
                         double2 *haloed, int fill_halo, hipStream_t s);
int launch_pack_links(const Dims &d, const double2 *haloed, double2 *dst64, float2 *dst32,
                      hipStream_t s);
int convert_spinors(const Dims &d, int nrhs, const void *src, const float *snrm, int sprec,
                    void *dst, float *dnrm, int dprec, hipStream_t s);

// u = 1 + i theta with antiperiodic t-links (gauge.F90:12-43), device layout
int launch_coef(const Dims &d, const float *theta, int ibound, double2 *links, hipStream_t s);

// one (R, X2) pair of a batched derivs with its weights fl32(tzi/2) and
// +-tzi (tzi = fl32(2 anum), + for iflag 0)
struct DerivPair {
  const double2 *r, *x2;
  double w1, w2;
};
DerivPair derivs_pair(const double2 *R, const double2 *X2, double anum, int iflag);
// dSdpi += the Shamir derivs of every pair, in one pass (Ls <= 64);
// pairs_dev: device buffer of npair entries (filled from pairs_host); part:
// device buffer of npair * 3 V4 doubles (the pairs' terms before their sum)
int launch_derivs_batch(const Dims &d, int npair, const DerivPair *pairs_host,
                        DerivPair *pairs_dev, double *part, double *dsdpi, hipStream_t s);
// fermion force (derivs_module.F90:99-196), FP64, accumulates into dsdpi [mu][t][y][x]
int launch_derivs(const Dims &d, const double2 *R, const double2 *X2, double anum, int iflag,
                  double *dsdpi, hipStream_t s);

// Wilson-kernel force (derivs_module.F90:41-97, 198-410): shared part plus the
// s-coupling slice terms; pair_buf is device scratch for 2*Ls pair records
int launch_derivs_wilson(const Dims &d, const double2 *R, const double2 *X2, double anum,
                         int iflag, float am, int imass, double *dsdpi, void *pair_buf,
                         hipStream_t s);
constexpr int SLICE_PAIR_BYTES = 48;

}  // namespace thr
