// Krylov solvers on device-resident fields.
#pragma once
#include "kernels.h"

namespace thr {

constexpr int MAXSHIFT = 32;

// One operator configuration: which DWF kernel, wall mass, links.
struct OpDesc {
  Dims d;
  int kernel;         // 0 Shamir, 1 Wilson kernel
  int imass;
  float am3;
  float am;           // single-rhs mass (multi-rhs masses live in the launch)
  const double2 *u64; // device links [mu][t][y][x]
  const float2 *u32;  // FP32 copy (may be null when only FP64 is used)
};

// growable device workspace, keyed by slot, owned by the calling thread's
// context (or the default context)
struct Context;
Context *context_create();
void context_destroy(Context *c);
void context_bind(Context *c);
size_t workspace_bytes();
void *workspace(int slot, size_t bytes);
// pinned host buffer of the bound context (state read-backs and uploads go
// through these: pageable async copies are staged through runtime buffers
// shared by all threads)
void *host_workspace(int slot, size_t bytes);
double *partial_buffer(int nrhs_blocks);
hipStream_t default_stream();
// whether solvers may capture their blocks as hipGraphs in the bound context
bool context_graphs();
// reliable-update factor of the mixed multishift in the bound context (0: off)
// and the number of reliable updates its last solve took
double context_reliable();
void context_set_reliable(double delta);
void context_note_reliable(int n);
int context_last_reliable();

// reference-faithful CG on M^+M (measure_module.F90:9-140); x in/out (FP64 device)
int cg_reference(const OpDesc &op, const double2 *phi, double2 *x, float res, int niterc,
                 int *itercg, int *iterations, hipStream_t s);

// the same CG on caller-provided device buffers: three spinors, the solver
// state (cg_state_bytes()) and dslash_blocks(d) partial sums.  None needs
// clearing (every one is written before it is read).  host_flag (optional)
// is the host int the stop flag is polled into; given one, nothing is
// allocated (then only the Shamir kernel is allowed: the Wilson kernel's
// unfused epilogue needs library scratch).
struct CgBuffers {
  double2 *x1, *r, *p;
  void *state;
  double *partial;
};
size_t cg_state_bytes();
int cg_reference_bufs(const OpDesc &op, const double2 *phi, double2 *x, float res, int niterc,
                      int *itercg, int *iterations, hipStream_t s, const CgBuffers &B,
                      int *host_flag);

// reference-faithful multishift CG (multishift_module.F90:13-284), FP64 vectors.
// out: [n][4*V5] shift solutions.
// plain = 1: corr_i = 1 and the threshold is res*|b| (stop when the shifted
// residual |zeta_i||r| < res |b|), the rule used for the FP64 comparison path
int multishift_reference(const OpDesc &op, int n, const double *aden, const double *anum,
                         const double2 *input, double2 *out, float res, int maxcg,
                         int *cg_return, int *cg_returns, hipStream_t s, int plain = 0);

// FP32-vector multishift with FP64 scalars (multishift_module.F90:289-534);
// input/output FP64 device fields.
int multishift_reference_sp(const OpDesc &op, int n, const double *aden, const double *anum,
                            const double2 *input, double2 *out, float res, int maxcg,
                            int *cg_return, int *cg_returns, hipStream_t s);

// element-wise FP64 -> FP32 complex conversion
int convert_d2f(long n, const double2 *a, float2 *b, hipStream_t s);
// streaming device-to-device copy (bytes % 16 == 0, 16-byte aligned pointers)
int stream_copy(const void *src, void *dst, size_t bytes, hipStream_t s);

// psibarpsi measurement (measure.hip): wall source sign * eta on (spin, s_wall)
// from a haloed real32 noise field ps(0:Lx+1, 0:Ly+1, 0:Lt+1, 2), and
// per-block partial sums (re, im pairs, wall_dot_blocks(d) of them) of
// sum_x conj(eta(x)) (X - Xsub)(s_read, x, spin), Xsub optional
int launch_wall_source(const Dims &d, const float *ps, int spin, int s_wall, double sign,
                       double2 *out, hipStream_t s);
long wall_dot_blocks(const Dims &d);
int launch_wall_dot(const Dims &d, const float *ps, const double2 *X, const double2 *Xsub, int spin,
                    int s_read, double *partial, hipStream_t s);
// X(s_read) with spins [lo, lo + 2) kept, as a 4-D spinor [spin][t][y][x]
int launch_slice_project(const Dims &d, const double2 *X, int s_read, int lo, double2 *out,
                         hipStream_t s);

// apply M or M^+ (FP64) into out
int apply_op(const OpDesc &op, int dagger, const double2 *in, double2 *out, hipStream_t s);

}  // namespace thr
