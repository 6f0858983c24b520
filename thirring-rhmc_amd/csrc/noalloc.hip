// Allocation-free congrad on device fields: the solver of thr_congrad_dev
// (measure_module.F90:9-140, Shamir kernel) with its scratch -- three
// spinors, the solver state and the per-block partial sums -- carved out of
// one file-scope device array instead of the library's growable workspace.
// For callers that may not allocate device or pinned host memory (code under
// hipGraph capture, tight serving loops): nothing here calls hipMalloc,
// hipFree, hipHostMalloc or a stream constructor.
//
// Each GPU holds its own copy of `scratch` (NOALLOC_SCRATCH_BYTES) from the
// library's load until the process ends.  Its device address is taken once per
// device with hipGetSymbolAddress; host code uses only that address.  Calls
// on one GPU share the array, and the last call's contents stay in it: every
// buffer is written before it is read within a call (the state by its init
// kernel, the partial sums by each Dslash launch, which also zero-fills the
// entries past its grid), so no clearing is needed, and calls that overlap on
// one GPU race.
#include <mutex>

#include "../../include/thirring_hip.h"
#include "solvers.h"

namespace thr {

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAX_DEVICES = 64;

__device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

std::mutex g_noalloc_mu;
void *g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
// the stop-flag read-back of the solver loop (pageable: no pinned allocation)
thread_local int t_host_flag;

inline size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// device address of this GPU's scratch array (looked up on its first use)
int scratch_address(void **ptr) {
  int dev = -1;
  THR_HIP(hipGetDevice(&dev));
  THR_CHECK(dev >= 0 && dev < NOALLOC_MAX_DEVICES, THR_ERR_ARG, "device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    THR_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    THR_CHECK(p, THR_ERR_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_ptr[dev] = p;
  }
  *ptr = g_scratch_ptr[dev];
  return THR_OK;
}

}  // namespace

size_t noalloc_scratch_bytes(const Dims &d) {
  const size_t spinor = align256(sizeof(double2) * (size_t)d.nspinor());
  return 3 * spinor + align256(cg_state_bytes()) +
         align256(sizeof(double) * (size_t)dslash_blocks(d));
}

}  // namespace thr

using namespace thr;

extern "C" {

unsigned long long thr_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int thr_congrad_dev_noalloc(const void *phi, void *x, const void *links64, const int dims[4],
                            float res, float am, int imass, float am3, int niterc, int *itercg,
                            int *iterations, void *stream) {
  THR_CHECK(dims, THR_ERR_ARG, "dims is NULL");
  for (int k = 0; k < 4; k++) THR_CHECK(dims[k] > 0, THR_ERR_ARG, "dims must be positive");
  THR_CHECK((long)dims[0] * dims[1] * dims[2] * dims[3] < (1L << 31) / 4, THR_ERR_SIZE,
            "lattice too large for 32-bit site indexing");
  THR_CHECK(imass == 1 || imass == 3 || imass == 5, THR_ERR_ARG, "imass must be 1, 3 or 5");
  THR_CHECK(phi && x && links64 && itercg, THR_ERR_ARG, "NULL argument");
  const Dims d{dims[0], dims[1], dims[2], dims[3]};
  const size_t need = noalloc_scratch_bytes(d);
  THR_CHECK(need <= NOALLOC_SCRATCH_BYTES, THR_ERR_SIZE,
            "lattice too large for the allocation-free congrad: it needs " + std::to_string(need) +
                " bytes of scratch, the library holds " + std::to_string(NOALLOC_SCRATCH_BYTES));
  void *base = nullptr;
  TRY_RC(scratch_address(&base));
  char *p = (char *)base;
  const size_t spinor = align256(sizeof(double2) * (size_t)d.nspinor());
  CgBuffers B{};
  B.x1 = (double2 *)p;
  B.r = (double2 *)(p + spinor);
  B.p = (double2 *)(p + 2 * spinor);
  B.state = p + 3 * spinor;
  B.partial = (double *)(p + 3 * spinor + align256(cg_state_bytes()));
  OpDesc op{};
  op.d = d;
  op.kernel = 0;
  op.imass = imass;
  op.am = am;
  op.am3 = am3;
  op.u64 = (const double2 *)links64;
  // no stream given: the null stream (creating one could allocate)
  return cg_reference_bufs(op, (const double2 *)phi, (double2 *)x, res, niterc, itercg,
                           iterations, (hipStream_t)stream, B, &t_host_flag);
}

}  // extern "C"
