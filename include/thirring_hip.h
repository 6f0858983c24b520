/*
 * thirring_hip.h -- C ABI of the MI355X (gfx950) domain-wall hot path.
 *
 * Drop-in boundary for j-ostmeyer/thirring-rhmc's Fortran subroutine API.
 * Every entry point returns 0 (THR_OK) or a THR_ERR_* code; the message of
 * the last failure is available from thr_last_error().  Plain pointers and
 * sizes only.  `dims` is always {Lx, Ly, Lt, Ls} (ksizex_l, ksizey_l,
 * ksizet_l, kthird_l); `kernel` selects the generation kernel: 0 Shamir
 * (GENERATE_WITH_SHAMIR), 1 Wilson kernel (GENERATE_WITH_WILSON).
 *
 * Two families of entry points:
 *
 *  (1) host-array entry points taking the reference's own haloed
 *      column-major complex(dp) arrays (interleaved re/im doubles):
 *        spinor  Phi(0:Ls+1, 0:Lx+1, 0:Ly+1, 0:Lt+1, 4)
 *        links   u(0:Lx+1, 0:Ly+1, 0:Lt+1, 3)
 *        shifts  x1(Ls, Lx, Ly, Lt, 4, n)        (halo-free)
 *        force   dSdpi(Lx, Ly, Lt, 3)            real(dp)
 *      Output spinors come back with periodic halos filled (what the
 *      reference's callers do with update_halo_5 after every call).
 *      These are what the Fortran shim modules (dirac, measure_module,
 *      multishift_module, qmrherm_module, derivs_module, gauge) bind.
 *
 *  (2) device-resident entry points (suffix _dev) on halo-free device
 *      buffers in the layout [rhs][spin][t][y][x][s] (s fastest; FP16
 *      spinors keep the 4 spins of a site together, [rhs][t][y][x][s][spin],
 *      plus one float scale per site), links
 *      [mu][t][y][x]; `stream` is a hipStream_t.  With a stream, entry
 *      points that only enqueue work (coef, links, dslash) return at once;
 *      with NULL they run on the calling context's own non-blocking stream
 *      and return when the work is complete.  Solvers always return with
 *      their results complete.
 *
 * Scalars the reference declares as default real (am, am3, res, theta) are
 * float here; coefficient arrays are double.
 */
#ifndef THIRRING_HIP_H
#define THIRRING_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define THR_OK 0
#define THR_ERR_ARG 1
#define THR_ERR_HIP 2
#define THR_ERR_NOCONV 3
#define THR_ERR_SIZE 4
#define THR_ERR_STATE 5
#define THR_ERR_NONFINITE 6
/* Failure reporting.  THR_ERR_NOCONV: a solver with an FP64 true-residual
 * contract (thr_multishift_fast_dev, qmrherm/force with use_sp = 2) left a
 * (source, shift) pair above its target after its correction rounds or at
 * maxcg; outputs hold the last iterates.  THR_ERR_NONFINITE: a dot product
 * of a solver came out NaN or Inf (a non-finite source, guess or link field);
 * the solve stops at that iteration instead of running to its cap.  The
 * reference-faithful solvers (congrad, multishift, qmrherm use_sp 0/1) keep
 * the reference's contract for the cap: they return THR_OK with the
 * iteration count equal to the cap, which the caller checks
 * (measure_module.F90:124-135, dwf3d_lib.F90:860-881). */

/* precisions of device spinors */
#define THR_F64 0
#define THR_F32 1
#define THR_F16 2 /* half2 per complex + one float scale per 5-D site; the 4 spins
                     of a site are stored together: [rhs][site][spin].  F32 stores
                     the spins of a site together too ([rhs][site][spin], 32 bytes
                     per site); F64 stores spin planes: [rhs][spin][site] */

/* ---- lifetime ------------------------------------------------------------ */
/* select the device and create the library stream; replaces init_MPI and
 * init_gammas (comms.F90:38-95, gammamatrices.F90:17-132) on one GPU */
int thr_init(int device);
int thr_finalize(void);
const char *thr_last_error(void);

/* Solver contexts: each owns its device scratch buffers and its stream.  A
 * thread that binds a context (NULL: the default one) runs every later call
 * in it, so concurrent solves on different threads share no scratch memory.
 * No counterpart in the reference (one MPI rank per process there); used by
 * the serving loop to overlap independent requests on one GPU.  Solvers in a
 * created context launch eagerly (no hipGraph capture, which allocations or
 * copies on other threads would invalidate); the default context captures. */
int thr_context_create(void **ctx);
int thr_context_destroy(void *ctx);
int thr_context_bind(void *ctx);
/* device bytes currently held by the bound context's scratch buffers */
int thr_context_bytes(unsigned long long *bytes);
/* release the bound context's device scratch buffers once its stream is idle
 * (later calls allocate them again); lets a long-lived worker give back the
 * scratch of a large request instead of holding its high-water mark */
int thr_context_trim(void);
/* Reliable updates in the bound context's batched mixed-precision multishift
 * (thr_multishift_fast_dev and the mixed qmrherm/force): while the Krylov
 * storage is FP32 or FP16, each time the iterated residual of the smallest
 * active shift has fallen by the factor delta since the last update, the base
 * residual is replaced by the FP64 true residual of that shift (scaled by
 * 1/zeta of that shift) and the shift directions are moved with it, so that
 * shift's solution does not inherit the low-precision drift.
 * delta in [0, 1); 0 disables; default 0.1.  No counterpart in the reference,
 * whose FP32 multishift (multishift_module.F90:289-534) has no correction. */
int thr_context_set_reliable(double delta);
/* reliable updates taken by the bound context's last multishift solve
 * (its main solve; corrections not counted) */
int thr_context_reliable_updates(int *count);

/* ---- (1) host-array drop-in entry points -------------------------------- */

/* gauge.F90:12-43  coef(u, theta): u = 1 + i theta, u(:,:,Lt,3) *= -1 when
 * ibound == -1, periodic halo */
int thr_coef(double *u, const float *theta, const int dims[4], int ibound);

/* dirac.F90:11-27, 29-179  dslash / dslash_shamir / dslash_wilson (Phi = M R) */
int thr_dslash(double *phi, const double *r, const double *u, const int dims[4], float am,
               int imass, float am3, int kernel);

/* dirac.F90:409-494, 538-625  dslashd / dslashd_shamir / dslashd_wilson
 * (Phi = M^dagger R); the optional MPI request argument has no meaning here */
int thr_dslashd(double *phi, const double *r, const double *u, const int dims[4], float am,
                int imass, float am3, int kernel);

/* inverter_utils.F90:9-54  dirac_op_shifted / dirac_operator:
 * xout = M^dagger M xin + shift xin */
int thr_dirac_op_shifted(double *xout, const double *xin, const double *u, const int dims[4],
                         float am, int imass, float am3, double shift, int kernel);

/* measure_module.F90:9-140  congrad(Phi, res, itercg, am, imass, iterations):
 * x is vector::X (initial guess in, solution out); niterc as params.F90:57 */
int thr_congrad(const double *phi, double *x, const double *u, const int dims[4], float res,
                float am, int imass, float am3, int niterc, int kernel, int *itercg,
                int *iterations);

/* measure_module.F90:170-423  one noise vector of measure_shamir: the four
 * wall-source solves of the inner loop (measure_module.F90:210-372).  ps and
 * pt are gauss0's haloed real32 fields ps(0:Lx+1, 0:Ly+1, 0:Lt+1, 2) (only
 * the interior is read).  Each source xi goes through dslashd_shamir, then
 * congrad (operator `kernel`, initial guess xi, cap niterc) solves
 * M^dagger M xi = M^dagger xi; pbp1 / pbp2 receive the unnormalised sums
 * psibarpsi1 / psibarpsi2 (re, im) of the reference before its division by
 * kvol, iters the sum of the four itercg.  imass 1, 3 and 5. */
int thr_measure_shamir_noise(const float *ps, const float *pt, const double *u, const int dims[4],
                             float res, float am, int imass, float am3, int niterc, int kernel,
                             double pbp1[2], double pbp2[2], int *iters);

/* measure_module.F90:431-640  one noise vector of measure_wilson (the
 * Wilson-kernel measurement, imass 1 and 3): sources on s = 1 for all four
 * spins from ps, then on s = Ls from pt; each goes through dslashd_wilson,
 * then congrad (operator `kernel`, initial guess the source); the read slice
 * (s = Ls through spins 1, 2; s = 1 through spins 3, 4) is replaced by
 * DWilson(islice, -am3) - islice before the contraction
 * (measure_module.F90:521-531, 594-604).  Outputs as thr_measure_shamir_noise
 * (eight solves' iterations in iters). */
int thr_measure_wilson_noise(const float *ps, const float *pt, const double *u, const int dims[4],
                             float res, float am, int imass, float am3, int niterc, int kernel,
                             double pbp1[2], double pbp2[2], int *iters);

/* multishift_module.F90:13-284 (use_sp = 0) and 289-534 (use_sp = 1)
 * multishift_solver(u, am, imass, ndiagq, aden, anum, output, input, res,
 *                   maxcg, cg_return, cg_returns) */
int thr_multishift(const double *u, const int dims[4], float am, int imass, float am3,
                   int ndiagq, const double *aden, const double *anum, double *output,
                   const double *input, float res, int maxcg, int use_sp, int kernel,
                   int *cg_return, int *cg_returns);

/* derivs_module.F90:8-26, 99-196  derivs(R, X2, anum, iflag, am, imass):
 * accumulates into dsdpi (gforce::dSdpi) */
int thr_derivs(const double *r, const double *x2, double anum, int iflag, const int dims[4],
               float am, int imass, float am3, int kernel, double *dsdpi);

/* qmrherm_module.F90:20-194  qmrherm(Phi, x, res, itercg, am, imass, anum,
 * aden, ndiagq, iflag, use_sp, cg_returns).  x1 (Ls,Lx,Ly,Lt,4,ndiagq) and
 * phi0 (haloed, 25 slots) are the module arrays; dsdpi accumulates for
 * iflag 2/3; max_qmr_iters as params.F90:56.  use_sp 0: the reference's FP64
 * multishift; 1: its FP32 variant (multishift_module.F90:289-534); 2 (an
 * addition, Shamir kernel): the batched solver with FP64-then-FP32 Krylov
 * storage and FP64 defect correction, every shift to true residual
 * |Phi - (M^+M + aden_i) x_i| <= res |Phi|.  The Wilson kernel solves in FP64
 * for every use_sp, as the reference does. */
int thr_qmrherm(const double *phi, double *x, float res, int *itercg, float am, int imass,
                float am3, const double *anum, const double *aden, int ndiagq, int iflag,
                int use_sp, const double *u, const int dims[4], int max_qmr_iters, int kernel,
                double *x1, double *phi0, double *dsdpi, int *cg_returns);

/* force_module.F90:9-113  force(Phi, res1, am, imass) for Nf = 1 (rational
 * branch): four qmrherm calls with the 12-pole guidance coefficients
 * (anum4g/aden4g, bnum2g/bden2g, anum2g/aden2g, remezg), the antiperiodic
 * flip of dSdpi(:,:,Lt,3) when ibound == -1, plus beta*Nf*theta.  theta is
 * trial::theta (Lx,Ly,Lt,3) float; dsdpi is overwritten; itercg[4] (may be
 * NULL) receives the four qmrherm iteration counts; use_sp as for
 * thr_qmrherm (the reference's spmd switch is use_sp = 1).  kernel 0 = Shamir, 1 =
 * Wilson kernel (the GENERATE_WITH_* switch; single precision is not used with
 * the Wilson kernel, qmrherm_module.F90:54-63). */
int thr_force(const double *phi, float res1, float am, int imass, float am3, float beta,
              int ibound, const float *theta, const double *u, const int dims[4],
              const double *anum4g, const double *aden4g, const double *bnum2g,
              const double *bden2g, const double *anum2g, const double *aden2g, int ndiagg,
              int use_sp, int max_qmr_iters, int kernel, double *dsdpi, int *itercg);

/* dwf3d_lib.F90:511-616  hamilton(Phi, h, hg, hp, s, res2, am, imass) for
 * Nf = 1 (rational branch); out = {h, hg, hp, s, hf}; itercg[2] (may be NULL);
 * kernel as for thr_force.  use_sp selects the solver of its two qmrherm
 * calls as for thr_qmrherm: 0 is the reference (FP64); 2 solves with
 * mixed-precision storage and FP64 correction of every shift to true
 * residual res2 |Phi|, the action sum staying FP64. */
int thr_hamilton(const double *phi, float res2, float am, int imass, float am3, float beta,
                 const float *theta, const float *pp, const double *u, const int dims[4],
                 const double *anum4, const double *aden4, const double *bnum2,
                 const double *bden2, int ndiag, int use_sp, int max_qmr_iters, int kernel,
                 double *out, int *itercg);

/* Source of the uniform deviates that end a molecular-dynamics trajectory:
 * returns the caller's next rano(yran, idum, 1, 1, 1) (evolution.F90:159-161). */
typedef float (*thr_uniform_fn)(void *ctx);

/* evolution.F90:113-179  evolve_theta_pp(iterl, dt, pp, phi, res, am, imass, itot)
 * for Nf = 1 (rational branch of force):
 *   force; pp -= real(0.5*dt*dSdpi);
 *   for iter = 1..4*iterl: theta += dt*pp; u = coef(theta); force;
 *     ytest = ytest_fn(ctx); if ytest < 1/iterl: pp -= real(0.5*dt*dSdpi), stop;
 *     else pp -= real(dt*dSdpi).
 * theta (trial::theta, (Lx,Ly,Lt,3) float) and pp ((Lx,Ly,Lt,3) float) are
 * updated in place; u (trial::u, haloed as thr_coef writes it) is rewritten
 * from the final theta; dsdpi receives the last force; *itot is increased by
 * min(iter, 4*iterl) (178).  The force arguments are those of thr_force;
 * ytest_fn is called once per MD step, in the reference's order. */
int thr_evolve_theta_pp(int iterl, float dt, float *pp, const double *phi, float res, float am,
                        int imass, float am3, float beta, int ibound, float *theta, double *u,
                        const int dims[4], const double *anum4g, const double *aden4g,
                        const double *bnum2g, const double *bden2g, const double *anum2g,
                        const double *aden2g, int ndiagg, int use_sp, int max_qmr_iters,
                        int kernel, thr_uniform_fn ytest_fn, void *ctx, double *dsdpi, int *itot);

/* ---- (2) device-resident entry points ------------------------------------ */

/* links from theta on device (gauge.F90:12-43); links64 [3][Lt][Ly][Lx] complex128 */
int thr_coef_dev(void *links64, const float *theta, const int dims[4], int ibound, void *stream);

/* convert FP64 device links to the FP32 copy used by the FP32/FP16 operators */
int thr_links_to_f32_dev(void *links32, const void *links64, const int dims[4], void *stream);

/* streaming device-to-device copy of `bytes` (a multiple of 16; 16-byte
 * aligned pointers), non-temporal.  No counterpart in the reference: it is
 * the plain read+write stream that bench.py times as the practical HBM
 * ceiling next to the Dslash, and a field copy that avoids the runtime's
 * blit path. */
int thr_device_copy(void *dst, const void *src, unsigned long long bytes, void *stream);

/* Synthetic fields on the device (no counterpart in the reference, whose
 * noise comes from host generators): n real values, each sigma * N(0, 1),
 * written as double (prec THR_F64) or float (THR_F32).  Counter-based:
 * element i depends only on (seed, i), so a field is the same for any stream
 * or launch.  Used for seeded sources (sigma = 1/sqrt(2) per component gives
 * unit complex variance) and theta (sigma = 1/sqrt(beta)) without host
 * staging. */
int thr_gaussian_dev(void *out, unsigned long long n, int prec, double sigma,
                     unsigned long long seed, void *stream);

/* norms[k] = |x_k| (FP64 L2 norm) of nvec complex128 vectors of n elements
 * each, stored one after another; deterministic summation order; returns
 * with the norms on the host (no counterpart in the reference). */
int thr_norms_dev(const void *x, int nvec, unsigned long long n, double *norms, void *stream);

/* out = M in (dagger = 0) or M^dagger in (dagger = 1) for nrhs right-hand
 * sides with per-rhs masses am[nrhs]; prec THR_F64/F32/F16 (links must be
 * complex128 for F64, complex64 otherwise; norms only for F16) */
int thr_dslash_dev(void *out, float *out_norm, const void *in, const float *in_norm,
                   const void *links, const int dims[4], int nrhs, const float *am, int imass,
                   float am3, int dagger, int prec, void *stream);

/* Launch shape of the Dslash kernel for (dims, prec): each thread marches
 * through `tchunk` t-slices (t hops from registers); `flags` is a bit set:
 *   THR_SHAPE_XCD_REMAP      give every XCD a contiguous range of blocks
 *   THR_SHAPE_UNIFORM_LINKS  Ls = 64 (a wave is one s-line): links through
 *                            scalar loads, wall couplings as lane reads;
 *                            Ls = 32 / 16 (two / four s-lines per wave): two /
 *                            four scalar loads per link (needs lane shifts)
 *   THR_SHAPE_LANESHIFT      s hops as DPP lane shifts (any Ls: when s-lines
 *                            cross waves the wave-edge lanes load theirs)
 *   THR_SHAPE_XCD_BOXES      every XCD owns one box of a split of (x, y, t)
 *                            into 8 (used when such a split divides the
 *                            lattice; the remap bit is then ignored)
 *   THR_SHAPE_LDS_X          with uniform links and Lx % 4, 8, 16 == 0 (Ls =
 *                            64, 32, 16): the inner x neighbours of a block
 *                            come through LDS
 *   THR_SHAPE_NT_STORES      non-temporal output stores
 *   k << THR_SHAPE_OCC_SHIFT occupancy cap, k = 1..3: at most k + 1 resident
 *                            blocks per CU (LDS padding; 0 = no cap)
 *   THR_SHAPE_LDS_Y          with THR_SHAPE_LDS_X at Ls = 64, Lx % 4 == 0 and
 *                            Ly % 2 == 0 (per XCD box with boxes): blocks of
 *                            4 x by 2 y s-lines (512 threads), the inner y
 *                            neighbours also through LDS
 * No counterpart in the reference (a tuning knob of this library).
 * thr_dslash_autotune times the candidates (tchunk 1..16 x flags) on scratch
 * fields with the given links (complex128 for F64, complex64 otherwise) and
 * pins the fastest for the process; thr_dslash_set_config pins a given shape;
 * both apply to every later launch (solvers included) with these dims and
 * precision. */
#define THR_SHAPE_XCD_REMAP 1
#define THR_SHAPE_UNIFORM_LINKS 2
#define THR_SHAPE_LANESHIFT 4
#define THR_SHAPE_XCD_BOXES 8
#define THR_SHAPE_LDS_X 16
#define THR_SHAPE_NT_STORES 32
#define THR_SHAPE_OCC_SHIFT 6
#define THR_SHAPE_LDS_Y 256
int thr_dslash_autotune(const void *links, const int dims[4], int prec, int *tchunk, int *flags,
                        void *stream);
/* thr_dslash_autotune timing launches over nrhs right-hand sides (1..32), for
 * callers that apply the operator to batches: nrhs > 1 pins the shape of the
 * batch class of these dims and precision (launches of 2..32 right-hand
 * sides), nrhs = 1 the single one (no counterpart in the reference) */
int thr_dslash_autotune_nrhs(const void *links, const int dims[4], int prec, int nrhs,
                             int *tchunk, int *flags, void *stream);
int thr_dslash_set_config(const int dims[4], int prec, int tchunk, int flags);
int thr_dslash_get_config(const int dims[4], int prec, int *tchunk, int *flags);
/* the shape a launch of nrhs right-hand sides uses: shapes are pinned per
 * batch class (one right-hand side, or several); thr_dslash_autotune_nrhs
 * pins the batch class, thr_dslash_autotune the single one,
 * thr_dslash_set_config both; a batch without a shape of its own takes the
 * single-RHS one */
int thr_dslash_get_config_nrhs(const int dims[4], int prec, int nrhs, int *tchunk, int *flags);

/* reference-faithful FP64 CG on device fields (phi, x: 4*V5 complex128) */
int thr_congrad_dev(const void *phi, void *x, const void *links64, const int dims[4], float res,
                    float am, int imass, float am3, int niterc, int kernel, int *itercg,
                    int *iterations, void *stream);

/* No counterpart in the reference: thr_congrad_dev (Shamir kernel) for
 * callers that cannot allocate -- code under hipGraph capture, tight serving
 * loops, harnesses that forbid allocation.  It calls no hipMalloc, hipFree,
 * hipHostMalloc or stream constructor: its scratch (three spinors, solver
 * state, partial sums) lives in a file-scope device array of
 * thr_noalloc_scratch_bytes() bytes (64 MiB) that each GPU holds from the
 * library's load to the end of the process.  Lattices needing more return
 * THR_ERR_SIZE (4*V5 complex128 spinors x 3 + small change: up to about
 * 3.4e5 5-D sites, e.g. 16^3 x 64).  Calls on one GPU share the array and
 * must not overlap; every call writes its scratch before reading it.  A NULL
 * stream means the null stream.  Results are those of thr_congrad_dev. */
unsigned long long thr_noalloc_scratch_bytes(void);
int thr_congrad_dev_noalloc(const void *phi, void *x, const void *links64, const int dims[4],
                            float res, float am, int imass, float am3, int niterc, int *itercg,
                            int *iterations, void *stream);

/* derivs_module.F90:8-26, 99-196 on device fields: r, x2 are spinors,
 * dsdpi [mu][t][y][x] float64 accumulates */
int thr_derivs_dev(const void *r, const void *x2, double anum, int iflag, const int dims[4],
                   float am, int imass, int kernel, void *dsdpi, void *stream);

/* dSdpi += the Shamir derivs of npair pairs (r[k], x2[k], anum[k], iflag[k])
 * in one pass (Ls <= 64); r and x2 are host arrays of device spinor
 * pointers.  No counterpart in the reference, which calls derivs once per
 * pole of the rational approximation (qmrherm_module.F90:143-181); the sum
 * equals those calls' up to the order of the additions.  qmrherm iflag 2/3
 * and force use it for the Shamir kernel. */
int thr_derivs_batch_dev(const void *const *r, const void *const *x2, const double *anum,
                         const int *iflag, int npair, const int dims[4], void *dsdpi,
                         void *stream);

/* reference-faithful multishift on device fields; out [ndiagq][4*V5] complex128 */
int thr_multishift_dev(const void *links64, const int dims[4], float am, int imass, float am3,
                       int ndiagq, const double *aden, const double *anum, void *out,
                       const void *input, float res, int maxcg, int use_sp, int kernel,
                       int *cg_return, int *cg_returns, void *stream);

/* Plain FP64 GPU path for comparisons: the reference's multishift recurrences
 * (multishift_module.F90:13-284) with corr_i = 1 and every shift stopped once
 * its shifted residual |zeta_i| |r| falls below tol |b|; res is that tol. */
int thr_multishift_plain_dev(const void *links64, const int dims[4], float am, int imass,
                             float am3, int ndiagq, const double *aden, const void *input,
                             void *out, float tol, int maxcg, int *cg_return, int *cg_returns,
                             void *stream);

/* Batched mixed-precision multishift with blocked shift updates: nsrc sources
 * b[nsrc][4 V5] (complex128) with masses am[nsrc], the same nshift shifts for
 * all; solutions x[nsrc][nshift][4 V5] (complex128).  tol is the FP64 true
 * residual asked for: each shift iterates until |b - (M^+M + aden_i) x_i|
 * (iterated) <= 0.75 tol |b|; Krylov storage drops to FP32 below
 * |r|/|b| < th32 and to FP16 below th16 (<= 0 disables).
 * max_corrections (0..16) > 0 adds FP64 defect correction: the true residual of
 * every (source, shift) is computed in FP64 and each one above tol |b| is solved
 * again for its residual (FP32 Krylov storage, to tol |b| / 2) and the
 * correction added, up to max_corrections rounds, so low-precision storage
 * still meets tol in FP64.
 * iterations: iterations run; cg_return[nsrc]: per-source iterations;
 * corr_iterations (optional): iterations spent in corrections; relres
 * (optional, nsrc*nshift): final true FP64 residuals |b - (M^+M + sigma_i) x_i| / |b|;
 * switches (optional, 3): iteration at which the Krylov storage entered FP64,
 * FP32, FP16 (-1: never).
 * Status: THR_ERR_NOCONV when a true residual (computed whenever relres is
 * asked for or corrections run) stays above tol, or, without them, when a
 * shift is still above its iterated target at maxcg; THR_ERR_NONFINITE when a
 * source's recurrences produce NaN/Inf (that source stops at once).
 * Same recurrences as multishift_module.F90:103-282 (see thr_multishift). */
int thr_multishift_fast_dev(const void *links64, const void *links32, const int dims[4],
                            int nsrc, const float *am, int imass, float am3, int nshift,
                            const double *aden, const void *b, void *x, double tol, int maxcg,
                            double th32, double th16, int max_corrections, int *iterations,
                            int *cg_return, int *corr_iterations, double *relres,
                            int *switches, void *stream);

/* True FP64 residuals relres[src*nshift + i] = |b - (M^+M + aden_i) x_i| / |b|
 * for solutions laid out as thr_multishift_fast_dev's x (check of any solver) */
int thr_true_residuals_dev(const void *links64, const int dims[4], int nsrc, const float *am,
                           int imass, float am3, int nshift, const double *aden, const void *b,
                           const void *x, double *relres, void *stream);

/* evolution.F90:113-179  thr_evolve_theta_pp on device-resident fields: phi
 * [4][Lt][Ly][Lx][Ls] complex128 (halo-free), theta and pp [3][Lt][Ly][Lx]
 * float, links64 [3][Lt][Ly][Lx] complex128 (rewritten by coef each step),
 * dsdpi [3][Lt][Ly][Lx] float64 (the last force); coefficients are host
 * arrays.  Returns when the trajectory has ended and its work is complete. */
int thr_evolve_theta_pp_dev(int iterl, float dt, float *pp, const void *phi, float res, float am,
                            int imass, float am3, float beta, int ibound, float *theta,
                            void *links64, const int dims[4], const double *anum4g,
                            const double *aden4g, const double *bnum2g, const double *bden2g,
                            const double *anum2g, const double *aden2g, int ndiagg, int use_sp,
                            int max_qmr_iters, int kernel, thr_uniform_fn ytest_fn, void *ctx,
                            double *dsdpi, int *itot, void *stream);

#ifdef __cplusplus
}
#endif
#endif
