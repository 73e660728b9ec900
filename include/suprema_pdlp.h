/*
 * suprema_pdlp.h -- C ABI of the MI355X (gfx950) restarted-PDHG (PDLP) engine.
 *
 * This is the drop-in boundary for the reference's GPU PDLP engine.  Every
 * entry point takes plain pointers and sizes (no C++ or framework types),
 * returns 0 on success or a negative SUPR_E* code, and leaves a message for
 * supr_last_error() on failure.  Nothing here throws.
 *
 * Entry points and the reference interfaces they replace (paths relative to
 * the reference repository):
 *   supr_solve_pdhg_gpu   solve_pdhg_gpu()    include/pdlp_gpu.hpp:22-25,
 *                                             gpu/pdlp_gpu_solve.cu:261-267
 *                         (GpuBackend + pdlp_run, include/pdlp_algo.hpp:49-185)
 *   supr_verify           verify(ranged, x, y) include/verify.hpp:13-20,
 *                                             src/verify.cpp:18-103
 *   supr_scale            scale()             include/scaling.hpp:13,
 *                                             src/scaling.cpp:25-84
 *   supr_pdlp_form_sizes  build_pdlp_form()   include/pdlp.hpp:26-27,
 *   supr_pdlp_form                            src/pdlp.cpp:8-82
 *   supr_extract_solution extract_solution()  include/pdlp.hpp:27-28,
 *                                             src/pdlp.cpp:84-90
 *   supr_gen_transport    tools/gen_lp.cpp:49-135 (same LP family, own PRNG)
 *   supr_gen_blend[_sizes] no reference counterpart: seeded blending LPs (the
 *                         refinery-blending family of tools/gen_refinery.py, LP
 *                         only) for batched workloads
 *   supr_gen_random_lp    no reference counterpart: seeded equality-form LP
 *                         generator (6 nonzeros per column, feasible and bounded
 *                         by construction) for synthetic workloads
 *   supr_batch_*          no reference counterpart: many independent LPs per
 *                         launch (callers: batch drivers, request loops)
 *   supr_server_*         no reference counterpart: continuous batching of LP
 *                         requests with admission by device footprint (callers:
 *                         request loops)
 *   supr_engine_*         no reference counterpart: a resident engine for repeated
 *                         solves of one LP (callers: benchmarks, request loops)
 *   supr_spmv             matvec() include/sparse.hpp:59-71 on the device, i.e.
 *                         GpuBackend::spmv, gpu/pdlp_gpu_solve.cu:211-222, with a
 *                         chosen schedule (callers: kernel parity tests)
 *   supr_bench_kernels    no reference counterpart: times the engine's two
 *                         iteration kernels at a problem's real shapes (callers:
 *                         benchmarks reporting achieved HBM bandwidth)
 *   supr_bench_copy       no reference counterpart: a plain device copy (callers:
 *                         benchmarks calibrating HBM counters and bandwidth)
 *   supr_mps_*            read_mps() + to_ranged_lp()  src/mps_reader.cpp:54-308,
 *                         include/ranged_lp.hpp:27-62 (supr_mps_write: the MPS
 *                         writing of tools/gen_lp.cpp:89-131, for any LP)
 *   supr_last_error       the what() text of the reference's std::runtime_error
 *                         (SOVEREIGN_CUDA_CHECK, gpu/pdlp_gpu_solve.cu:32-38)
 *
 * Threading: every entry point is re-entrant; solves on different threads
 * use independent device state and streams.  The error text is thread-local.
 */
#ifndef SUPREMA_PDLP_H
#define SUPREMA_PDLP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ----------------------------------------------------------- status ----- */
#define SUPR_OK 0
#define SUPR_EINVAL (-1)    /* bad argument (sizes, indices, options) */
#define SUPR_EDEVICE (-2)   /* HIP runtime / device error */
#define SUPR_ENOMEM (-3)    /* host or device allocation failed */
#define SUPR_ENODEV (-4)    /* no gfx950 device visible */
#define SUPR_EINTERNAL (-5)

const char* supr_last_error(void);
/* Library version string. */
const char* supr_version(void);
/* Number of HIP devices visible (0 when none); never fails. */
int supr_device_count(void);

/* ------------------------------------------------------------- LP view -- */
/* One LP "min c'x s.t. rL <= Ax <= rU, l <= x <= u" with A in COO order
 * (row_idx, col_idx, val), the reference's RangedLP (include/ranged_lp.hpp:
 * 10-25).  Infinite bounds are +-INFINITY.  All host memory, read only. */
typedef struct supr_lp {
    int32_t m, n, nnz;
    const int32_t* row_idx;
    const int32_t* col_idx;
    const double* val;
    const double* c;
    const double* rL;
    const double* rU;
    const double* l;
    const double* u;
    double obj_offset;
} supr_lp;

/* --------------------------------------------------------- solve modes -- */
/* SUPR_MODE_PLAIN_FP64: the reference GPU engine's design on MI355X: FP64
 *   everywhere, the host decides every step (one blocking read of the step
 *   statistics per attempt) and runs the KKT check on the host every
 *   check_every accepted steps.
 * SUPR_MODE_FAST_FP64: device-resident loop (no host sync inside the loop,
 *   KKT residuals, restarts and step-size rule on the device, graph-replayed
 *   iterations), FP64 iterates.
 * SUPR_MODE_MIXED: device-resident loop with FP32 iterates / FP32 SpMV and
 *   FP64 step statistics, KKT, restart rule; switches to the FP64 iterate
 *   loop (the polish) once the FP32 iterate stalls or reaches switch_tol. */
#define SUPR_MODE_PLAIN_FP64 0
#define SUPR_MODE_FAST_FP64 1
#define SUPR_MODE_MIXED 2

typedef struct supr_pdlp_options {
    int32_t mode;              /* SUPR_MODE_* */
    int32_t max_iterations;    /* attempts, accepted or not (reference default 200000) */
    int32_t check_every;       /* KKT check cadence in accepted steps (reference: 64) */
    double tol;                /* converged iff max(eps_P, eps_D, eps_G) <= tol */
    double switch_tol;         /* MIXED: FP32 -> FP64 when merit <= switch_tol (0: auto) */
    int32_t stall_checks;      /* MIXED: ... or after this many checks without a new best (0: auto) */
    int32_t graph_attempts;    /* device-resident modes: attempts per graph replay, each replay ending with the KKT
                                  check group; < 0: the same kernel sequence launched on the stream directly,
                                  -graph_attempts attempts per poll; 0 (auto): A with 2^20+ nonzeros launches
                                  directly, 16 attempts per poll, smaller LPs replay 8 x (16 attempts + check group)
                                  per graph (single-LP engine; the batched engine and the server: 16 per replay).
                                  Every choice runs the same kernels in the same order: the same bits. */
    int32_t spmv_lanes_A;      /* lanes per row for A: 0 = a rule of the matrix shape (deterministic: every
                                  run sums in the same order); -1 = timed autotune (may differ run to run);
                                  1..64 powers of two, or 256 = one block per row.  The batched engine:
                                  0 tunes its lanes rule by timing, see supr_batch_create */
    int32_t spmv_lanes_AT;     /* lanes per row for A^T (as spmv_lanes_A) */
    int32_t device;            /* HIP device ordinal */
    int32_t verbose;           /* 0 quiet, 1 phase summary on stderr, 2 also a progress line every 5 s, 3 also per-kernel timing (batched) */
} supr_pdlp_options;

/* Fills the reference's defaults: mode FAST_FP64 (FP64 arithmetic), max_iterations 200000,
 * check_every 64, tol 1e-6, the rest automatic. */
void supr_pdlp_default_options(supr_pdlp_options* opt);

typedef struct supr_pdlp_info {
    int32_t iterations;        /* attempts, as PdlpResult::iterations */
    int32_t restarts;
    int32_t converged;
    int32_t has_point;         /* 0: no finite-merit point was seen; x, y untouched */
    double eps_P, eps_D, eps_G;
    double objective;          /* c'x + obj_offset of the returned point (unscaled form), as the
                                  device's KKT check summed it (FP64, fixed order) */
    int32_t iterations_fp32;   /* MIXED: attempts spent in the FP32 loop */
    int32_t checks;            /* KKT checks performed */
    int32_t lanes_A, lanes_AT; /* SpMV schedules used */
    double seconds_setup;      /* device upload + layout + autotune */
    double seconds_solve;      /* first iteration to verified result on the host clock */
    double seconds_total;
} supr_pdlp_info;

/* Stop callback: polled by the host loop (at least every check_every
 * attempts for the plain mode, every graph replay otherwise); nonzero stops
 * the solve, which then returns the best point seen with converged = 0.
 * The reference's std::atomic<bool>* stop_flag (include/pdlp.hpp:30-38). */
typedef int (*supr_stop_fn)(void* ctx);

/* solve_pdhg_gpu(): unscaled and scaled PDLP forms share one COO pattern
 * (scaled = diag(Dr) A diag(Dc) with scaled c, bounds, rows as produced by
 * supr_scale); x_out (n) and y_out (m) receive the returned point in
 * UNSCALED PDLP-form coordinates, exactly like PdlpResult::x / y.  Forms of
 * 2^16 or more nonzeros are laid out (CSR / ELL of A and A^T) on the device,
 * smaller ones on the host, with the same layout either way; the environment
 * variable SUPR_HOST_LAYOUT=1 forces the host build (also for
 * supr_engine_create).  An LP with 2^16 or more columns whose largest column
 * degree is at least 4x the mean is solved with its columns relabelled
 * (sorted by degree, in groups of 16 dealt over 256-column tiles: the dual
 * half's gathers of x then share cache lines); x_out comes back in the
 * caller's column order.  SUPR_COLUMN_ORDER=given turns this off
 * (also for supr_engine_create and supr_bench_kernels); SUPR_MODE_PLAIN_FP64,
 * the reference engine's design, always solves the columns as given.  When
 * A^T is ELL with one lane per row, an even row count and more than 2^19 rows,
 * the primal half takes two adjacent rows per lane: each row's values are
 * the same bits, but the step statistic's partial sums group the rows
 * differently, so the trajectory differs from one row per lane in the last
 * bits (SUPR_ROW_PAIRS=off). */
int supr_solve_pdhg_gpu(const supr_lp* unscaled_form, const supr_lp* scaled_form,
                        const double* Dr, const double* Dc, const supr_pdlp_options* opt,
                        supr_stop_fn stop, void* stop_ctx, double* x_out, double* y_out,
                        supr_pdlp_info* info);

/* A resident engine: the LP is uploaded, laid out and its SpMV schedules
 * tuned once (supr_engine_create); every supr_engine_solve then runs one
 * solve from the cold start with the data already in HBM and returns what
 * supr_solve_pdhg_gpu returns.  One engine is used by one thread at a time;
 * supr_engine_destroy frees its device memory.  (No reference counterpart:
 * the reference rebuilds GpuBackend per call; callers: benchmarks and
 * request loops that solve an LP repeatedly.) */
typedef struct supr_engine supr_engine;
int supr_engine_create(const supr_lp* unscaled_form, const supr_lp* scaled_form, const double* Dr,
                       const double* Dc, const supr_pdlp_options* opt, supr_engine** engine);
int supr_engine_solve(supr_engine* engine, supr_stop_fn stop, void* stop_ctx, double* x_out,
                      double* y_out, supr_pdlp_info* info);
void supr_engine_destroy(supr_engine* engine);

/* scale() then solve_pdhg_gpu() with the scaling on the device (the
 * reference's src/solve.cpp:195-207 sequence: scale(pdlp_form), then
 * solve_pdhg_gpu(pdlp_form, scaled, Dr, Dc, ...)).  form is the UNSCALED
 * PDLP form; integer (n flags, may be NULL) and ruiz_passes are scale()'s
 * arguments (src/scaling.cpp:25).  The matrix is uploaded once, laid out as
 * CSR / CSC and scaled on the device (bit-identical to supr_scale), so no
 * host CSR build or host scaling precedes the solve; the result is the one
 * supr_solve_pdhg_gpu gives for supr_scale's output, bit for bit. */
int supr_solve_form_gpu(const supr_lp* form, const char* integer, int32_t ruiz_passes,
                        const supr_pdlp_options* opt, supr_stop_fn stop, void* stop_ctx,
                        double* x_out, double* y_out, supr_pdlp_info* info);
/* The reference's solve_mps() PDLP stage after presolve (src/solve.cpp:
 * 194-224) with every O(nnz) step on the GPU: build_pdlp_form, scale(), the
 * solve, extract_solution, z = c - A^T y and verify() on lp itself; while the
 * verifier's worst residual misses opt->tol (and the solve converged), it
 * solves again with a 10x tighter internal tolerance, three attempts at most.
 * lp is the LP as the reference's RangedLP (any row bounds); x_out (n),
 * y_out (m) and z_out (n, may be NULL) are in lp's coordinates; report7 is
 * verify()'s report on lp (as supr_verify's out7); info->iterations sums the
 * attempts of every solve.  Device-resident modes only. */
int supr_solve_lp_gpu(const supr_lp* lp, const supr_pdlp_options* opt, supr_stop_fn stop,
                      void* stop_ctx, double* x_out, double* y_out, double* z_out,
                      supr_pdlp_info* info, double* report7);
/* A resident engine built the same way (see supr_engine_create). */
int supr_engine_create_form(const supr_lp* form, const char* integer, int32_t ruiz_passes,
                            const supr_pdlp_options* opt, supr_engine** engine);

/* y = A x (transposed = 0; y has m entries) or A^T x (transposed = 1; y has
 * n entries) on the device, with lanes-per-row `lanes` (0: the shape rule,
 * -1: timed autotune) and
 * FP32 (fp32 = 1: values, x and the accumulation in FP32) or FP64 arithmetic. */
int supr_spmv(const supr_lp* lp, int32_t transposed, int32_t lanes, int32_t fp32,
              const double* x, double* y);

/* Times the two iteration kernels (primal half: A^T y + projection +
 * averages; dual half: A x + dual step + step decision) at this LP's shapes,
 * each launched reps times after a warm-up with every launch doing an
 * accepted step's full work.  out9 = ms per primal-half launch (FP32), ms per
 * dual-half launch (FP32), the same two in FP64, lanes-per-row of A and of
 * A^T (FP64 choice), ELL width of A and of A^T (0 = CSR), 1 if the FP64
 * primal half forms A^T's values from the scaling factors (every unscaled
 * entry the same value) instead of reading them. */
int supr_bench_kernels(const supr_lp* unscaled_form, const supr_lp* scaled_form,
                       const double* Dr, const double* Dc, const supr_pdlp_options* opt,
                       int32_t reps, double* out9);

/* Times a plain device-to-device copy of `bytes` (a positive multiple of
 * 16) with a 16-byte-vector stream kernel, reps launches after one warm-up;
 * *ms_per_copy = ms per launch (reads `bytes`, writes `bytes`). */
int supr_bench_copy(int64_t bytes, int32_t reps, int32_t device, double* ms_per_copy);

/* No reference counterpart (callers: long-lived processes that make many
 * calls).  Device memory a call releases is kept per device, up to
 * SUPR_DEVICE_CACHE_MB megabytes (default 4096; 0 turns the cache off), and
 * reused by later calls, which then skip hipMalloc / hipFree.  This frees
 * every kept block now (engines, batches and servers keep what they hold). */
void supr_release_cached_memory(void);

/* ------------------------------------------------------- host-side form -- */
/* verify(): out7 = eps_P, eps_D, eps_G, objective, dual_objective,
 * max_row_violation, max_bound_violation (KKTReport, include/verify.hpp:7-11). */
int supr_verify(const supr_lp* lp, const double* x, const double* y, double* out7);

/* scale(): Ruiz passes then Pock-Chambolle; integer may be NULL.  Outputs
 * have the input's sizes (val_s: nnz in the input's COO order). */
int supr_scale(const supr_lp* lp, const char* integer, int32_t ruiz_passes, double* val_s,
               double* c_s, double* rL_s, double* rU_s, double* l_s, double* u_s, double* Dr,
               double* Dc);

/* scale() on the GPU (replaces src/scaling.cpp:25-84 in front of
 * solve_pdhg_gpu, src/solve.cpp:194-195): the same arguments and outputs as
 * supr_scale, bit for bit.  The matrix is laid out on the device as CSR and
 * CSC by a stable sort (COO order kept inside every row and column), then
 * scaled there.  seconds3 (may be NULL) receives the upload + layout, scaling
 * and download times. */
int supr_scale_gpu(const supr_lp* lp, const char* integer, int32_t ruiz_passes, int32_t device,
                   double* val_s, double* c_s, double* rL_s, double* rU_s, double* l_s,
                   double* u_s, double* Dr, double* Dc, double* seconds3);

/* The CSR of A (transposed = 0: the reference's to_csr, include/sparse.hpp:
 * 31-47) or of A^T (transposed = 1: to_csc_as_transposed_csr, :50-57) built
 * on the GPU by a stable radix sort: ptr (rows + 1), idx and perm (nnz; perm
 * = COO position of each entry), equal to the host counting sort's. */
int supr_csr_gpu(const supr_lp* lp, int32_t transposed, int32_t device, int32_t* ptr,
                 int32_t* idx, int32_t* perm);

/* build_pdlp_form() on the GPU (src/pdlp.cpp:8-82): the outputs of
 * supr_pdlp_form (sized by supr_pdlp_form_sizes), bit for bit; dims4 (may be
 * NULL) receives m2, n2, nnz2, n_slack as computed on the device. */
int supr_pdlp_form_gpu(const supr_lp* lp, int32_t device, int32_t* row_idx2, int32_t* col_idx2,
                       double* val2, double* c2, double* rL2, double* rU2, double* l2, double* u2,
                       int32_t* row_perm, double* row_signs, int32_t* dims4);

/* verify() (src/verify.cpp:24-103) with A x and A^T y formed on the GPU in
 * the host's term order: the same out7 as supr_verify, bit for bit. */
int supr_verify_gpu(const supr_lp* lp, const double* x, const double* y, int32_t device,
                    double* out7);

/* An LP in DEVICE memory, laid out for verify(): the CSR of A (m rows:
 * A_ptr m + 1 entries, A_idx / A_val nnz) and of A^T (n rows: AT_ptr n + 1,
 * AT_idx / AT_val nnz; supr_csr_gpu builds both layouts), c, l, u (n) and
 * rL, rU (m); obj_offset is a host value.  The layouts must be valid CSR
 * (row pointers nondecreasing from 0 to nnz, column indices in range): the
 * device arrays are not checked.  The reference's verify() takes the same
 * CSR pair (src/verify.cpp:24, verify(ranged, A, AT, x, y)). */
typedef struct supr_dev_lp {
    int32_t m, n, nnz;
    const int32_t* A_ptr;
    const int32_t* A_idx;
    const double* A_val;
    const int32_t* AT_ptr;
    const int32_t* AT_idx;
    const double* AT_val;
    const double* c;
    const double* rL;
    const double* rU;
    const double* l;
    const double* u;
    double obj_offset;
} supr_dev_lp;

/* supr_verify_gpu for a point and an LP already in device memory, allocating
 * nothing.  No reference counterpart: a variant of supr_verify_gpu for the
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  x (n), y (m) and
 * out7 (7 doubles: the out7 of supr_verify) are device pointers; the kernels
 * run on `stream` (a hipStream_t; NULL: the null stream) of the current
 * device and the call returns without waiting for them.  The same report as
 * supr_verify_gpu to within the summation order (partial sums per block,
 * added in a fixed order: the same bits every call).  The per-block partial
 * sums live in one fixed array per GPU (88 KiB, held until the process ends),
 * sized for the 1,024 blocks per pass at most that the grid-stride passes use,
 * so any LP fits; calls that overlap on one GPU (two streams at once) race
 * on it. */
int supr_verify_gpu_noalloc(const supr_dev_lp* lp, const double* x, const double* y,
                            double* out7, void* stream);

/* build_pdlp_form(): dims4 = m2, n2, nnz2, n_slack. */
int supr_pdlp_form_sizes(const supr_lp* lp, int32_t* dims4);
int supr_pdlp_form(const supr_lp* lp, int32_t* row_idx2, int32_t* col_idx2, double* val2,
                   double* c2, double* rL2, double* rU2, double* l2, double* u2,
                   int32_t* row_perm, double* row_signs);
int supr_extract_solution(int32_t n_orig, int32_t m_orig, int32_t m2, const int32_t* row_perm,
                          const double* row_signs, const double* x2, const double* y2,
                          double* x, double* y);

/* ----------------------------------------------------------- generators -- */
/* Balanced transportation LP (tools/gen_lp.cpp:49-135 family): m = S + D
 * equality rows (suppliers, then demanders), n = S*D columns x_ij in i-major
 * order, 2 nonzeros (+1, +1) per column, nnz = 2*S*D.  Supplies and raw
 * demands are integers U[50,150]; demands are rescaled so the totals match
 * exactly (remainder on the last).  Costs are integers U[1, cost_max];
 * 0 <= x < inf.  Arrays sized by the caller. */
int supr_gen_transport(int32_t S, int32_t D, uint64_t seed, int32_t cost_max, int32_t* row_idx,
                       int32_t* col_idx, double* val, double* c, double* rL, double* rU,
                       double* l, double* u);

/* Random equality-form LP: A is m x n with exactly col_nnz nonzeros per
 * column at distinct uniform rows, values N(0,1) (or +-log-uniform
 * magnitudes in [10^-mag_decades, 10^mag_decades] when mag_decades > 0);
 * x0 ~ U[0, x0_hi]^n, b = A x0 (rL = rU = b); y0 ~ N(0,1)^m, s0 ~ U[0,1]^n,
 * c = A'y0 + s0; bounds [0, u_hi].  nnz = n * col_nnz.  Feasible (x0) and
 * bounded (box).  If col_deg is non-NULL it gives per-column degrees instead
 * of col_nnz (nnz = sum). */
int supr_gen_random_lp(int32_t m, int32_t n, int32_t col_nnz, const int32_t* col_deg,
                       uint64_t seed, double x0_hi, double u_hi, int32_t mag_decades,
                       int32_t* row_idx, int32_t* col_idx, double* val, double* c, double* rL,
                       double* rU, double* l, double* u);

/* ------------------------------------------------------- batched mode -- */
/* Many independent LPs solved together: the instances are laid out
 * block-diagonally on the device once (supr_batch_create); every
 * supr_batch_solve solves a subset of them (ids, or all when ids is NULL; no
 * id twice) from the cold start.  Each launch of the device-resident loop
 * covers every running instance, a fixed number of blocks per instance, each
 * instance with its own step sizes, checks, restarts and (MIXED) FP32 -> FP64
 * hand-over, so an instance's result does not depend on the other instances
 * of the solve.  Modes: FAST_FP64 (default) or MIXED.  opt->spmv_lanes_A == 0
 * tunes the lanes-per-row rule (lanes_bias) on a short run at creation;
 * supr_batch_configure sets it and the per-instance block cap (max_blocks)
 * explicitly (defaults 1.0 and 512).  x_out[r] / y_out[r] (may be NULL) receive
 * the r-th requested instance's point, infos[r] its information; *seconds
 * (may be NULL) the solve's wall time on the device.  (No reference
 * counterpart; callers: batch drivers and request loops.) */
typedef struct supr_batch supr_batch;
int supr_batch_create(int32_t count, const supr_lp* unscaled_forms, const supr_lp* scaled_forms,
                      const double* const* Dr, const double* const* Dc,
                      const supr_pdlp_options* opt, supr_batch** batch);
int supr_batch_configure(supr_batch* batch, double lanes_bias, int32_t max_blocks);
int supr_batch_config(const supr_batch* batch, double* lanes_bias, int32_t* max_blocks);
int supr_batch_solve(supr_batch* batch, const int32_t* ids, int32_t nids, double* const* x_out,
                     double* const* y_out, supr_pdlp_info* infos, double* seconds);
void supr_batch_destroy(supr_batch* batch);

/* ------------------------------------------------------- serving mode -- */
/* Continuous batching: requests are submitted at any time (thread-safe; the
 * host-side preparation runs on the caller's thread) and wait in a FIFO
 * queue; every supr_server_step retires finished instances, admits queued
 * requests whose column, row and nonzero ranges fit the device arenas
 * (capacities fixed at creation: admission by HBM footprint), starts them
 * cold, and launches replays_per_step graph replays of the device-resident
 * loop over all running instances.  A step does not wait for its replays (at
 * most two steps' replays are in flight); admitted requests are uploaded on a
 * stream of their own while the running replays continue, and the running set
 * changes only after the replays in flight have drained.  An instance's block
 * counts and lanes are fixed at admission, so its result is the same alone and
 * in any mix.
 * supr_server_result returns a finished request's point (unscaled PDLP-form
 * coordinates) and information once (*ready = 1; then the result is
 * released); with x_out, y_out and info all NULL it only reports readiness
 * and the sizes m, n.  info->seconds_total is the time from submission to
 * retirement, seconds_solve from admission to retirement.  A request larger
 * than the capacities is refused at submission (SUPR_EINVAL).
 * supr_server_request_bytes gives the device bytes a request occupies while
 * resident; supr_server_device_bytes the bytes the server holds.  (No
 * reference counterpart; callers: request loops.) */
typedef struct supr_server_options {
    supr_pdlp_options pdlp;    /* mode MIXED or FAST_FP64, tol, max_iterations, check_every, graph_attempts, device, verbose */
    int64_t max_cols;          /* column arena: sum over resident requests of n + 1 */
    int64_t max_rows;          /* row arena: sum of m + 1 */
    int64_t max_nnz;           /* nonzero arena: sum of nnz */
    int32_t max_instances;     /* resident requests */
    int32_t max_blocks;        /* blocks per request and side of a launch (0: 1024) */
    double lanes_bias;         /* lanes-per-row rule (0: 1.0) */
    int32_t replays_per_step;  /* graph replays per supr_server_step (0: 1) */
} supr_server_options;
typedef struct supr_server supr_server;
void supr_server_default_options(supr_server_options* opt);
int64_t supr_server_request_bytes(int32_t m, int32_t n, int32_t nnz, int32_t mode);
int supr_server_create(const supr_server_options* opt, supr_server** server);
int supr_server_device_bytes(const supr_server* server, int64_t* bytes);
int supr_server_submit(supr_server* server, const supr_lp* unscaled_form, const supr_lp* scaled_form,
                       const double* Dr, const double* Dc, int64_t* ticket);
int supr_server_step(supr_server* server, int32_t* running, int32_t* queued, int32_t* finished);
int supr_server_result(supr_server* server, int64_t ticket, int32_t* m, int32_t* n, double* x_out,
                       double* y_out, supr_pdlp_info* info, int32_t* ready);
void supr_server_destroy(supr_server* server);

/* Blending LP: K components blended into P products with two quality
 * attributes.  Columns x_kp >= 0 in product-major order (col = p*K + k),
 * n = K*P; rows: P demand equalities sum_k x_kp = d_p (d_p ~ U[50,150]),
 * K capacity rows sum_p x_kp <= 1.1 * (use of k by a random feasible blend),
 * 2P ranged quality rows 0.9 Q_pq d_p <= sum_k a_kq x_kp <= 1.1 Q_pq d_p with
 * a_kq ~ U[0.5,1.5] and Q_pq the quality of that blend; costs U[10,100].
 * m = 3P + K, nnz = 4 K P.  Feasible (the blend) and bounded (x >= 0, c > 0). */
int supr_gen_blend(int32_t K, int32_t P, uint64_t seed, int32_t* row_idx, int32_t* col_idx,
                   double* val, double* c, double* rL, double* rU, double* l, double* u);
int supr_gen_blend_sizes(int32_t count, uint64_t seed, int32_t K_lo, int32_t K_hi, int32_t P_lo,
                         int32_t P_hi, int32_t* K, int32_t* P);

/* ------------------------------------------------------------ MPS ----- */
/* read_mps() + to_ranged_lp() (src/mps_reader.cpp:54-308, include/ranged_lp.hpp:
 * 27-62): the MPS file at path (free format, then fixed format if the free
 * parse fails) as the reference's RangedLP.  Returns a handle that holds the
 * LP; on a parse error, SUPR_EINVAL with "MPS parse failed: ..." (the text of
 * the reference's std::runtime_error).  The COLUMNS body of a free-format
 * file is parsed on several threads (the hardware's, at most 16, at least
 * 1 MiB of body each; the environment variable SUPR_MPS_THREADS sets the
 * count, 1 for the one-thread parse, and SUPR_MPS_MIN_CHUNK the bytes per
 * thread) with the one-thread parse's result; markers, rows not declared
 * before COLUMNS and every error take the one-thread path. */
typedef struct supr_mps supr_mps;
int supr_mps_read(const char* path, supr_mps** out);
/* m, n, nnz of A; q_nnz of the QP's lower triangle (0 for an LP); maximize
 * = 1 when OBJSENSE is MAX (c and Q are then negated, as to_ranged_lp does).
 * Any pointer may be NULL. */
int supr_mps_sizes(const supr_mps* h, int32_t* m, int32_t* n, int32_t* nnz, int32_t* q_nnz,
                   int32_t* maximize);
/* Copies the LP into caller arrays sized by supr_mps_sizes (any may be NULL):
 * A in COO order (row_idx, col_idx, val: nnz), c, l, u (n), rL, rU (m),
 * obj_offset, integer flags (n bytes), Q (q_nnz). */
int supr_mps_get(const supr_mps* h, int32_t* row_idx, int32_t* col_idx, double* val, double* c,
                 double* rL, double* rU, double* l, double* u, double* obj_offset, char* integer,
                 int32_t* q_row, int32_t* q_col, double* q_val);
/* Names held by the handle: kind 0 the problem (NAME), 1 column index, 2 row
 * index, 3 the objective row; NULL when out of range. */
const char* supr_mps_name(const supr_mps* h, int32_t kind, int32_t index);
/* All column (kind 1) or row (kind 2) names at once, each followed by a NUL,
 * in index order: *bytes gets the total size; with buf == NULL nothing is
 * copied (ask for the size first), else buf must hold cap >= *bytes bytes.
 * (No reference counterpart: the reference's RangedLP carries the names;
 * this fills a caller's name table with one call instead of one per name.) */
int supr_mps_names(const supr_mps* h, int32_t kind, char* buf, int64_t cap, int64_t* bytes);
void supr_mps_free(supr_mps* h);
/* Writes lp as a free-format MPS file that supr_mps_read (and the
 * reference's read_mps) reads back to the same LP: row kinds E / L / G from
 * rL, rU (a ranged row as G or L with RANGES, whichever reproduces both
 * bounds; a free row as L with an infinite right-hand side), every column
 * with its objective coefficient, bounds, the objective offset as the
 * objective row's RHS, integer columns between INTORG / INTEND markers
 * (integer may be NULL).  Names: col_names / row_names (no white space), or
 * C<j> / R<i> when NULL.  Values print with 17 significant digits, so they
 * read back exactly; explicit zero coefficients are dropped, as the reader
 * drops them.  (No reference counterpart: tools/gen_lp.cpp:89-131 writes
 * the transportation family this way.) */
int supr_mps_write(const char* path, const supr_lp* lp, const char* name,
                   const char* const* col_names, const char* const* row_names,
                   const char* integer);

#ifdef __cplusplus
}
#endif
#endif /* SUPREMA_PDLP_H */
