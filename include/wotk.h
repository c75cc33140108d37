/*
 * wotk.h -- C ABI of the MI355X unbalanced entropic OT library (libwotk.so).
 *
 * Drop-in boundary for wot's optimal-transport hot path.  The reference is
 * pure Python; its "FFI" is the solver callable that OTModel invokes as
 * ``compute_transport_matrix(solver=self.solver, **config)``
 * (wot/ot/ot_model.py:318 -> wot/ot/optimal_transport.py:30).  Each entry point
 * below replaces one reference call site; the Python package wot_amd binds
 * them with ctypes (wot_amd/_native.py) behind the reference's own signatures.
 *
 * Conventions
 *   - Every pointer argument is DEVICE memory unless stated otherwise.
 *   - Matrices are row-major with an explicit leading dimension (elements).
 *     Single-precision cost matrices must have ld % 4 == 0 and 16-byte
 *     aligned base pointers (the sweep kernels read float4).
 *   - Every call takes a hipStream_t (passed as void*) and enqueues on it.
 *     Solver calls synchronise that stream where the algorithm needs a value
 *     on the host (duality-gap checks) and before returning.
 *   - The library never allocates device memory: scratch space comes in a
 *     caller-owned workspace whose size the matching *_workspace_size()
 *     returns.
 *   - Return value: WOTK_OK (0) or an error code; wotk_last_error() gives the
 *     message of the last failure in the calling thread.
 *   - Calls that share a workspace or outputs must not overlap.
 */
#ifndef WOTK_H
#define WOTK_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WOTK_OK 0
#define WOTK_EINVAL 1     /* bad argument (ValueError on the Python side) */
#define WOTK_EHIP 2       /* HIP runtime failure                          */
#define WOTK_ENAN 3       /* NaN duality gap (optimal_transport.py:162-163) */
#define WOTK_EWORKSPACE 4 /* workspace too small                          */

/* cost modes for wotk_cost_sqeuclid */
#define WOTK_COST_F64 0       /* FP64 direct difference sum_k (x_k - y_k)^2     */
#define WOTK_COST_BF16X3 1    /* split-BF16 3-term MFMA cross term, FP32 accum  */
#define WOTK_COST_BF16X6 2    /* split-BF16 6-term MFMA cross term, FP32 accum  */

typedef struct wotk_stats {
    int64_t iterations;       /* Sinkhorn (a, b) iterations performed            */
    int64_t stabilisations;   /* tau-triggered absorptions                      */
    int64_t gap_checks;       /* duality-gap / dual-evolution checks            */
    double gap;               /* last duality gap (or dual-evolution measure)  */
    double epsilon_final;     /* epsilon the returned plan is built with        */
    int32_t max_iter_reached; /* 1: early un-normalised return (warning)        */
    int32_t nonfinite;        /* 1: a potential became non-finite               */
    double primal;            /* last primal objective (duality-gap solver)     */
    double dual;              /* last dual objective                            */
    int64_t sweeps;           /* row log-sum-exp sweeps launched (2 per iteration) */
    double sweep_ms;          /* device time of the fixed-iteration sweep loop (hipEvents) */
} wotk_stats;

const char *wotk_last_error(void);
int wotk_abi_version(void);

/* ---------------------------------------------------------------------------
 * Cost matrix.  Replaces OTModel.compute_default_cost_matrix
 * (wot/ot/ot_model.py:242-253): C = cdist(X, Y, 'sqeuclidean') / median(C).
 * X: (I, d) and Y: (J, d) float64 row-major (ld = d).  Writes the normalised
 * cost in FP32 twice, C (I x J, ldc) and its transpose CT (J x I, ldct), and the
 * FP64 median to *median_out (device scalar) and, if median_index_out is not
 * NULL, the flat indices (i*J + j) of the middle element(s) as in
 * wotk_median_f64.  mode = WOTK_COST_F64 computes
 * the FP64 direct difference (scipy's formula); WOTK_COST_BF16X3 / _BF16X6
 * compute ||x - m||^2 + ||y - m||^2 - 2 (x - m).(y - m) (m = mean of both
 * days) with the cross term as a 3- or 6-term split-BF16 MFMA GEMM
 * accumulated in FP32 (x = hi + mid + lo in BF16; 3 terms hi.hi + hi.mid +
 * mid.hi, 6 terms add hi.lo + lo.hi + mid.mid).  The median is exact (radix select over the FP64
 * bit patterns; even count -> mean of the two middle values, numpy.median).
 * Layout every sweep over C / CT assumes: ldc and ldct multiples of 4, and
 * columns J .. round_up(J, 4) - 1 of every row of C (I .. round_up(I, 4) - 1
 * of CT) holding finite values -- the sweeps read whole 16-byte groups.  This
 * function and wotk_cost_import write those columns as 0; a caller that
 * builds C / CT itself must do the same.
 * ------------------------------------------------------------------------- */
size_t wotk_cost_workspace_size(int64_t I, int64_t J, int64_t d, int32_t mode);
int wotk_cost_sqeuclid(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                       int32_t mode, float *C, int64_t ldc, float *CT, int64_t ldct,
                       double *median_out, int64_t *median_index_out, void *workspace,
                       size_t workspace_bytes, void *stream);

/* Un-normalised FP64 squared distances (I x J, ldc64), no median: the raw
 * pairwise_distances(..., 'sqeuclidean') of ot_model.py:249-251. */
int wotk_sqeuclid_f64(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                      double *C64, int64_t ldc64, void *stream);

/* Exact median of n FP64 values >= 0 (np.median, ot_model.py:252).
 * median_index_out (device, 2 entries, may be NULL) receives the flat index of
 * an element holding the lower and the upper middle value (equal for odd n):
 * the elements that carry the median's subgradient. */
size_t wotk_median_workspace_size(int64_t n);
int wotk_median_f64(const double *values, int64_t n, double *median_out,
                    int64_t *median_index_out, void *workspace, size_t workspace_bytes,
                    void *stream);

/* A caller-supplied cost (OTModel.compute_transport_map(..., cost_matrix=C),
 * ot_model.py:203-240, used as-is by the solvers) converted to the FP32 pair
 * (C, CT).  C64 is (I x J) with leading dimension ld64, scale multiplies every
 * entry (1.0 for an as-is cost). */
int wotk_cost_import(const double *C64, int64_t I, int64_t J, int64_t ld64, double scale,
                     float *C, int64_t ldc, float *CT, int64_t ldct, void *stream);

/* ---------------------------------------------------------------------------
 * Solvers.  C (I x J, ldc) and CT (J x I, ldct) hold the same FP32 cost.
 * G (I) float64 growth.  plan (I x J, ldp) float64 output, may be NULL;
 * row_sums (I) float64 output of the returned plan, may be NULL (the growth
 * loop's next G, optimal_transport.py:27).  f_out (I) / g_out (J) receive the
 * final dual potentials u + eps log a, v + eps log b, may be NULL.
 * ------------------------------------------------------------------------- */
size_t wotk_solver_workspace_size(int64_t I, int64_t J);

/* transport_stablev2 (wot/ot/optimal_transport.py:167-236), log-domain. */
int wotk_transport_stablev2(const float *C, int64_t ldc, const float *CT, int64_t ldct, int64_t I,
                            int64_t J, const double *G, double lambda1, double lambda2,
                            double epsilon, int64_t scaling_iter, double tau, double epsilon0,
                            int64_t extra_iter, int64_t inner_iter_max, double *plan, int64_t ldp,
                            double *row_sums, double *f_out, double *g_out, void *workspace,
                            size_t workspace_bytes, wotk_stats *stats, void *stream);

/* No counterpart in the reference: the same fixed-iteration solve for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Arguments and outputs are
 * those of wotk_transport_stablev2 without the workspace: the solver state
 * lives in one static device array per GPU of wotk_scratch_capacity() bytes
 * (16 MiB: wotk_solver_workspace_size(I, J) must fit, i.e. I, J up to about
 * 95,000; larger problems are rejected with WOTK_EINVAL).  The call allocates
 * nothing (no hipMalloc / hipFree, no events: stats->sweep_ms is 0); calls that
 * overlap on one GPU share the array and race. */
size_t wotk_scratch_capacity(void);
int wotk_transport_stablev2_noalloc(const float *C, int64_t ldc, const float *CT, int64_t ldct, int64_t I,
                                    int64_t J, const double *G, double lambda1, double lambda2,
                                    double epsilon, int64_t scaling_iter, double tau, double epsilon0,
                                    int64_t extra_iter, int64_t inner_iter_max, double *plan, int64_t ldp,
                                    double *row_sums, double *f_out, double *g_out, wotk_stats *stats,
                                    void *stream);

/* optimal_transport_duality_gap (wot/ot/optimal_transport.py:67-164).  The
 * last epsilon scaling's gap passes run on a second, non-blocking stream that
 * the call creates and destroys (ordered after the caller's stream by an
 * event); the call returns with both drained.  Results and counts are those
 * of the sequential loop. */
int wotk_transport_duality_gap(const float *C, int64_t ldc, const float *CT, int64_t ldct,
                               int64_t I, int64_t J, const double *G, double lambda1,
                               double lambda2, double epsilon, int64_t batch_size,
                               double tolerance, double tau, double epsilon0, int64_t max_iter,
                               double *plan, int64_t ldp, double *row_sums, double *f_out,
                               double *g_out, void *workspace, size_t workspace_bytes,
                               wotk_stats *stats, void *stream);

/* The plan exp((f_i + g_j - C_ij)/eps) / divisor (I x J, ldp, FP64) and/or its
 * row sums from potentials: optimal_transport.py:234-236 (R / J) and :145
 * (divisor 1).  plan or row_sums may be NULL. */
int wotk_plan(const float *C, int64_t ldc, int64_t I, int64_t J, const double *f, const double *g,
              double epsilon, double divisor, double *plan, int64_t ldp, double *row_sums,
              void *stream);

/* ---------------------------------------------------------------------------
 * Batched fixed-iteration solves: every day pair of a time course
 * (ot_model.py:124-201 solves them one after another) times n_combos
 * (epsilon, lambda1) combinations (the reference sweeps a grid as one process
 * per combination), transport_stablev2 semantics for each solve.  All costs
 * live in one FP32 buffer; pair p's C_p and C_p^T start at c_offset /
 * ct_offset (multiples of 4 elements), laid out as wotk_cost_sqeuclid writes
 * them (finite padding columns up to a multiple of 4).  G holds the concatenated growth
 * vectors.  pairs, epsilon and lambda1 are HOST arrays.  f_out / g_out
 * (device, may be NULL) receive the final potentials as consecutive blocks
 * [pair][combination][I_p] (resp. [J_p]).  graph_iters > 0 captures that many
 * iterations in a hipGraph once and replays it.
 * flags & WOTK_BATCH_SHARE_EXP: combinations are taken in consecutive groups
 * of equal epsilon (n_combos = groups x group size, e.g. an epsilon x lambda1
 * grid in epsilon-major order); within a group one exponential per cost entry
 * serves every combination (exp2(y_ref) times a per-column factor); rows
 * whose factored sums could have lost mass to FP32 underflow, or overflowed,
 * are recomputed exactly.  Ignored (the per-combination kernel runs) when the
 * grid has no such grouping.
 * flags & WOTK_BATCH_SHARE_EPS (with WOTK_BATCH_SHARE_EXP, three groups of
 * three, on the LDS-staged tiles; ignored otherwise): groups 0 and 1 reuse the
 * last group's exponent.  At every iteration where k = log2(e)/eps of a group
 * is an integer multiple m >= 1 of the last group's (within 1e-10 relative,
 * for every pair) -- e.g. eps {0.01, 0.05, 0.1} once the eps schedule has
 * converged -- it takes the last group's exponential to the power m (FP32
 * products) instead of its own; otherwise its exponent is r = k_g / k_last
 * times the last group's (ratios up to 2^16), with its own exponential.
 * Compiled (group 1, group 0): (2, 10), (2, 4), (2, r), (1, 1) and (r, r).
 * wotk_batch_modes reports the mode per iteration.
 * stats: sweeps / sweep_ms as for the single solver; stabilisations = rows
 * recomputed exactly (fixed shift or factors out of range); gap_checks =
 * combinations per shared-exponential group (0: one exponential per
 * combination).
 * ------------------------------------------------------------------------- */
#define WOTK_BATCH_SHARE_EXP 1
#define WOTK_BATCH_SHARE_EPS 2

typedef struct wotk_batch_pair {
    int64_t c_offset;   /* element offset of C_p (I_p x J_p, ldc) in the cost buffer */
    int64_t ct_offset;  /* element offset of C_p^T (J_p x I_p, ldct)                 */
    int64_t ldc, ldct;
    int64_t n_i, n_j;
    int64_t g_offset;   /* element offset of G_p (I_p) in the growth buffer         */
} wotk_batch_pair;

size_t wotk_batch_workspace_size(const wotk_batch_pair *pairs, int32_t n_pairs, int32_t n_combos,
                                 int64_t n_iterations);
int wotk_batch_stablev2(const float *cost, const wotk_batch_pair *pairs, int32_t n_pairs,
                        const double *G, int32_t n_combos, const double *epsilon,
                        const double *lambda1, double lambda2, int64_t scaling_iter, double tau,
                        double epsilon0, int64_t extra_iter, int64_t inner_iter_max, double *f_out,
                        double *g_out, int32_t graph_iters, int32_t flags, void *workspace,
                        size_t workspace_bytes, wotk_stats *stats, void *stream);

/* Stepwise form of the same batched solve, with per-pair parameters (the
 * per-pair epsilon / lambda1 / lambda2 of a day-pair configuration,
 * wot/ot/initializer.py:135-183, merged at ot_model.py:231, 238).
 * wotk_batch_begin sets up the run in `workspace` (sized by
 * wotk_batch_workspace_size) and captures the graph; epsilon, lambda1 and
 * lambda2 are HOST arrays [pair][combination].  growth_per_combo = 0: G holds
 * one growth vector per pair (at g_offset); 1: one per (pair, combination),
 * [combination][I_p] starting at g_offset (the learned growth of the
 * previous growth iteration, optimal_transport.py:26-28).
 * wotk_batch_advance runs the next n_iterations of the schedule (3000 + 1000
 * for the reference defaults; clamped to what is left) on `stream` and
 * reports the iterations left.  wotk_batch_potentials copies the current
 * potentials ([pair][combination][I_p] / [J_p], device), the final epsilon of
 * every (pair, combination) (HOST, may be NULL) and the stats (sweep_ms over
 * every advance so far); it synchronises `stream` when stats is given.
 * wotk_batch_end releases the run's host objects (graph, events, streams);
 * the caller keeps `workspace` alive until then.  A run is driven from one
 * host thread at a time. */
typedef struct wotk_batch_run wotk_batch_run;
int wotk_batch_begin(const float *cost, const wotk_batch_pair *pairs, int32_t n_pairs, const double *G,
                     int32_t growth_per_combo, int32_t n_combos, const double *epsilon,
                     const double *lambda1, const double *lambda2, int64_t scaling_iter,
                     double epsilon0, int64_t extra_iter, int64_t inner_iter_max,
                     int32_t graph_iters, int32_t flags, void *workspace, size_t workspace_bytes,
                     wotk_batch_run **run, void *stream);
int wotk_batch_advance(wotk_batch_run *run, int64_t n_iterations, int64_t *remaining, void *stream);
int wotk_batch_potentials(wotk_batch_run *run, double *f_out, double *g_out, double *eps_final,
                          wotk_stats *stats, void *stream);
/* Per iteration of the run's schedule (n >= its iteration count, HOST
 * array): the powers groups 1 (bits 0-7) and 0 (bits 8-15) take of the last
 * group's exponential under WOTK_BATCH_SHARE_EPS (255: the real ratio of
 * their exponents), 0 where each group uses its own exponent.  No counterpart
 * in the reference (it solves each combination alone). */
int wotk_batch_modes(const wotk_batch_run *run, int32_t *modes, int64_t n);
void wotk_batch_end(wotk_batch_run *run);

/* ---------------------------------------------------------------------------
 * transport_stablev2 (optimal_transport.py:167-236) on the cost of
 * ot_model.py:242-253 computed from the coordinates X (I x d), Y (J x d) and
 * never stored: every sweep regenerates C from the coordinates on the matrix
 * cores (split-BF16 cross term, `terms` = 3 or 6 products, FP32 accumulate)
 * with FP64 potentials; the median is the exact median of the FP64 distances.
 * d <= 128.  f_out (I), g_out (J), row_sums (I: row sums of the returned plan
 * R / J, the growth loop's next G), median_out (1): device, may be NULL.
 * ------------------------------------------------------------------------- */
size_t wotk_fused_workspace_size(int64_t I, int64_t J, int64_t d);
int wotk_transport_stablev2_coords(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                                   const double *G, double lambda1, double lambda2, double epsilon,
                                   int64_t scaling_iter, double tau, double epsilon0,
                                   int64_t extra_iter, int64_t inner_iter_max, int32_t terms,
                                   double *f_out, double *g_out, double *row_sums,
                                   double *median_out, void *workspace, size_t workspace_bytes,
                                   wotk_stats *stats, void *stream);

/* optimal_transport_duality_gap (wot/ot/optimal_transport.py:67-164) on the
 * never-stored cost of wotk_transport_stablev2_coords: the same schedule as
 * wotk_transport_duality_gap (epsilon scalings, dual-evolution checks, tau
 * stabilisation counts, batches of batch_size on the last scaling, max_iter
 * early return with un-normalised row sums, NaN gap -> WOTK_ENAN), every sweep
 * regenerating C from X, Y on the matrix cores; the gap pass forms the plan's
 * row and column sums with FP64 exponentials over the same regenerated C.
 * Outputs as wotk_transport_stablev2_coords; stats carry iterations,
 * stabilisations, gap checks, the final gap, primal and dual. */
size_t wotk_duality_gap_coords_workspace_size(int64_t I, int64_t J, int64_t d);
int wotk_transport_duality_gap_coords(const double *X, const double *Y, int64_t I, int64_t J, int64_t d,
                                      const double *G, double lambda1, double lambda2, double epsilon,
                                      int64_t batch_size, double tolerance, double tau, double epsilon0,
                                      int64_t max_iter, int32_t terms, double *f_out, double *g_out,
                                      double *row_sums, double *median_out, void *workspace,
                                      size_t workspace_bytes, wotk_stats *stats, void *stream);

/* Primal / dual objectives of optimal_transport.py:45-62 at the plan
 * R = exp((f_i + g_j - C_ij)/eps) (un-normalised), FP64 reductions.
 * out[0] = primal, out[1] = dual (device or host pointer to 2 doubles: host). */
int wotk_primal_dual(const float *C, int64_t ldc, const float *CT, int64_t ldct, int64_t I,
                     int64_t J, const double *G, const double *f, const double *g,
                     double lambda1, double lambda2, double epsilon, double *out_host,
                     void *workspace, size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Long-range coupling chain (wot/tmap/util.py:74-94 glue_transport_maps,
 * np.dot(tmap_0.X, tmap_1.X[idx]); chained by wot/tmap/chaining.py:4-40):
 * C (M x N, ldc) = A (M x K, lda) * B (K x N, ldb), FP32 row-major on the
 * matrix cores (MFMA), FP32 accumulation.
 * ------------------------------------------------------------------------- */
int wotk_gemm_f32(const float *A, int64_t lda, const float *B, int64_t ldb, float *C, int64_t ldc,
                  int64_t M, int64_t N, int64_t K, void *stream);

/* ---------------------------------------------------------------------------
 * Adjoint (no counterpart in the reference: the gradient of its primal
 * objective, optimal_transport.py:45-53, at the converged plan).  Given the
 * converged potentials f, g of a solve on C = sqeuclid(X, Y) / median, returns
 * d OT / d X (I x d) and d OT / d log g (I) where G = g^dt
 * (ot_model.py:314-315), by the envelope theorem with
 * dP/dC_ij = (R_ij - exp(-C_ij/eps)) / (I J), including the median's
 * subgradient through the middle element(s) median_index (may be NULL: the
 * median is then treated as a constant).  grad_x or grad_log_g may be NULL.
 * ------------------------------------------------------------------------- */
size_t wotk_adjoint_workspace_size(int64_t I, int64_t J, int64_t d);
int wotk_ot_adjoint(const float *C, int64_t ldc, const float *CT, int64_t ldct, int64_t I,
                    int64_t J, const double *X, const double *Y, int64_t d, const double *G,
                    double dt, const double *f, const double *g, double lambda1, double lambda2,
                    double epsilon, const double *median, const int64_t *median_index,
                    double *grad_x, double *grad_log_g, void *workspace, size_t workspace_bytes,
                    void *stream);

/* ---------------------------------------------------------------------------
 * Populations through stored transport maps (TransportMapModel,
 * wot/tmap/transport_map_model.py):
 *   WOTK_PUSH_FORWARD  p <- p T        push_forward, :284-293
 *   WOTK_PULL_BACK     p <- (T p^T)^T  pull_back,    :350-359
 * maps[0..n_maps-1] are applied in that order (a pull back from day i to
 * day i-2 passes T_{i-1,i} first, then T_{i-2,i-1}); map s is row-major
 * (rows = source-day cells, cols = target-day cells), FP64 or FP32 entries,
 * any leading dimension.  Consecutive maps must chain (push: cols of map s =
 * rows of map s+1; pull: rows of map s = cols of map s+1).
 * Populations are a CELLS x POPULATIONS matrix (ld ldp >= n_pops), i.e. the
 * transpose of the reference's np.vstack([pop.p ...]).  With normalize != 0
 * every population is divided by its sum after every map (the reference's
 * normalize=True); a population summing to 0 becomes NaN there, as in numpy.
 * outs is a HOST array of n_maps device pointers: outs[s] (may be NULL,
 * except the last) receives the populations after map s as an n_s x n_pops
 * matrix with leading dimension ldo.  Products and sums are FP64 (matrix
 * cores); the split-reduction order is fixed, so results are deterministic.
 * A pull back through a map with a transposed copy streams the copy with the
 * push kernel (rows of the copy along the lanes, fully coalesced); without
 * one it reads the map's rows directly.  Both give the same sums up to
 * rounding order.
 * stats (host, may be NULL): when given, every map's product kernel is timed
 * with hipEvents and the call synchronises the stream before returning;
 * sweeps = maps applied, sweep_ms = summed product-kernel time.
 * ------------------------------------------------------------------------- */
#define WOTK_DTYPE_F64 0
#define WOTK_DTYPE_F32 1
#define WOTK_PUSH_FORWARD 0
#define WOTK_PULL_BACK 1

typedef struct wotk_tmap {
    const void *data;       /* device, row-major rows x cols, leading dimension ld      */
    int32_t dtype;          /* WOTK_DTYPE_F64 or WOTK_DTYPE_F32                         */
    int32_t reserved;       /* 0                                                        */
    int64_t ld, rows, cols;
    const void *transposed; /* optional: the same map as cols x rows (same dtype), or NULL */
    int64_t ld_t;           /* its leading dimension                                    */
} wotk_tmap;

/* 0 when the arguments are invalid (wotk_tmap_propagate then reports why) */
size_t wotk_tmap_workspace_size(const wotk_tmap *maps, int32_t n_maps, int32_t direction,
                                int32_t n_pops);
int wotk_tmap_propagate(const wotk_tmap *maps, int32_t n_maps, int32_t direction,
                        const double *pops, int64_t ldp, int32_t n_pops, int32_t normalize,
                        double *const *outs, int64_t ldo, void *workspace, size_t workspace_bytes,
                        wotk_stats *stats, void *stream);

/* dst (cols x rows, ld_dst) = src^T (src rows x cols, ld_src), FP64 or FP32
 * (dtype): the transposed copy of a map for wotk_tmap.transposed. */
int wotk_tmap_transpose(const void *src, int32_t dtype, int64_t ld_src, int64_t rows, int64_t cols,
                        void *dst, int64_t ld_dst, void *stream);

/* X[i, 0:k] /= sum(X[i, 0:k]) for i < n (the fates matrix normalisation,
 * transport_map_model.py:66 X /= X.sum(axis=1, keepdims=1)); a row summing
 * to 0 becomes NaN, as in numpy. */
int wotk_rows_normalize_f64(double *X, int64_t n, int32_t k, int64_t ld, void *stream);

/* ---------------------------------------------------------------------------
 * Local PCA (wot/ot/util.py:240-255 compute_pca, used by ot_model.py:296-301):
 * the stacked cells of both days x (n cells x g genes, FP64 or FP32 by dtype,
 * leading dimension ldx) centered per gene and then per cell (sklearn's PCA is
 * fitted on x^T and centers its columns):
 *   A[c][j] = x[c][j] - gene_mean[j] - cell_mean[c]
 * gene_mean (g) is the reference's mean_shift; cell_mean (n) is the per-cell
 * mean of the gene-centered data.  A (n x g, lda, FP64) may be NULL (means
 * only).  The singular vectors of A give the PCA coordinates.
 * ------------------------------------------------------------------------- */
size_t wotk_pca_center_workspace_size(int64_t n, int64_t g);
int wotk_pca_center(const void *x, int32_t dtype, int64_t ldx, int64_t n, int64_t g, double *A,
                    int64_t lda, double *gene_mean, double *cell_mean, void *workspace,
                    size_t workspace_bytes, void *stream);

/* ---------------------------------------------------------------------------
 * Interpolated populations for OT validation (wot/ot/util.py:109-200).
 * wotk_interp_weights: W (I x J, contiguous) = T / colsum(T)^(1 - frac)
 *   (util.py:142; colsum (J) receives the column sums).
 * wotk_interp_sample: idx[s] = first k < n with cdf[k] > u[s] (numpy's
 *   cdf.searchsorted(u, side='right') inside RandomState.choice); u are the
 *   caller's uniforms, so a caller drawing them from numpy's stream gets the
 *   reference's samples.
 * wotk_interp_gather: out[s] = p0[idx/J] (1 - frac) + p1[idx%J] frac (d columns).
 * ------------------------------------------------------------------------- */
int wotk_interp_weights(const double *T, int64_t ld, int64_t I, int64_t J, double frac, double *W,
                        double *colsum, void *stream);
int wotk_interp_sample(const double *cdf, int64_t n, const double *u, int64_t size, int64_t *idx,
                       void *stream);
int wotk_interp_gather(const int64_t *idx, int64_t size, const double *p0, int64_t ld0, int64_t I,
                       const double *p1, int64_t ld1, int64_t J, int64_t d, double frac, double *out,
                       int64_t ldo, void *stream);

/* ---------------------------------------------------------------------------
 * Exact earth mover's distance (POT's emd2 as wot/ot/util.py:202-237 calls it,
 * numItermax): min sum_ij F_ij C_ij over F >= 0 with row sums a (n) and column
 * sums b (m), by a primal network simplex.  HOST pointers (the only host-side
 * entry point: exact pivoting is sequential; the device builds C).  C is
 * n x m with leading dimension ldc; flow_out (n x m, may be NULL) receives the
 * optimal plan; max_iter <= 0: no limit.  *status_out: 0 optimal, 1 stopped at
 * max_iter (the plan so far: part of the mass may still sit on the artificial
 * start arcs, so its marginals can fall short).  a and b must be
 * non-negative with equal sums (relative 1e-9).
 * ------------------------------------------------------------------------- */
int wotk_emd(const double *a, int64_t n, const double *b, int64_t m, const double *C, int64_t ldc,
             int64_t max_iter, double *cost_out, double *flow_out, int64_t *iterations_out,
             int32_t *status_out);

/* ---------------------------------------------------------------------------
 * Tile shapes of the sweep kernels (no counterpart in the reference, which has
 * no device kernels).  A tile is (rows of the cost matrix per workgroup,
 * threads per workgroup).  The library keeps one tile per (kernel, shape
 * bucket); launches look the tile up by the bucket of their shape.  The
 * project's autotuner (wot_amd.tune) times every variant on the caller's
 * workload and stores the fastest here; unset buckets use the built-in
 * default.  Settings are process-wide and may be changed between calls (not
 * during one).
 *   WOTK_KERNEL_ROW_SWEEP:   one stored-cost half-iteration (transport_stablev2,
 *                            duality-gap solver); bucket from (rows, columns)
 *                            of the swept matrix.
 *   WOTK_KERNEL_BATCH_SWEEP: the time-course half-iteration; bucket from the
 *                            number of combinations and the mean row length
 *                            sum(2 I J) / sum(I + J) over the pairs.
 *   WOTK_KERNEL_BATCH_SHARED: the same with WOTK_BATCH_SHARE_EXP; same buckets.
 * ------------------------------------------------------------------------- */
#define WOTK_KERNEL_ROW_SWEEP 0
#define WOTK_KERNEL_BATCH_SWEEP 1
#define WOTK_KERNEL_BATCH_SHARED 2

typedef struct wotk_tile {
    int32_t rows;    /* cost-matrix rows per workgroup  */
    int32_t threads; /* threads per workgroup (wave64s) */
} wotk_tile;

/* number of compiled variants of `kernel`; copies up to `capacity` into out */
int32_t wotk_tile_variants(int32_t kernel, wotk_tile *out, int32_t capacity);
/* shape bucket of a launch: ROW_SWEEP (n_rows, n_cols); BATCH_SWEEP
 * (n_combos, mean row length in n_cols; n_rows ignored).  -1 if kernel unknown */
int32_t wotk_shape_bucket(int32_t kernel, int64_t n_rows, int64_t n_cols, int32_t n_combos);
int32_t wotk_num_buckets(int32_t kernel);
/* tile for one bucket, or every bucket with bucket = -1; tile NULL restores the
 * default; a tile that is not a compiled variant is rejected (WOTK_EINVAL) */
int wotk_set_tile(int32_t kernel, int32_t bucket, const wotk_tile *tile);
int wotk_get_tile(int32_t kernel, int32_t bucket, wotk_tile *out);

/* ---------------------------------------------------------------------------
 * Instruction-rate probe (measurement only; no counterpart in the reference):
 * the issue rate of the instructions the compute-bound batched sweep is built
 * from, for its roofline.  rates_host (HOST, >= 7 entries) receives, in
 * operations per second over the whole device: [0] v_exp_f32, [1] FP64 FMA,
 * [2] FP32 FMA, [3] FP64 add + FP64->FP32 + FP32->FP64 converts, [4] FP64 round-to-nearest (+ add),
 * [5] v_ldexp_f32, [6] the sweep's exact split exponential of an FP64
 * exponent (one whole entry: 2 FP64 FMAs, clamp, rint, subtract, two
 * converts, exp2, ldexp).  sink: device scratch of wotk_probe_sink_bytes()
 * bytes (never written in practice).
 * ------------------------------------------------------------------------- */
size_t wotk_probe_sink_bytes(void);
int wotk_probe_valu_rates(double *rates_host, int32_t n_rates, void *sink, size_t sink_bytes, void *stream);
/* Issue rate (lane-operations per second over the device) of an FP32 mix of
 * 29 FMAs and one v_exp_f32 per step, about the batched sweep's ratio, at 1,
 * 2, 4 and 8 waves per SIMD: rates_host (HOST, >= 4 entries).  Same sink. */
int wotk_probe_issue_mix(double *rates_host, int32_t n_rates, void *sink, size_t sink_bytes, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* WOTK_H */
