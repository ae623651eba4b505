/*
 * numbagg_amd — C ABI of the MI355X (gfx950) backend for numbagg's hot path.
 *
 * The reference (numbagg) has no native code: each public function is a
 * numba-compiled NumPy gufunc whose per-row loop lives in a Python kernel
 * (numbagg/moving.py, moving_exp.py, grouped.py, funcs.py) and whose argument
 * handling lives in the decorator classes of numbagg/decorators.py. Each entry
 * point below replaces one such gufunc: the caller (the Python host layer in
 * numbagg_amd/, which re-states the decorator semantics) hands over device
 * pointers for a row-major (rows, n) view whose core axis is contiguous, plus
 * the scalar arguments the gufunc received.
 *
 * Conventions shared by every entry point:
 *   - pointers are device pointers; rows/n/ld_* are element counts; ld_* is
 *     the distance between consecutive rows (>= n for rows > 1);
 *   - `stream` is a hipStream_t (NULL = the default stream); calls are
 *     asynchronous with respect to the host and never allocate device memory;
 *   - `ws`/`ws_bytes` is caller-owned scratch sized by the matching
 *     *_workspace_bytes() query (its contents need not be initialised);
 *   - the return value is NBG_OK or an NBG_ERR_* code; nbg_last_error() then
 *     returns a message (thread-local).
 *   - dtype codes: NBG_F32/NBG_F64 for values, NBG_I32/NBG_I64 also accepted
 *     where the reference compiles integer loops, NBG_I32/NBG_I64 for labels.
 */
#ifndef NUMBAGG_AMD_H
#define NUMBAGG_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NBG_OK 0
#define NBG_ERR_INVALID 1  /* invalid argument (shape, window, dtype, workspace) */
#define NBG_ERR_HIP 2      /* HIP runtime error */
#define NBG_ERR_INJECTED 3 /* failure forced through NBG_INJECT_ERROR (tests) */

#define NBG_F32 0
#define NBG_F64 1
#define NBG_I32 2
#define NBG_I64 3

/* ---- library ----------------------------------------------------------- */
const char* nbg_version(void);
const char* nbg_last_error(void);
/* Number of bounded device-side waits that expired since the last reset on
 * the current device. Always 0 unless a cross-workgroup protocol was broken. */
int nbg_spin_timeouts(unsigned int* count);
int nbg_reset_spin_timeouts(void);

/* Tuning knobs of this host thread (no reference counterpart): tile shapes,
 * grid caps and path choices the autotuner picks per call shape, by NBG_* name
 * (the same names are read from the environment when no override is set).
 * Overrides are thread-local, so concurrent callers do not see each other's
 * choices; results never depend on them beyond the float rounding the parity
 * bars allow. nbg_tune_get returns the effective value (override, environment,
 * else dflt). */
int nbg_tune_set(const char* name, int64_t value);
int nbg_tune_clear(void);
int64_t nbg_tune_get(const char* name, int64_t dflt);

/* Streaming copy of `bytes` (16-byte aligned, multiple of 16). No reference
 * counterpart: a calibration kernel for bandwidth and traffic measurements. */
int nbg_stream_copy(const void* src, void* dst, size_t bytes, void* stream);

/* ---- fill (numbagg/funcs.py:294-326 bfill/ffill; ndfill.__call__ at
 *      numbagg/decorators.py:465-487) ------------------------------------- */
/* limit: the reference's `lives_remaining` budget (>= 0; None -> n).
 * Workspace: per-tile summaries (two-pass path, 8 rows or more) or
 * look-back status words and tickets (ordered single pass); the query and the
 * call choose the same path for the same arguments and environment. */
size_t nbg_fill_workspace_bytes(int64_t rows, int64_t n, int dtype);
int nbg_ffill(int dtype, const void* x, void* out, int64_t rows, int64_t n,
              int64_t ld_x, int64_t ld_out, int64_t limit, void* ws,
              size_t ws_bytes, void* stream);
int nbg_bfill(int dtype, const void* x, void* out, int64_t rows, int64_t n,
              int64_t ld_x, int64_t ld_out, int64_t limit, void* ws,
              size_t ws_bytes, void* stream);

/* ---- moving windows (numbagg/moving.py:12-275; ndmove.__call__ at
 *      numbagg/decorators.py:305-341) ------------------------------------- */
/* out[i] = statistic of the valid inputs in [i-window+1, i] (pairwise-complete
 * for cov/corr), NaN when fewer than the effective min_count (sum: min_count;
 * mean/corr: max(min_count,1); std/var/cov: max(min_count,2)).
 * dtype: NBG_F32 or NBG_F64 for inputs and output; 0 < window <= n. */
int nbg_move_sum(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                 int64_t ld_x, int64_t ld_out, int64_t window,
                 int64_t min_count, void* stream);     /* moving.py:58-99 */
int nbg_move_mean(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                  int64_t ld_x, int64_t ld_out, int64_t window,
                  int64_t min_count, void* stream);    /* moving.py:12-55 */
int nbg_move_std(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                 int64_t ld_x, int64_t ld_out, int64_t window,
                 int64_t min_count, void* stream);     /* moving.py:123-151 */
int nbg_move_var(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                 int64_t ld_x, int64_t ld_out, int64_t window,
                 int64_t min_count, void* stream);     /* moving.py:154-181 */
int nbg_move_cov(int dtype, const void* a, const void* b, void* out,
                 int64_t rows, int64_t n, int64_t ld_a, int64_t ld_b,
                 int64_t ld_out, int64_t window, int64_t min_count,
                 void* stream);                        /* moving.py:184-220 */
int nbg_move_corr(int dtype, const void* a, const void* b, void* out,
                  int64_t rows, int64_t n, int64_t ld_a, int64_t ld_b,
                  int64_t ld_out, int64_t window, int64_t min_count,
                  void* stream);                       /* moving.py:223-275 */

/* ---- exponentially weighted windows (numbagg/moving_exp.py:12-335;
 *      ndmoveexp.__call__ at numbagg/decorators.py:374-414) --------------- */
/* State is float64, as in the reference. `dtype` is the loop dtype (NBG_F32 or
 * NBG_F64): alpha values, min_weight comparison and output use it. `x_dtype`
 * is the input dtype: equal to dtype, or NBG_F32 with dtype NBG_F64 (float32
 * data on the float64 loop, widened in the kernel, which is what NumPy's loop
 * selection does for float32 data with a Python-float alpha).
 * alpha: NULL -> the scalar alpha_scalar for every position; otherwise an
 * array in the loop dtype with row stride ld_alpha (0 = one row shared by all
 * rows). */
size_t nbg_move_exp_workspace_bytes(int64_t rows, int64_t n, int x_dtype);
int nbg_move_exp_nancount(int x_dtype, int dtype, const void* x,
                          const void* alpha, int64_t ld_alpha,
                          double alpha_scalar, void* out, int64_t rows,
                          int64_t n, int64_t ld_x, int64_t ld_out,
                          double min_weight, void* ws, size_t ws_bytes,
                          void* stream); /* moving_exp.py:12-38 */
int nbg_move_exp_nanmean(int x_dtype, int dtype, const void* x,
                         const void* alpha, int64_t ld_alpha,
                         double alpha_scalar, void* out, int64_t rows,
                         int64_t n, int64_t ld_x, int64_t ld_out,
                         double min_weight, void* ws, size_t ws_bytes,
                         void* stream); /* moving_exp.py:41-72 */
int nbg_move_exp_nansum(int x_dtype, int dtype, const void* x,
                        const void* alpha, int64_t ld_alpha,
                        double alpha_scalar, void* out, int64_t rows,
                        int64_t n, int64_t ld_x, int64_t ld_out,
                        double min_weight, void* ws, size_t ws_bytes,
                        void* stream); /* moving_exp.py:75-103 */
int nbg_move_exp_nanvar(int x_dtype, int dtype, const void* x,
                        const void* alpha, int64_t ld_alpha,
                        double alpha_scalar, void* out, int64_t rows,
                        int64_t n, int64_t ld_x, int64_t ld_out,
                        double min_weight, void* ws, size_t ws_bytes,
                        void* stream); /* moving_exp.py:106-156 */
int nbg_move_exp_nanstd(int x_dtype, int dtype, const void* x,
                        const void* alpha, int64_t ld_alpha,
                        double alpha_scalar, void* out, int64_t rows,
                        int64_t n, int64_t ld_x, int64_t ld_out,
                        double min_weight, void* ws, size_t ws_bytes,
                        void* stream); /* moving_exp.py:159-224 */
int nbg_move_exp_nancov(int x_dtype, int dtype, const void* a, const void* b,
                        const void* alpha, int64_t ld_alpha,
                        double alpha_scalar, void* out, int64_t rows,
                        int64_t n, int64_t ld_a, int64_t ld_b, int64_t ld_out,
                        double min_weight, void* ws, size_t ws_bytes,
                        void* stream); /* moving_exp.py:227-273 */
int nbg_move_exp_nancorr(int x_dtype, int dtype, const void* a, const void* b,
                         const void* alpha, int64_t ld_alpha,
                         double alpha_scalar, void* out, int64_t rows,
                         int64_t n, int64_t ld_a, int64_t ld_b, int64_t ld_out,
                         double min_weight, void* ws, size_t ws_bytes,
                         void* stream); /* moving_exp.py:276-335 */

/* ---- grouped reductions (numbagg/grouped.py:7-270; groupndreduce.__call__
 *      at numbagg/decorators.py:558-674) ---------------------------------- */
/* values: (rows, n) with row stride ld_v; labels: (n,) shared by all rows,
 * int32 or int64; out: (rows, num_labels) contiguous, in the values dtype.
 * Labels < 0 are skipped, and so are labels >= num_labels.
 * vdtype: NBG_F32/NBG_F64, or NBG_I32/NBG_I64 for every function but mean,
 * var and std (which take floats; the host widens integer input to float64 as
 * the reference does). Integer sums, sums of squares and products wrap like
 * the reference's fixed-width integer arithmetic; float ones are formed in
 * float64 and rounded once.
 * strategy: per-label accumulators privatised in LDS, or global atomics. */
#define NBG_GROUP_AUTO 0
#define NBG_GROUP_LDS 1
#define NBG_GROUP_GLOBAL 2
size_t nbg_group_workspace_bytes(int vdtype, int64_t rows, int64_t n,
                                 int64_t num_labels, int strategy);
int nbg_group_nansum(int vdtype, const void* values, int ldtype,
                     const void* labels, void* out, int64_t rows, int64_t n,
                     int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                     size_t ws_bytes, void* stream); /* grouped.py:30-40 */
int nbg_group_nanmean(int vdtype, const void* values, int ldtype,
                      const void* labels, void* out, int64_t rows, int64_t n,
                      int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                      size_t ws_bytes, void* stream); /* grouped.py:7-27 */
int nbg_group_nancount(int vdtype, const void* values, int ldtype,
                       const void* labels, void* out, int64_t rows, int64_t n,
                       int64_t ld_v, int64_t num_labels, int strategy,
                       void* ws, size_t ws_bytes,
                       void* stream); /* grouped.py:43-51 */
int nbg_group_nanargmax(int vdtype, const void* values, int ldtype,
                        const void* labels, void* out, int64_t rows, int64_t n,
                        int64_t ld_v, int64_t num_labels, int strategy,
                        void* ws, size_t ws_bytes,
                        void* stream); /* grouped.py:54-75 */
int nbg_group_nanargmin(int vdtype, const void* values, int ldtype,
                        const void* labels, void* out, int64_t rows, int64_t n,
                        int64_t ld_v, int64_t num_labels, int strategy,
                        void* ws, size_t ws_bytes,
                        void* stream); /* grouped.py:78-92 */
int nbg_group_nanfirst(int vdtype, const void* values, int ldtype,
                       const void* labels, void* out, int64_t rows, int64_t n,
                       int64_t ld_v, int64_t num_labels, int strategy,
                       void* ws, size_t ws_bytes,
                       void* stream); /* grouped.py:95-110 */
int nbg_group_nanlast(int vdtype, const void* values, int ldtype,
                      const void* labels, void* out, int64_t rows, int64_t n,
                      int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                      size_t ws_bytes, void* stream); /* grouped.py:113-121 */
int nbg_group_nanvar(int vdtype, const void* values, int ldtype,
                     const void* labels, void* out, int64_t rows, int64_t n,
                     int64_t ld_v, int64_t num_labels, int64_t ddof,
                     int strategy, void* ws, size_t ws_bytes,
                     void* stream); /* grouped.py:148-176 */
int nbg_group_nanstd(int vdtype, const void* values, int ldtype,
                     const void* labels, void* out, int64_t rows, int64_t n,
                     int64_t ld_v, int64_t num_labels, int64_t ddof,
                     int strategy, void* ws, size_t ws_bytes,
                     void* stream); /* grouped.py:179-208 */

int nbg_group_nanmin(int vdtype, const void* values, int ldtype,
                     const void* labels, void* out, int64_t rows, int64_t n,
                     int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                     size_t ws_bytes, void* stream); /* grouped.py:211-225 */
int nbg_group_nanmax(int vdtype, const void* values, int ldtype,
                     const void* labels, void* out, int64_t rows, int64_t n,
                     int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                     size_t ws_bytes, void* stream); /* grouped.py:228-244 */
int nbg_group_nanprod(int vdtype, const void* values, int ldtype,
                      const void* labels, void* out, int64_t rows, int64_t n,
                      int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                      size_t ws_bytes, void* stream); /* grouped.py:124-132 */
int nbg_group_nansum_of_squares(int vdtype, const void* values, int ldtype,
                                const void* labels, void* out, int64_t rows,
                                int64_t n, int64_t ld_v, int64_t num_labels,
                                int strategy, void* ws, size_t ws_bytes,
                                void* stream); /* grouped.py:135-145 */
int nbg_group_nanany(int vdtype, const void* values, int ldtype,
                     const void* labels, void* out, int64_t rows, int64_t n,
                     int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                     size_t ws_bytes, void* stream); /* grouped.py:247-257 */
int nbg_group_nanall(int vdtype, const void* values, int ldtype,
                     const void* labels, void* out, int64_t rows, int64_t n,
                     int64_t ld_v, int64_t num_labels, int strategy, void* ws,
                     size_t ws_bytes, void* stream); /* grouped.py:260-270 */

/* ---- plain reductions (numbagg/funcs.py:23-228; ndaggregate.__call__ at
 *      numbagg/decorators.py:231-256, ndreduce.__call__ at :1002-1028) ----- */
/* x: (rows, n) with row stride ld; one output per row. dtype: NBG_F32,
 * NBG_F64, NBG_I32 or NBG_I64 (mean/var/std: floats only). Outputs: sum in
 * the input dtype (integers wrap); mean/var/std in the input dtype; count,
 * argmax, argmin as int64 (argmax/argmin: -1 for a row without a valid value,
 * which the host reports as the reference's ValueError); allnan/anynan as
 * one byte (0/1); max/min in the input dtype, int64 for integers. n == 0 is
 * rejected for max/min/argmax/argmin (the reference raises). Deterministic:
 * fixed-order merges, no atomics. */
size_t nbg_reduce_workspace_bytes(int64_t rows, int64_t n, int dtype);
int nbg_nansum(int dtype, const void* x, void* out, int64_t rows, int64_t n,
               int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:69-82 */
int nbg_nanmean(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:85-102 */
int nbg_nancount(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                 int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:53-66 */
int nbg_allnan(int dtype, const void* x, void* out, int64_t rows, int64_t n,
               int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:23-35 */
int nbg_anynan(int dtype, const void* x, void* out, int64_t rows, int64_t n,
               int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:38-50 */
int nbg_nanvar(int dtype, const void* x, void* out, int64_t rows, int64_t n,
               int64_t ld, int64_t ddof, void* ws, size_t ws_bytes,
               void* stream); /* funcs.py:105-130 */
int nbg_nanstd(int dtype, const void* x, void* out, int64_t rows, int64_t n,
               int64_t ld, int64_t ddof, void* ws, size_t ws_bytes,
               void* stream); /* funcs.py:133-157 */
int nbg_nanmax(int dtype, const void* x, void* out, int64_t rows, int64_t n,
               int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:200-216 */
int nbg_nanmin(int dtype, const void* x, void* out, int64_t rows, int64_t n,
               int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:219-234 */
int nbg_nanargmax(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                  int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:160-178 */
int nbg_nanargmin(int dtype, const void* x, void* out, int64_t rows, int64_t n,
                  int64_t ld, void* ws, size_t ws_bytes, void* stream); /* funcs.py:181-197 */

/* ---- label factorisation (the caller-side step before the grouped
 *      reductions: numbagg/decorators.py:570-573 requires factorised labels,
 *      numbagg/test/conftest.py:175-189 factorises with pd.factorize(sort=True))
 * codes[i] = position of values[i] among the uniques (-1 for NaN); uniques
 * ascending (sort != 0) or in order of first appearance; -0.0 and +0.0 are one
 * value, represented by whichever comes first. Two calls: _count leaves the
 * number of uniques in *num_uniques (device memory; the caller reads it and
 * allocates the uniques), _finish writes codes (int64, n) and uniques (values
 * dtype). The workspace must be left untouched between the two calls and
 * num_uniques must be the count _count wrote for the same values: on a
 * mismatch _finish writes nothing; a value absent from the table (values
 * changed in between) gets code -1. dtype: NBG_F32, NBG_F64, NBG_I32, NBG_I64. */
size_t nbg_factorize_workspace_bytes(int64_t n);
int nbg_factorize_count(int dtype, const void* values, int64_t n, int sort,
                        void* ws, size_t ws_bytes, int64_t* num_uniques,
                        void* stream);
int nbg_factorize_finish(int dtype, const void* values, int64_t n,
                         int64_t num_uniques, int sort, void* codes,
                         void* uniques, void* ws, size_t ws_bytes,
                         void* stream);

/* ---- correlation / covariance matrices (numbagg/funcs.py:338-532;
 *      numbagg/moving_matrix.py:16-432; ndmatrix / ndmovematrix /
 *      ndmoveexpmatrix at numbagg/decorators.py:677-818, :1031-1112)
 * Pairwise complete: an observation counts for (i, j) where both variables
 * are valid. float64 sums of values shifted by a per-variable constant.
 * Static: x (batch, nvars, nobs) contiguous -> out (batch, nvars, nvars),
 * NaN where a pair has fewer than 2 observations (corr also where a variance
 * is not positive); nobs must be > 0 (every pair is NaN otherwise). Moving and
 * exponential: x (batch, nobs, nvars) contiguous -> out (batch, nobs, nvars,
 * nvars). dtype NBG_F32 or NBG_F64; results are deterministic. */
size_t nbg_matrix_workspace_bytes(int64_t batch, int64_t nvars, int64_t nobs);
int nbg_nancorrmatrix(int dtype, const void* x, void* out, int64_t batch,
                      int64_t nvars, int64_t nobs, void* ws, size_t ws_bytes,
                      void* stream); /* funcs.py:338-441 */
int nbg_nancovmatrix(int dtype, const void* x, void* out, int64_t batch,
                     int64_t nvars, int64_t nobs, void* ws, size_t ws_bytes,
                     void* stream); /* funcs.py:444-532 */
/* min_count as the caller passes it (>= 0); the kernels apply the reference's
 * max(min_count, 1) and, for corr, the two-observation minimum. */
size_t nbg_move_matrix_workspace_bytes(int64_t batch, int64_t nobs,
                                       int64_t nvars, int64_t window);
int nbg_move_corrmatrix(int dtype, const void* x, void* out, int64_t batch,
                        int64_t nobs, int64_t nvars, int64_t window,
                        int64_t min_count, void* ws, size_t ws_bytes,
                        void* stream); /* moving_matrix.py:16-114 */
int nbg_move_covmatrix(int dtype, const void* x, void* out, int64_t batch,
                       int64_t nobs, int64_t nvars, int64_t window,
                       int64_t min_count, void* ws, size_t ws_bytes,
                       void* stream); /* moving_matrix.py:117-201 */
/* x_dtype / dtype as for the exponential windows above (float32 data on the
 * float64 loop allowed); alpha: loop dtype, (batch, nobs) with row stride
 * ld_alpha (0: one row shared by the batch); min_weight already rounded to
 * the loop dtype. */
size_t nbg_move_exp_matrix_workspace_bytes(int64_t batch, int64_t nobs,
                                           int64_t nvars);
int nbg_move_exp_nancorrmatrix(int x_dtype, int dtype, const void* x,
                               const void* alpha, int64_t ld_alpha,
                               double min_weight, void* out, int64_t batch,
                               int64_t nobs, int64_t nvars, void* ws,
                               size_t ws_bytes,
                               void* stream); /* moving_matrix.py:204-328 */
int nbg_move_exp_nancovmatrix(int x_dtype, int dtype, const void* x,
                              const void* alpha, int64_t ld_alpha,
                              double min_weight, void* out, int64_t batch,
                              int64_t nobs, int64_t nvars, void* ws,
                              size_t ws_bytes,
                              void* stream); /* moving_matrix.py:331-432 */

/* No counterpart in the reference. Allocation-free variant of
 * nbg_group_nansum for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * same arguments and results, but the scratch comes from a static per-device
 * buffer of NBG_GROUP_SCRATCH_BYTES instead of a workspace argument. Calls
 * whose nbg_group_workspace_bytes(...) exceeds it fail with NBG_ERR_INVALID
 * (32 bytes per (row, label): rows x num_labels <= 524288). Every call on one
 * device shares that device's buffer, so calls that overlap on one device
 * (different streams or host threads without ordering) race. */
#define NBG_GROUP_SCRATCH_BYTES (16 * 1024 * 1024)
int nbg_group_nansum_noalloc(int vdtype, const void* values, int ldtype,
                             const void* labels, void* out, int64_t rows,
                             int64_t n, int64_t ld_v, int64_t num_labels,
                             int strategy, void* stream);

/* ---- ragged batched serving path (no reference counterpart: its reference
 *      semantics are per-series plain calls of move_mean, moving.py:12-55,
 *      and group_nanmean, grouped.py:7-27) ---------------------------------- */
/* nseries float32 series concatenated in `values` (series s occupies
 * [offsets[s], offsets[s+1])), with int32 labels laid out alike; series s has
 * label_offsets[s+1] - label_offsets[s] labels and its group means go to
 * group_out[label_offsets[s] : label_offsets[s+1]]. Work items are tiles of
 * nbg_ragged_tile_size() positions of each series: tile_offsets[s] is the first item of series s and
 * tile_series[t] the series of item t (total_tiles items). strategy[s]: 0 =
 * labels below lds_labels accumulate in LDS-privatised slots (the rest, if
 * any, in global float64/uint32 accumulators), 1 = global atomics only.
 * lds_labels (1..nbg_ragged_max_lds_labels(), 0 = 1024) sets the LDS slots of
 * the launch: more slots keep more labels out of global atomics but leave
 * fewer workgroups resident per CU. One persistent launch with an atomic
 * work queue; move_out is bitwise equal to per-series move_mean calls. All
 * arrays are device arrays. */
/* Work-item (tile) length; the same tiles a plain float32 move_mean uses. */
int64_t nbg_ragged_tile_size(void);
int64_t nbg_ragged_max_lds_labels(void);
size_t nbg_ragged_workspace_bytes(int64_t nseries, int64_t total_labels);
int nbg_ragged_move_mean_group_nanmean(
    const float* values, const int32_t* labels, const int64_t* offsets,
    int64_t nseries, const int64_t* label_offsets, int64_t total_labels,
    const int64_t* tile_offsets, const int32_t* tile_series,
    int64_t total_tiles, const uint8_t* strategy, int64_t lds_labels,
    int64_t window, int64_t min_count, float* move_out, float* group_out,
    void* ws, size_t ws_bytes, void* stream);

/* ---- strategy cost model (no reference counterpart): fp32 MLP kernels ---- */
/* Row-major x (batch, in_dim), w (out_dim, in_dim), b (out_dim), y (batch,
 * out_dim); act: 0 identity, 1 ReLU. Backward takes the forward output y and
 * dy, writes dz (batch, out_dim) scratch, dw, db and (if dx != NULL) dx.
 * Weight gradients are fixed-order reductions (deterministic). */
int nbg_dense_forward(const float* x, const float* w, const float* b, float* y,
                      int64_t batch, int64_t in_dim, int64_t out_dim, int act,
                      void* stream);
int nbg_dense_backward(const float* x, const float* w, const float* y,
                       const float* dy, float* dx, float* dw, float* db,
                       float* dz, int64_t batch, int64_t in_dim,
                       int64_t out_dim, int act, void* stream);
/* dy = 2 (y - t) / n; *loss += mean squared error (loss may be NULL). */
int nbg_mse_grad(const float* y, const float* t, float* dy, float* loss,
                 int64_t n, void* stream);
/* One Adam update of n parameters (step counts from 1). */
int nbg_adam_step(float* p, const float* g, float* m, float* v, int64_t n,
                  float lr, float beta1, float beta2, float eps, int64_t step,
                  void* stream);

#ifdef __cplusplus
}
#endif

#endif /* NUMBAGG_AMD_H */
