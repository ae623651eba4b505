/*
 * TEST INFRASTRUCTURE ONLY — never linked into or called by the numbagg_amd
 * product path. Imported solely by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py, as the checker.
 *
 * A plain-C restatement of numbagg's per-row kernels for the hot path, with the
 * accumulator types and operation order numba infers for them (fastmath off, no
 * FMA contraction: build with -ffp-contract=off), and OpenMP over rows standing
 * in for numba's `parallel` target (numbagg/decorators.py:132-162).
 *
 * numba typing facts this file relies on (numba 0.67, uv.lock:503-504):
 *   - a float literal `0.0` types an accumulator as float64, so rolling and
 *     exponential accumulators are float64 even for float32 input;
 *   - float32 (op) float32 stays float32, including `x ** 2` with an integer
 *     exponent (numba's power typing keeps float32 ** int in single precision),
 *     so squares and cross products of float32 inputs are rounded to float32
 *     before being added to a float64 accumulator;
 *   - float32 (op) int promotes to float64;
 *   - grouped kernels accumulate in the output array's dtype (the values'
 *     dtype), counts in the labels' dtype (int64 here; the host widens labels).
 *
 * Each function cites the reference kernel it restates.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PRAGMA_OMP _Pragma("omp parallel for schedule(static)")

/* ------------------------------------------------------------------------ */
/* moving windows: numbagg/moving.py                                         */
/* ------------------------------------------------------------------------ */

/* move_mean, numbagg/moving.py:12-55 */
#define DEF_MOVE_MEAN(T, SFX)                                                           \
  void orc_move_mean_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ldx,   \
                           int64_t ldo, int64_t window, int64_t min_count) {           \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ldx;                                                         \
      T* o = out + r * ldo;                                                             \
      double asum = 0.0;                                                                \
      int64_t count = 0;                                                                \
      int64_t mc = min_count > 1 ? min_count : 1;                                       \
      int64_t w0 = window < n ? window : n;                                             \
      for (int64_t i = 0; i < w0; ++i) {                                                \
        T ai = a[i];                                                                    \
        if (!isnan(ai)) {                                                               \
          asum += (double)ai;                                                           \
          count += 1;                                                                   \
        }                                                                               \
        o[i] = (T)(count >= mc ? asum / (double)count : NAN);                           \
      }                                                                                 \
      for (int64_t i = window; i < n; ++i) {                                            \
        T ai = a[i];                                                                    \
        T aold = a[i - window];                                                         \
        if (!isnan(aold)) {                                                             \
          asum -= (double)aold;                                                         \
          count -= 1;                                                                   \
        }                                                                               \
        if (!isnan(ai)) {                                                               \
          asum += (double)ai;                                                           \
          count += 1;                                                                   \
        }                                                                               \
        o[i] = (T)(count >= mc ? asum / (double)count : NAN);                           \
      }                                                                                 \
    }                                                                                   \
  }
DEF_MOVE_MEAN(float, f32)
DEF_MOVE_MEAN(double, f64)

/* move_sum, numbagg/moving.py:58-99 (add the new value first, then remove) */
#define DEF_MOVE_SUM(T, SFX)                                                            \
  void orc_move_sum_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ldx,    \
                          int64_t ldo, int64_t window, int64_t min_count) {            \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ldx;                                                         \
      T* o = out + r * ldo;                                                             \
      double asum = 0.0;                                                                \
      int64_t count = 0;                                                                \
      int64_t w0 = window < n ? window : n;                                             \
      for (int64_t i = 0; i < w0; ++i) {                                                \
        T ai = a[i];                                                                    \
        if (!isnan(ai)) {                                                               \
          asum += (double)ai;                                                           \
          count += 1;                                                                   \
        }                                                                               \
        o[i] = (T)(count >= min_count ? asum : NAN);                                    \
      }                                                                                 \
      for (int64_t i = window; i < n; ++i) {                                            \
        T ai = a[i];                                                                    \
        T aold = a[i - window];                                                         \
        if (!isnan(ai)) {                                                               \
          asum += (double)ai;                                                           \
          count += 1;                                                                   \
        }                                                                               \
        if (!isnan(aold)) {                                                             \
          asum -= (double)aold;                                                         \
          count -= 1;                                                                   \
        }                                                                               \
        o[i] = (T)(count >= min_count ? asum : NAN);                                    \
      }                                                                                 \
    }                                                                                   \
  }
DEF_MOVE_SUM(float, f32)
DEF_MOVE_SUM(double, f64)

/* move_std (SQRT=1, numbagg/moving.py:123-151) and move_var (SQRT=0, :154-181) */
#define DEF_MOVE_VAR(T, SFX, NAME, SQRT)                                                \
  void orc_##NAME##_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ldx,    \
                          int64_t ldo, int64_t window, int64_t min_count) {            \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ldx;                                                         \
      T* o = out + r * ldo;                                                             \
      double asum = 0.0, asum_sq = 0.0;                                                 \
      int64_t count = 0;                                                                \
      int64_t mc = min_count > 2 ? min_count : 2;                                       \
      for (int64_t i = 0; i < n; ++i) {                                                 \
        T ai = a[i];                                                                    \
        if (i >= window) {                                                              \
          T aold = a[i - window];                                                       \
          if (!isnan(aold)) {                                                           \
            asum -= (double)aold;                                                       \
            asum_sq -= (double)(T)(aold * aold);                                        \
            count -= 1;                                                                 \
          }                                                                             \
        }                                                                               \
        if (!isnan(ai)) {                                                               \
          asum += (double)ai;                                                           \
          asum_sq += (double)(T)(ai * ai);                                              \
          count += 1;                                                                   \
        }                                                                               \
        if (count >= mc) {                                                              \
          double variance = (asum_sq - asum * asum / (double)count) / (double)(count - 1); \
          o[i] = (T)(SQRT ? sqrt(variance) : variance);                                 \
        } else {                                                                        \
          o[i] = (T)NAN;                                                                \
        }                                                                               \
      }                                                                                 \
    }                                                                                   \
  }
DEF_MOVE_VAR(float, f32, move_std, 1)
DEF_MOVE_VAR(double, f64, move_std, 1)
DEF_MOVE_VAR(float, f32, move_var, 0)
DEF_MOVE_VAR(double, f64, move_var, 0)

/* move_cov, numbagg/moving.py:184-220 (pairwise-complete) */
#define DEF_MOVE_COV(T, SFX)                                                            \
  void orc_move_cov_##SFX(const T* x, const T* y, T* out, int64_t rows, int64_t n,     \
                          int64_t ldx, int64_t ldy, int64_t ldo, int64_t window,       \
                          int64_t min_count) {                                          \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ldx;                                                         \
      const T* b = y + r * ldy;                                                         \
      T* o = out + r * ldo;                                                             \
      double asum = 0.0, bsum = 0.0, prodsum = 0.0;                                     \
      int64_t count = 0;                                                                \
      int64_t mc = min_count > 2 ? min_count : 2;                                       \
      for (int64_t i = 0; i < n; ++i) {                                                 \
        T ai = a[i], bi = b[i];                                                         \
        if (i >= window) {                                                              \
          T aold = a[i - window], bold = b[i - window];                                 \
          if (!(isnan(aold) || isnan(bold))) {                                          \
            asum -= (double)aold;                                                       \
            bsum -= (double)bold;                                                       \
            prodsum -= (double)(T)(aold * bold);                                        \
            count -= 1;                                                                 \
          }                                                                             \
        }                                                                               \
        if (!(isnan(ai) || isnan(bi))) {                                                \
          asum += (double)ai;                                                           \
          bsum += (double)bi;                                                           \
          prodsum += (double)(T)(ai * bi);                                              \
          count += 1;                                                                   \
        }                                                                               \
        if (count >= mc)                                                                \
          o[i] = (T)((prodsum - asum * bsum / (double)count) / (double)(count - 1));    \
        else                                                                            \
          o[i] = (T)NAN;                                                                \
      }                                                                                 \
    }                                                                                   \
  }
DEF_MOVE_COV(float, f32)
DEF_MOVE_COV(double, f64)

/* move_corr, numbagg/moving.py:223-275 */
#define DEF_MOVE_CORR(T, SFX)                                                           \
  void orc_move_corr_##SFX(const T* x, const T* y, T* out, int64_t rows, int64_t n,    \
                           int64_t ldx, int64_t ldy, int64_t ldo, int64_t window,      \
                           int64_t min_count) {                                         \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ldx;                                                         \
      const T* b = y + r * ldy;                                                         \
      T* o = out + r * ldo;                                                             \
      double asum = 0.0, bsum = 0.0, prodsum = 0.0, asum_sq = 0.0, bsum_sq = 0.0;       \
      int64_t count = 0;                                                                \
      int64_t mc = min_count > 1 ? min_count : 1;                                       \
      for (int64_t i = 0; i < n; ++i) {                                                 \
        T ai = a[i], bi = b[i];                                                         \
        if (i >= window) {                                                              \
          T aold = a[i - window], bold = b[i - window];                                 \
          if (!(isnan(aold) || isnan(bold))) {                                          \
            asum -= (double)aold;                                                       \
            bsum -= (double)bold;                                                       \
            prodsum -= (double)(T)(aold * bold);                                        \
            asum_sq -= (double)(T)(aold * aold);                                        \
            bsum_sq -= (double)(T)(bold * bold);                                        \
            count -= 1;                                                                 \
          }                                                                             \
        }                                                                               \
        if (!(isnan(ai) || isnan(bi))) {                                                \
          asum += (double)ai;                                                           \
          bsum += (double)bi;                                                           \
          prodsum += (double)(T)(ai * bi);                                              \
          asum_sq += (double)(T)(ai * ai);                                              \
          bsum_sq += (double)(T)(bi * bi);                                              \
          count += 1;                                                                   \
        }                                                                               \
        if (count >= mc) {                                                              \
          double cr = 1.0 / (double)count;                                              \
          double avg_a = asum * cr, avg_b = bsum * cr;                                  \
          double var_a = asum_sq * cr - avg_a * avg_a;                                  \
          double var_b = bsum_sq * cr - avg_b * avg_b;                                  \
          double cov_ab = prodsum * cr - avg_a * avg_b;                                 \
          double vv = var_a * var_b;                                                    \
          o[i] = (T)(vv > 0 ? cov_ab / sqrt(vv) : NAN);                                 \
        } else {                                                                        \
          o[i] = (T)NAN;                                                                \
        }                                                                               \
      }                                                                                 \
    }                                                                                   \
  }
DEF_MOVE_CORR(float, f32)
DEF_MOVE_CORR(double, f64)

/* ------------------------------------------------------------------------ */
/* exponential moving windows: numbagg/moving_exp.py                         */
/* alpha is a per-position array (ld_alpha = 0 shares one row across rows);  */
/* alpha and min_weight arrive in the input dtype, as in the reference's     */
/* signatures.                                                               */
/* ------------------------------------------------------------------------ */

#define EXP_LOOP_HEAD(T)                                                                \
  ORC_PRAGMA_OMP                                                                        \
  for (int64_t r = 0; r < rows; ++r) {                                                  \
    const T* a = x + r * ldx;                                                           \
    const T* al = alpha + r * lda;                                                      \
    T* o = out + r * ldo;

/* move_exp_nancount, numbagg/moving_exp.py:12-38 */
#define DEF_EXP_NANCOUNT(T, SFX)                                                        \
  void orc_move_exp_nancount_##SFX(const T* x, const T* alpha, T* out, int64_t rows,   \
                                   int64_t n, int64_t ldx, int64_t lda, int64_t ldo,   \
                                   T min_weight) {                                      \
    EXP_LOOP_HEAD(T)                                                                    \
    double count = 0.0, weight = 0.0;                                                   \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      T a_i = a[i];                                                                     \
      T alpha_i = al[i];                                                                \
      double decay = 1.0 - (double)alpha_i;                                             \
      count *= decay;                                                                   \
      weight *= decay;                                                                  \
      if (!isnan(a_i)) {                                                                \
        count += 1;                                                                     \
        weight += (double)alpha_i;                                                      \
      }                                                                                 \
      o[i] = (T)(weight >= (double)min_weight ? count : NAN);                           \
    }                                                                                   \
  }                                                                                     \
  }
DEF_EXP_NANCOUNT(float, f32)
DEF_EXP_NANCOUNT(double, f64)

/* move_exp_nanmean, numbagg/moving_exp.py:41-72 */
#define DEF_EXP_NANMEAN(T, SFX)                                                         \
  void orc_move_exp_nanmean_##SFX(const T* x, const T* alpha, T* out, int64_t rows,    \
                                  int64_t n, int64_t ldx, int64_t lda, int64_t ldo,    \
                                  T min_weight) {                                       \
    EXP_LOOP_HEAD(T)                                                                    \
    double numer = 0.0, denom = 0.0, weight = 0.0;                                      \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      T a_i = a[i];                                                                     \
      T alpha_i = al[i];                                                                \
      double decay = 1.0 - (double)alpha_i;                                             \
      numer *= decay;                                                                   \
      denom *= decay;                                                                   \
      weight *= decay;                                                                  \
      if (!isnan(a_i)) {                                                                \
        numer += (double)a_i;                                                           \
        denom += 1.0;                                                                   \
        weight += (double)alpha_i;                                                      \
      }                                                                                 \
      o[i] = (T)(weight >= (double)min_weight ? numer / denom : NAN);                   \
    }                                                                                   \
  }                                                                                     \
  }
DEF_EXP_NANMEAN(float, f32)
DEF_EXP_NANMEAN(double, f64)

/* move_exp_nansum, numbagg/moving_exp.py:75-103 */
#define DEF_EXP_NANSUM(T, SFX)                                                          \
  void orc_move_exp_nansum_##SFX(const T* x, const T* alpha, T* out, int64_t rows,     \
                                 int64_t n, int64_t ldx, int64_t lda, int64_t ldo,     \
                                 T min_weight) {                                        \
    EXP_LOOP_HEAD(T)                                                                    \
    double numer = 0.0, weight = 0.0;                                                   \
    int zero_count = 1;                                                                 \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      T a_i = a[i];                                                                     \
      T alpha_i = al[i];                                                                \
      double decay = 1.0 - (double)alpha_i;                                             \
      numer *= decay;                                                                   \
      weight *= decay;                                                                  \
      if (!isnan(a_i)) {                                                                \
        zero_count = 0;                                                                 \
        numer += (double)a_i;                                                           \
        weight += (double)alpha_i;                                                      \
      }                                                                                 \
      o[i] = (T)((weight >= (double)min_weight && !zero_count) ? numer : NAN);          \
    }                                                                                   \
  }                                                                                     \
  }
DEF_EXP_NANSUM(float, f32)
DEF_EXP_NANSUM(double, f64)

/* move_exp_nanvar (SQRT=0, numbagg/moving_exp.py:106-156) and move_exp_nanstd
 * (SQRT=1, :159-224) */
#define DEF_EXP_NANVAR(T, SFX, NAME, SQRT)                                              \
  void orc_##NAME##_##SFX(const T* x, const T* alpha, T* out, int64_t rows, int64_t n, \
                          int64_t ldx, int64_t lda, int64_t ldo, T min_weight) {        \
    EXP_LOOP_HEAD(T)                                                                    \
    double sum_x_2 = 0.0, sum_x = 0.0, sum_weight = 0.0, sum_weight_2 = 0.0,            \
           weight = 0.0;                                                                \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      T a_i = a[i];                                                                     \
      T alpha_i = al[i];                                                                \
      double decay = 1.0 - (double)alpha_i;                                             \
      sum_x_2 *= decay;                                                                 \
      sum_x *= decay;                                                                   \
      sum_weight *= decay;                                                              \
      sum_weight_2 *= decay * decay;                                                    \
      weight *= decay;                                                                  \
      if (!isnan(a_i)) {                                                                \
        sum_x_2 += (double)(T)(a_i * a_i);                                              \
        sum_x += (double)a_i;                                                           \
        sum_weight += 1.0;                                                              \
        sum_weight_2 += 1.0;                                                            \
        weight += (double)alpha_i;                                                      \
      }                                                                                 \
      double mean = sum_x / sum_weight;                                                 \
      double var_biased = (sum_x_2 / sum_weight) - mean * mean;                         \
      double bias = 1 - sum_weight_2 / (sum_weight * sum_weight);                       \
      if (weight >= (double)min_weight && bias > 0) {                                   \
        double v = var_biased / bias;                                                   \
        o[i] = (T)(SQRT ? sqrt(v) : v);                                                 \
      } else {                                                                          \
        o[i] = (T)NAN;                                                                  \
      }                                                                                 \
    }                                                                                   \
  }                                                                                     \
  }
DEF_EXP_NANVAR(float, f32, move_exp_nanvar, 0)
DEF_EXP_NANVAR(double, f64, move_exp_nanvar, 0)
DEF_EXP_NANVAR(float, f32, move_exp_nanstd, 1)
DEF_EXP_NANVAR(double, f64, move_exp_nanstd, 1)

/* move_exp_nancov, numbagg/moving_exp.py:227-273 */
#define DEF_EXP_NANCOV(T, SFX)                                                          \
  void orc_move_exp_nancov_##SFX(const T* x, const T* y, const T* alpha, T* out,       \
                                 int64_t rows, int64_t n, int64_t ldx, int64_t ldy,    \
                                 int64_t lda, int64_t ldo, T min_weight) {             \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a1 = x + r * ldx;                                                        \
      const T* a2 = y + r * ldy;                                                        \
      const T* al = alpha + r * lda;                                                    \
      T* o = out + r * ldo;                                                             \
      double sum_x1 = 0.0, sum_x2 = 0.0, sum_x1x2 = 0.0, sum_weight = 0.0,              \
             sum_weight_2 = 0.0, weight = 0.0;                                          \
      for (int64_t i = 0; i < n; ++i) {                                                 \
        T a1_i = a1[i], a2_i = a2[i];                                                   \
        T alpha_i = al[i];                                                              \
        double decay = 1.0 - (double)alpha_i;                                           \
        sum_x1 *= decay;                                                                \
        sum_x2 *= decay;                                                                \
        sum_x1x2 *= decay;                                                              \
        sum_weight *= decay;                                                            \
        sum_weight_2 *= decay * decay;                                                  \
        weight *= decay;                                                                \
        if (!(isnan(a1_i) || isnan(a2_i))) {                                            \
          sum_x1 += (double)a1_i;                                                       \
          sum_x2 += (double)a2_i;                                                       \
          sum_x1x2 += (double)(T)(a1_i * a2_i);                                         \
          sum_weight += 1;                                                              \
          sum_weight_2 += 1;                                                            \
          weight += (double)alpha_i;                                                    \
        }                                                                               \
        double cov_biased = ((sum_x1x2) - (sum_x1 * sum_x2 / sum_weight)) / sum_weight; \
        double bias = 1 - sum_weight_2 / (sum_weight * sum_weight);                     \
        o[i] = (T)((weight >= (double)min_weight && bias > 0) ? cov_biased / bias : NAN); \
      }                                                                                 \
    }                                                                                   \
  }
DEF_EXP_NANCOV(float, f32)
DEF_EXP_NANCOV(double, f64)

/* move_exp_nancorr, numbagg/moving_exp.py:276-335 */
#define DEF_EXP_NANCORR(T, SFX)                                                         \
  void orc_move_exp_nancorr_##SFX(const T* x, const T* y, const T* alpha, T* out,      \
                                  int64_t rows, int64_t n, int64_t ldx, int64_t ldy,   \
                                  int64_t lda, int64_t ldo, T min_weight) {            \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a1 = x + r * ldx;                                                        \
      const T* a2 = y + r * ldy;                                                        \
      const T* al = alpha + r * lda;                                                    \
      T* o = out + r * ldo;                                                             \
      double sum_x1 = 0, sum_x2 = 0, sum_x1x2 = 0, sum_weight = 0, sum_weight_2 = 0;    \
      double weight = 0, sum_x1_2 = 0, sum_x2_2 = 0;                                    \
      for (int64_t i = 0; i < n; ++i) {                                                 \
        T a1_i = a1[i], a2_i = a2[i];                                                   \
        T alpha_i = al[i];                                                              \
        double decay = 1.0 - (double)alpha_i;                                           \
        sum_x1 *= decay;                                                                \
        sum_x2 *= decay;                                                                \
        sum_x1x2 *= decay;                                                              \
        sum_weight *= decay;                                                            \
        sum_weight_2 *= decay * decay;                                                  \
        weight *= decay;                                                                \
        sum_x1_2 *= decay;                                                              \
        sum_x2_2 *= decay;                                                              \
        if (!(isnan(a1_i) || isnan(a2_i))) {                                            \
          sum_x1 += (double)a1_i;                                                       \
          sum_x2 += (double)a2_i;                                                       \
          sum_x1x2 += (double)(T)(a1_i * a2_i);                                         \
          sum_weight += 1;                                                              \
          sum_weight_2 += 1;                                                            \
          weight += (double)alpha_i;                                                    \
          sum_x1_2 += (double)(T)(a1_i * a1_i);                                         \
          sum_x2_2 += (double)(T)(a2_i * a2_i);                                         \
        }                                                                               \
        double cov = sum_x1x2 - (sum_x1 * sum_x2 / sum_weight);                         \
        double var_a1 = sum_x1_2 - (sum_x1 * sum_x1 / sum_weight);                      \
        double var_a2 = sum_x2_2 - (sum_x2 * sum_x2 / sum_weight);                      \
        double bias = 1 - sum_weight_2 / (sum_weight * sum_weight);                     \
        if (weight >= (double)min_weight && bias > 0) {                                 \
          double denominator = sqrt(var_a1 * var_a2);                                   \
          o[i] = (T)(denominator > 0 ? cov / denominator : NAN);                        \
        } else {                                                                        \
          o[i] = (T)NAN;                                                                \
        }                                                                               \
      }                                                                                 \
    }                                                                                   \
  }
DEF_EXP_NANCORR(float, f32)
DEF_EXP_NANCORR(double, f64)

/* ------------------------------------------------------------------------ */
/* fill: numbagg/funcs.py:294-326                                            */
/* ------------------------------------------------------------------------ */

/* ffill, numbagg/funcs.py:313-326 */
#define DEF_FFILL(T, SFX)                                                               \
  void orc_ffill_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ldx,       \
                       int64_t ldo, int64_t limit) {                                   \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ldx;                                                         \
      T* o = out + r * ldo;                                                             \
      int64_t lives_remaining = limit;                                                  \
      T current = (T)NAN;                                                               \
      for (int64_t i = 0; i < n; ++i) {                                                 \
        T val = a[i];                                                                   \
        if (isnan(val)) {                                                               \
          if (lives_remaining <= 0) current = (T)NAN;                                   \
          lives_remaining -= 1;                                                         \
        } else {                                                                        \
          lives_remaining = limit;                                                      \
          current = val;                                                                \
        }                                                                               \
        o[i] = current;                                                                 \
      }                                                                                 \
    }                                                                                   \
  }
DEF_FFILL(float, f32)
DEF_FFILL(double, f64)

/* bfill, numbagg/funcs.py:294-310 */
#define DEF_BFILL(T, SFX)                                                               \
  void orc_bfill_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ldx,       \
                       int64_t ldo, int64_t limit) {                                   \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ldx;                                                         \
      T* o = out + r * ldo;                                                             \
      int64_t lives_remaining = limit;                                                  \
      T current = (T)NAN;                                                               \
      for (int64_t i = n - 1; i >= 0; --i) {                                            \
        T val = a[i];                                                                   \
        if (isnan(val)) {                                                               \
          if (lives_remaining <= 0) current = (T)NAN;                                   \
          lives_remaining -= 1;                                                         \
        } else {                                                                        \
          lives_remaining = limit;                                                      \
          current = val;                                                                \
        }                                                                               \
        o[i] = current;                                                                 \
      }                                                                                 \
    }                                                                                   \
  }
DEF_BFILL(float, f32)
DEF_BFILL(double, f64)

/* ------------------------------------------------------------------------ */
/* grouped reductions: numbagg/grouped.py                                    */
/* values (rows, n) share int64 labels (n,); out is (rows, num_labels).      */
/* Labels < 0 are skipped; labels >= num_labels are skipped too (the         */
/* reference leaves them undefined).                                         */
/* ------------------------------------------------------------------------ */

#define ISNAN_F(v) isnan(v)
#define ISNAN_I(v) 0
#define NAN_F(T) ((T)NAN)
/* numba's float->int conversion of NaN (fptosi) yields the minimum integer. */
#define NAN_I32 ((int32_t)INT32_MIN)
#define NAN_I64 ((int64_t)INT64_MIN)

#define GROUP_LOOP_HEAD(T, OT)                                                          \
  ORC_PRAGMA_OMP                                                                        \
  for (int64_t r = 0; r < rows; ++r) {                                                  \
    const T* v = values + r * ldv;                                                      \
    OT* o = out + r * num_labels;                                                       \
    (void)v;

/* group_nansum, numbagg/grouped.py:30-40 (accumulates in the values dtype) */
#define DEF_GROUP_NANSUM(T, SFX, ISNAN)                                                 \
  void orc_group_nansum_##SFX(const T* values, const int64_t* labels, T* out,          \
                              int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) { \
    GROUP_LOOP_HEAD(T, T)                                                               \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = 0;                                  \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      T value = v[i];                                                                   \
      if (!ISNAN(value)) o[label] = (T)(o[label] + value);                              \
    }                                                                                   \
  }                                                                                     \
  }
DEF_GROUP_NANSUM(float, f32, ISNAN_F)
DEF_GROUP_NANSUM(double, f64, ISNAN_F)
/* int32 sums wrap like numba's int32 adds (computed unsigned to stay defined C) */
void orc_group_nansum_i32(const int32_t* values, const int64_t* labels, int32_t* out,
                          int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) {
  ORC_PRAGMA_OMP
  for (int64_t r = 0; r < rows; ++r) {
    const int32_t* v = values + r * ldv;
    uint32_t* o = (uint32_t*)(out + r * num_labels);
    for (int64_t k = 0; k < num_labels; ++k) o[k] = 0;
    for (int64_t i = 0; i < n; ++i) {
      int64_t label = labels[i];
      if (label < 0 || label >= num_labels) continue;
      o[label] += (uint32_t)v[i];
    }
  }
}
void orc_group_nansum_i64(const int64_t* values, const int64_t* labels, int64_t* out,
                          int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) {
  ORC_PRAGMA_OMP
  for (int64_t r = 0; r < rows; ++r) {
    const int64_t* v = values + r * ldv;
    uint64_t* o = (uint64_t*)(out + r * num_labels);
    for (int64_t k = 0; k < num_labels; ++k) o[k] = 0;
    for (int64_t i = 0; i < n; ++i) {
      int64_t label = labels[i];
      if (label < 0 || label >= num_labels) continue;
      o[label] += (uint64_t)v[i];
    }
  }
}

/* group_nanmean, numbagg/grouped.py:7-27 (float values only: the reference runs
 * integer input through its float64 loop) */
#define DEF_GROUP_NANMEAN(T, SFX)                                                       \
  void orc_group_nanmean_##SFX(const T* values, const int64_t* labels, T* out,         \
                               int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) { \
    GROUP_LOOP_HEAD(T, T)                                                               \
    int64_t* counts = (int64_t*)calloc((size_t)num_labels + 1, sizeof(int64_t));       \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = 0;                                  \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      T value = v[i];                                                                   \
      if (!isnan(value)) {                                                              \
        counts[label] += 1;                                                             \
        o[label] = (T)(o[label] + value);                                               \
      }                                                                                 \
    }                                                                                   \
    for (int64_t k = 0; k < num_labels; ++k)                                            \
      o[k] = counts[k] == 0 ? (T)NAN : (T)((double)o[k] / (double)counts[k]);           \
    free(counts);                                                                       \
  }                                                                                     \
  }

/* group_nanvar (SQRT=0, numbagg/grouped.py:148-176) and group_nanstd (SQRT=1,
 * :179-208): sums and sums of squares in the values dtype. */
#define DEF_GROUP_NANVAR(T, SFX, NAME, SQRT)                                            \
  void orc_##NAME##_##SFX(const T* values, const int64_t* labels, T* out, int64_t rows, \
                          int64_t n, int64_t ldv, int64_t num_labels, int64_t ddof) {  \
    GROUP_LOOP_HEAD(T, T)                                                               \
    T* sums = (T*)calloc((size_t)num_labels + 1, sizeof(T));                            \
    T* sos = (T*)calloc((size_t)num_labels + 1, sizeof(T));                             \
    int64_t* counts = (int64_t*)calloc((size_t)num_labels + 1, sizeof(int64_t));       \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      T value = v[i];                                                                   \
      if (!isnan(value)) {                                                              \
        counts[label] += 1;                                                             \
        sums[label] = (T)(sums[label] + value);                                         \
        sos[label] = (T)(sos[label] + (T)(value * value));                              \
      }                                                                                 \
    }                                                                                   \
    for (int64_t k = 0; k < num_labels; ++k) {                                          \
      int64_t count = counts[k];                                                        \
      int64_t denom = count - ddof;                                                     \
      if (denom <= 0) {                                                                 \
        o[k] = (T)NAN;                                                                  \
      } else {                                                                          \
        /* `sums ** 2` stays in T; T / int64 and T - float64 promote to float64 */     \
        double var = ((double)sos[k] - (double)(T)(sums[k] * sums[k]) / (double)count) \
                     / (double)denom;                                                   \
        o[k] = (T)(SQRT ? sqrt(var) : var);                                             \
      }                                                                                 \
    }                                                                                   \
    free(sums);                                                                         \
    free(sos);                                                                          \
    free(counts);                                                                       \
  }                                                                                     \
  }

/* group_nancount, numbagg/grouped.py:43-51 */
#define DEF_GROUP_NANCOUNT(T, OT, SFX, ISNAN)                                           \
  void orc_group_nancount_##SFX(const T* values, const int64_t* labels, OT* out,       \
                                int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) { \
    GROUP_LOOP_HEAD(T, OT)                                                              \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = 0;                                  \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      if (!ISNAN(v[i])) o[label] = (OT)(o[label] + 1);                                  \
    }                                                                                   \
  }                                                                                     \
  }

/* group_nanargmax (CMP >, numbagg/grouped.py:54-75) and group_nanargmin
 * (CMP <, :78-92); running extremes are float64, the first index wins ties. */
#define DEF_GROUP_NANARG(T, SFX, NAME, CMP, ISNAN, NANV)                                \
  void orc_##NAME##_##SFX(const T* values, const int64_t* labels, T* out, int64_t rows, \
                          int64_t n, int64_t ldv, int64_t num_labels) {                 \
    GROUP_LOOP_HEAD(T, T)                                                               \
    double* best = (double*)malloc(((size_t)num_labels + 1) * sizeof(double));          \
    for (int64_t k = 0; k < num_labels; ++k) best[k] = NAN;                             \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      T value = v[i];                                                                   \
      int64_t label = labels[i];                                                        \
      if (label >= 0 && label < num_labels && !ISNAN(value)) {                          \
        if (isnan(best[label]) || (double)value CMP best[label]) {                      \
          best[label] = (double)value;                                                  \
          o[label] = (T)i;                                                              \
        }                                                                               \
      }                                                                                 \
    }                                                                                   \
    for (int64_t k = 0; k < num_labels; ++k)                                            \
      if (isnan(best[k])) o[k] = NANV;                                                  \
    free(best);                                                                         \
  }                                                                                     \
  }

/* group_nanfirst, numbagg/grouped.py:95-110 (unseen labels: NaN for floats; the
 * reference leaves integer outputs uninitialised, here they get NANV) */
#define DEF_GROUP_NANFIRST(T, SFX, ISNAN, NANV)                                         \
  void orc_group_nanfirst_##SFX(const T* values, const int64_t* labels, T* out,        \
                                int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) { \
    GROUP_LOOP_HEAD(T, T)                                                               \
    unsigned char* seen = (unsigned char*)calloc((size_t)num_labels + 1, 1);            \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      if (!seen[label] && !ISNAN(v[i])) {                                               \
        seen[label] = 1;                                                                \
        o[label] = v[i];                                                                \
      }                                                                                 \
    }                                                                                   \
    for (int64_t k = 0; k < num_labels; ++k)                                            \
      if (!seen[k]) o[k] = NANV;                                                        \
    free(seen);                                                                         \
  }                                                                                     \
  }

/* group_nanlast, numbagg/grouped.py:113-121 */
#define DEF_GROUP_NANLAST(T, SFX, ISNAN, NANV)                                          \
  void orc_group_nanlast_##SFX(const T* values, const int64_t* labels, T* out,         \
                               int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) { \
    GROUP_LOOP_HEAD(T, T)                                                               \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = NANV;                               \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      if (!ISNAN(v[i])) o[label] = v[i];                                                \
    }                                                                                   \
  }                                                                                     \
  }

DEF_GROUP_NANMEAN(float, f32)
DEF_GROUP_NANMEAN(double, f64)
DEF_GROUP_NANVAR(float, f32, group_nanvar, 0)
DEF_GROUP_NANVAR(double, f64, group_nanvar, 0)
DEF_GROUP_NANVAR(float, f32, group_nanstd, 1)
DEF_GROUP_NANVAR(double, f64, group_nanstd, 1)
DEF_GROUP_NANCOUNT(float, float, f32, ISNAN_F)
DEF_GROUP_NANCOUNT(double, double, f64, ISNAN_F)
DEF_GROUP_NANCOUNT(int32_t, int32_t, i32, ISNAN_I)
DEF_GROUP_NANCOUNT(int64_t, int64_t, i64, ISNAN_I)
DEF_GROUP_NANARG(float, f32, group_nanargmax, >, ISNAN_F, NAN_F(float))
DEF_GROUP_NANARG(double, f64, group_nanargmax, >, ISNAN_F, NAN_F(double))
DEF_GROUP_NANARG(int32_t, i32, group_nanargmax, >, ISNAN_I, NAN_I32)
DEF_GROUP_NANARG(int64_t, i64, group_nanargmax, >, ISNAN_I, NAN_I64)
DEF_GROUP_NANARG(float, f32, group_nanargmin, <, ISNAN_F, NAN_F(float))
DEF_GROUP_NANARG(double, f64, group_nanargmin, <, ISNAN_F, NAN_F(double))
DEF_GROUP_NANARG(int32_t, i32, group_nanargmin, <, ISNAN_I, NAN_I32)
DEF_GROUP_NANARG(int64_t, i64, group_nanargmin, <, ISNAN_I, NAN_I64)
DEF_GROUP_NANFIRST(float, f32, ISNAN_F, NAN_F(float))
DEF_GROUP_NANFIRST(double, f64, ISNAN_F, NAN_F(double))
DEF_GROUP_NANFIRST(int32_t, i32, ISNAN_I, NAN_I32)
DEF_GROUP_NANFIRST(int64_t, i64, ISNAN_I, NAN_I64)
DEF_GROUP_NANLAST(float, f32, ISNAN_F, NAN_F(float))
DEF_GROUP_NANLAST(double, f64, ISNAN_F, NAN_F(double))
DEF_GROUP_NANLAST(int32_t, i32, ISNAN_I, NAN_I32)
DEF_GROUP_NANLAST(int64_t, i64, ISNAN_I, NAN_I64)

/* group_nanprod, numbagg/grouped.py:124-132: products in the values dtype
 * (integer products wrap like numba's fixed-width multiplies; computed
 * unsigned to stay defined C). */
#define DEF_GROUP_NANPROD(T, U, SFX, ISNAN)                                             \
  void orc_group_nanprod_##SFX(const T* values, const int64_t* labels, T* out,         \
                               int64_t rows, int64_t n, int64_t ldv, int64_t num_labels) { \
    GROUP_LOOP_HEAD(T, T)                                                               \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = 1;                                  \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      T value = v[i];                                                                   \
      if (!ISNAN(value)) o[label] = (T)((U)o[label] * (U)value);                        \
    }                                                                                   \
  }                                                                                     \
  }
DEF_GROUP_NANPROD(float, float, f32, ISNAN_F)
DEF_GROUP_NANPROD(double, double, f64, ISNAN_F)
DEF_GROUP_NANPROD(int32_t, uint32_t, i32, ISNAN_I)
DEF_GROUP_NANPROD(int64_t, uint64_t, i64, ISNAN_I)

/* group_nansum_of_squares, numbagg/grouped.py:135-145: `value ** 2` in the
 * values dtype (float32 squares round to float32), accumulated in it. */
#define DEF_GROUP_NANSUMSQ(T, U, SFX, ISNAN)                                            \
  void orc_group_nansum_of_squares_##SFX(const T* values, const int64_t* labels, T* out, \
                                         int64_t rows, int64_t n, int64_t ldv,         \
                                         int64_t num_labels) {                          \
    GROUP_LOOP_HEAD(T, T)                                                               \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = 0;                                  \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      T value = v[i];                                                                   \
      if (!ISNAN(value)) {                                                              \
        U sq = (U)((U)value * (U)value);                                                \
        o[label] = (T)((U)o[label] + sq);                                               \
      }                                                                                 \
    }                                                                                   \
  }                                                                                     \
  }
DEF_GROUP_NANSUMSQ(float, float, f32, ISNAN_F)
DEF_GROUP_NANSUMSQ(double, double, f64, ISNAN_F)
DEF_GROUP_NANSUMSQ(int32_t, uint32_t, i32, ISNAN_I)
DEF_GROUP_NANSUMSQ(int64_t, uint64_t, i64, ISNAN_I)

/* group_nanmin (CMP <, numbagg/grouped.py:211-225) and group_nanmax (CMP >,
 * :228-244): running extremes in a float64 array initialised to NaN, copied
 * to the output at the end (NaN -> minimum integer for integer outputs). */
#define DEF_GROUP_NANEXT(T, SFX, NAME, CMP, ISNAN, NANV)                                \
  void orc_##NAME##_##SFX(const T* values, const int64_t* labels, T* out, int64_t rows, \
                          int64_t n, int64_t ldv, int64_t num_labels) {                 \
    GROUP_LOOP_HEAD(T, T)                                                               \
    double* best = (double*)malloc(((size_t)num_labels + 1) * sizeof(double));          \
    for (int64_t k = 0; k < num_labels; ++k) best[k] = NAN;                             \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      T value = v[i];                                                                   \
      if (!ISNAN(value) && (isnan(best[label]) || (double)value CMP best[label]))       \
        best[label] = (double)value;                                                    \
    }                                                                                   \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = isnan(best[k]) ? NANV : (T)best[k]; \
    free(best);                                                                         \
  }                                                                                     \
  }
DEF_GROUP_NANEXT(float, f32, group_nanmin, <, ISNAN_F, NAN_F(float))
DEF_GROUP_NANEXT(double, f64, group_nanmin, <, ISNAN_F, NAN_F(double))
DEF_GROUP_NANEXT(int32_t, i32, group_nanmin, <, ISNAN_I, NAN_I32)
DEF_GROUP_NANEXT(int64_t, i64, group_nanmin, <, ISNAN_I, NAN_I64)
DEF_GROUP_NANEXT(float, f32, group_nanmax, >, ISNAN_F, NAN_F(float))
DEF_GROUP_NANEXT(double, f64, group_nanmax, >, ISNAN_F, NAN_F(double))
DEF_GROUP_NANEXT(int32_t, i32, group_nanmax, >, ISNAN_I, NAN_I32)
DEF_GROUP_NANEXT(int64_t, i64, group_nanmax, >, ISNAN_I, NAN_I64)

/* group_nanany (numbagg/grouped.py:247-257: 0, then 1 on a valid truthy
 * value) and group_nanall (:260-270: 1, then 0 on a valid falsy value). */
#define DEF_GROUP_NANANYALL(T, SFX, NAME, INIT, HIT, ISNAN)                             \
  void orc_##NAME##_##SFX(const T* values, const int64_t* labels, T* out, int64_t rows, \
                          int64_t n, int64_t ldv, int64_t num_labels) {                 \
    GROUP_LOOP_HEAD(T, T)                                                               \
    for (int64_t k = 0; k < num_labels; ++k) o[k] = INIT;                               \
    for (int64_t i = 0; i < n; ++i) {                                                   \
      int64_t label = labels[i];                                                        \
      if (label < 0 || label >= num_labels) continue;                                   \
      T value = v[i];                                                                   \
      if (!ISNAN(value) && (HIT)) o[label] = 1 - INIT;                                  \
    }                                                                                   \
  }                                                                                     \
  }
DEF_GROUP_NANANYALL(float, f32, group_nanany, 0, value != 0, ISNAN_F)
DEF_GROUP_NANANYALL(double, f64, group_nanany, 0, value != 0, ISNAN_F)
DEF_GROUP_NANANYALL(int32_t, i32, group_nanany, 0, value != 0, ISNAN_I)
DEF_GROUP_NANANYALL(int64_t, i64, group_nanany, 0, value != 0, ISNAN_I)
DEF_GROUP_NANANYALL(float, f32, group_nanall, 1, value == 0, ISNAN_F)
DEF_GROUP_NANANYALL(double, f64, group_nanall, 1, value == 0, ISNAN_F)
DEF_GROUP_NANANYALL(int32_t, i32, group_nanall, 1, value == 0, ISNAN_I)
DEF_GROUP_NANANYALL(int64_t, i64, group_nanall, 1, value == 0, ISNAN_I)

/* ------------------------------------------------------------------------ */
/* Exact oracles (long double, 64-bit significand) used to state the error   */
/* bounds of float results. They follow the mathematical definitions the     */
/* reference kernels compute, not the reference's operation order, and       */
/* write float64 so float32 results can be measured in float32 ULPs.         */
/* ------------------------------------------------------------------------ */

/* Windowed statistics recomputed directly for every output (two-pass inside
 * each window). kind: 0 sum, 1 mean, 2 var, 3 std, 4 cov, 5 corr. Inputs are
 * float64 (exact copies of the float32 data when the data is float32). The
 * valid-count thresholds are the reference's: sum min_count; mean
 * max(min_count,1); var/std/cov max(min_count,2); corr max(min_count,1). */
void orc_exact_move(int kind, const double* x, const double* y, double* out, int64_t rows,
                    int64_t n, int64_t window, int64_t min_count) {
  int64_t mc = min_count;
  if (kind == 1 || kind == 5) mc = mc > 1 ? mc : 1;
  if (kind == 2 || kind == 3 || kind == 4) mc = mc > 2 ? mc : 2;
  ORC_PRAGMA_OMP
  for (int64_t r = 0; r < rows; ++r) {
    const double* a = x + r * n;
    const double* b = y ? y + r * n : NULL;
    double* o = out + r * n;
    for (int64_t i = 0; i < n; ++i) {
      int64_t j0 = i - window + 1 > 0 ? i - window + 1 : 0;
      long double sa = 0, sb = 0;
      int64_t c = 0;
      for (int64_t j = j0; j <= i; ++j) {
        if (isnan(a[j]) || (b && isnan(b[j]))) continue;
        sa += a[j];
        if (b) sb += b[j];
        c += 1;
      }
      if (c < mc) {
        o[i] = NAN;
        continue;
      }
      if (kind == 0) { o[i] = (double)sa; continue; }
      if (kind == 1) { o[i] = (double)(sa / c); continue; }
      long double ma = sa / c, mb = b ? sb / c : 0;
      long double saa = 0, sbb = 0, sab = 0;
      for (int64_t j = j0; j <= i; ++j) {
        if (isnan(a[j]) || (b && isnan(b[j]))) continue;
        long double da = a[j] - ma;
        long double db = b ? b[j] - mb : 0;
        saa += da * da;
        sbb += db * db;
        sab += da * db;
      }
      if (kind == 2) o[i] = (double)(saa / (c - 1));
      else if (kind == 3) o[i] = (double)sqrtl(saa / (c - 1));
      else if (kind == 4) o[i] = (double)(sab / (c - 1));
      else {
        long double vv = saa * sbb;
        o[i] = vv > 0 ? (double)(sab / sqrtl(vv)) : NAN;
      }
    }
  }
}

/* Exponentially weighted statistics by the reference's recurrences carried
 * in long double. kind: 0 nancount, 1 nanmean, 2 nansum, 3 nanvar, 4 nanstd,
 * 5 nancov, 6 nancorr (numbagg/moving_exp.py:12-335). alpha holds the loop
 * dtype's values widened to float64; ld_alpha 0 shares one alpha row.
 * With a single valid observation the bias correction 1 - sw2/sw^2 is exactly
 * zero, so var/std/cov/corr are NaN until two observations have been seen (the
 * reference evaluates the expression in floating point, where it can land a
 * few ULPs either side of zero when alpha varies by position). */
void orc_exact_move_exp(int kind, const double* x, const double* y, const double* alpha,
                        int64_t ld_alpha, double* out, int64_t rows, int64_t n,
                        double min_weight) {
  ORC_PRAGMA_OMP
  for (int64_t r = 0; r < rows; ++r) {
    const double* a = x + r * n;
    const double* b = y ? y + r * n : NULL;
    const double* al = alpha + r * ld_alpha;
    double* o = out + r * n;
    long double s1 = 0, s2 = 0, s11 = 0, s22 = 0, s12 = 0, sw = 0, sw2 = 0, wt = 0;
    int seen = 0, nobs = 0;
    for (int64_t i = 0; i < n; ++i) {
      long double d = 1.0L - (long double)al[i];
      s1 *= d; s2 *= d; s11 *= d; s22 *= d; s12 *= d; sw *= d; sw2 *= d * d; wt *= d;
      int valid = !isnan(a[i]) && !(b && isnan(b[i]));
      if (valid) {
        long double xa = a[i], xb = b ? b[i] : 0;
        s1 += xa; s2 += xb; s11 += xa * xa; s22 += xb * xb; s12 += xa * xb;
        sw += 1; sw2 += 1; wt += al[i];
        seen = 1;
        nobs += nobs < 2;
      }
      int ok = wt >= (long double)min_weight;
      long double bias = 1 - sw2 / (sw * sw);
      if (kind >= 3 && nobs < 2) bias = 0;
      double res = NAN;
      switch (kind) {
        case 0: res = ok ? (double)sw : NAN; break;
        case 1: res = ok ? (double)(s1 / sw) : NAN; break;
        case 2: res = (ok && seen) ? (double)s1 : NAN; break;
        case 3:
        case 4: {
          long double m = s1 / sw;
          long double vb = s11 / sw - m * m;
          if (ok && bias > 0) res = kind == 3 ? (double)(vb / bias) : (double)sqrtl(vb / bias);
          break;
        }
        case 5: {
          long double cb = (s12 - s1 * s2 / sw) / sw;
          if (ok && bias > 0) res = (double)(cb / bias);
          break;
        }
        case 6: {
          long double cv = s12 - s1 * s2 / sw;
          long double v1 = s11 - s1 * s1 / sw, v2 = s22 - s2 * s2 / sw;
          if (ok && bias > 0) {
            long double den = sqrtl(v1 * v2);
            res = den > 0 ? (double)(cv / den) : NAN;
          }
          break;
        }
      }
      o[i] = res;
    }
  }
}

/* ------------------------------------------------------------------------ */
/* plain reductions over the last axis of (rows, n): numbagg/funcs.py       */
/* ------------------------------------------------------------------------ */

/* allnan (numbagg/funcs.py:23-35) and anynan (:38-50): bool outputs */
#define DEF_ALLANY(T, SFX, ISNAN)                                                       \
  void orc_allnan_##SFX(const T* x, uint8_t* out, int64_t rows, int64_t n, int64_t ld) { \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      out[r] = 1;                                                                       \
      for (int64_t i = 0; i < n; ++i)                                                   \
        if (!ISNAN(a[i])) {                                                             \
          out[r] = 0;                                                                   \
          break;                                                                        \
        }                                                                               \
    }                                                                                   \
  }                                                                                     \
  void orc_anynan_##SFX(const T* x, uint8_t* out, int64_t rows, int64_t n, int64_t ld) { \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      out[r] = 0;                                                                       \
      for (int64_t i = 0; i < n; ++i)                                                   \
        if (ISNAN(a[i])) {                                                              \
          out[r] = 1;                                                                   \
          break;                                                                        \
        }                                                                               \
    }                                                                                   \
  }
DEF_ALLANY(float, f32, ISNAN_F)
DEF_ALLANY(double, f64, ISNAN_F)
DEF_ALLANY(int32_t, i32, ISNAN_I)
DEF_ALLANY(int64_t, i64, ISNAN_I)

/* nancount (numbagg/funcs.py:53-66): int64 output */
#define DEF_NANCOUNT(T, SFX, ISNAN)                                                     \
  void orc_nancount_##SFX(const T* x, int64_t* out, int64_t rows, int64_t n, int64_t ld) { \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      int64_t c = 0;                                                                    \
      for (int64_t i = 0; i < n; ++i) c += !ISNAN(a[i]);                                \
      out[r] = c;                                                                       \
    }                                                                                   \
  }
DEF_NANCOUNT(float, f32, ISNAN_F)
DEF_NANCOUNT(double, f64, ISNAN_F)
DEF_NANCOUNT(int32_t, i32, ISNAN_I)
DEF_NANCOUNT(int64_t, i64, ISNAN_I)

/* nansum (numbagg/funcs.py:69-82): accumulates in the input dtype (integer sums
 * wrap like numba's fixed-width adds; computed unsigned to stay defined C) */
#define DEF_NANSUM(T, U, SFX, ISNAN)                                                    \
  void orc_nansum_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ld) {      \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      U s = 0;                                                                          \
      for (int64_t i = 0; i < n; ++i)                                                   \
        if (!ISNAN(a[i])) s = (U)(s + (U)a[i]);                                         \
      out[r] = (T)s;                                                                    \
    }                                                                                   \
  }
DEF_NANSUM(float, float, f32, ISNAN_F)
DEF_NANSUM(double, double, f64, ISNAN_F)
DEF_NANSUM(int32_t, uint32_t, i32, ISNAN_I)
DEF_NANSUM(int64_t, uint64_t, i64, ISNAN_I)

/* nanmean (numbagg/funcs.py:85-102): float64 sum (`asum = 0.0`), int count */
#define DEF_NANMEAN(T, SFX)                                                             \
  void orc_nanmean_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ld) {     \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      double s = 0.0;                                                                   \
      int64_t c = 0;                                                                    \
      for (int64_t i = 0; i < n; ++i)                                                   \
        if (!isnan(a[i])) {                                                             \
          s += (double)a[i];                                                            \
          c += 1;                                                                       \
        }                                                                               \
      out[r] = c > 0 ? (T)(s / (double)c) : (T)NAN;                                     \
    }                                                                                   \
  }
DEF_NANMEAN(float, f32)
DEF_NANMEAN(double, f64)

/* nanvar (SQRT=0, numbagg/funcs.py:105-130) and nanstd (SQRT=1, :133-157): two
 * passes in float64 (the loop variable unifies to float64 after `ai -= amean`) */
#define DEF_NANVAR(T, SFX, NAME, SQRT)                                                  \
  void orc_##NAME##_##SFX(const T* x, T* out, int64_t rows, int64_t n, int64_t ld,      \
                          int64_t ddof) {                                               \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      double s = 0.0;                                                                   \
      int64_t c = 0;                                                                    \
      for (int64_t i = 0; i < n; ++i)                                                   \
        if (!isnan(a[i])) {                                                             \
          s += (double)a[i];                                                            \
          c += 1;                                                                       \
        }                                                                               \
      if (c > ddof) {                                                                   \
        const double m = s / (double)c;                                                 \
        double q = 0.0;                                                                 \
        for (int64_t i = 0; i < n; ++i)                                                 \
          if (!isnan(a[i])) {                                                           \
            const double d = (double)a[i] - m;                                          \
            q += d * d;                                                                 \
          }                                                                             \
        const double v = q / (double)(c - ddof);                                        \
        out[r] = (T)(SQRT ? sqrt(v) : v);                                               \
      } else {                                                                          \
        out[r] = (T)NAN;                                                                \
      }                                                                                 \
    }                                                                                   \
  }
DEF_NANVAR(float, f32, nanvar, 0)
DEF_NANVAR(double, f64, nanvar, 0)
DEF_NANVAR(float, f32, nanstd, 1)
DEF_NANVAR(double, f64, nanstd, 1)

/* nanmax (CMP >=, INIT -inf, numbagg/funcs.py:200-216) and nanmin (CMP <=, INIT
 * +inf, :219-234): float64 running extreme replaced on >= / <=; all-NaN -> NaN;
 * integer inputs return int64 (the float64 extreme converted). n > 0. */
#define DEF_NANEXT(T, OT, SFX, NAME, CMP, INIT)                                         \
  void orc_##NAME##_##SFX(const T* x, OT* out, int64_t rows, int64_t n, int64_t ld) {   \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      double best = INIT;                                                               \
      int missing = 1;                                                                  \
      for (int64_t i = 0; i < n; ++i)                                                   \
        if ((double)a[i] CMP best) {                                                    \
          best = (double)a[i];                                                          \
          missing = 0;                                                                  \
        }                                                                               \
      out[r] = missing ? (OT)NAN : (OT)best;                                            \
    }                                                                                   \
  }
DEF_NANEXT(float, float, f32, nanmax, >=, -INFINITY)
DEF_NANEXT(double, double, f64, nanmax, >=, -INFINITY)
DEF_NANEXT(int32_t, int64_t, i32, nanmax, >=, -INFINITY)
DEF_NANEXT(int64_t, int64_t, i64, nanmax, >=, -INFINITY)
DEF_NANEXT(float, float, f32, nanmin, <=, INFINITY)
DEF_NANEXT(double, double, f64, nanmin, <=, INFINITY)
DEF_NANEXT(int32_t, int64_t, i32, nanmin, <=, INFINITY)
DEF_NANEXT(int64_t, int64_t, i64, nanmin, <=, INFINITY)

/* nanargmax (CMP >, INIT -inf, numbagg/funcs.py:160-178) and nanargmin (CMP <,
 * INIT +inf, :181-197): the first index of the extreme; -1 for an all-NaN row
 * (the reference raises ValueError there). */
#define DEF_NANARGEXT(T, SFX, NAME, CMP, INIT)                                          \
  void orc_##NAME##_##SFX(const T* x, int64_t* out, int64_t rows, int64_t n, int64_t ld) { \
    ORC_PRAGMA_OMP                                                                      \
    for (int64_t r = 0; r < rows; ++r) {                                                \
      const T* a = x + r * ld;                                                          \
      double best = INIT;                                                               \
      int64_t idx = -1;                                                                 \
      for (int64_t i = 0; i < n; ++i) {                                                 \
        const double v = (double)a[i];                                                  \
        if (v CMP best || (idx == -1 && !isnan(v))) {                                   \
          best = v;                                                                     \
          idx = i;                                                                      \
        }                                                                               \
      }                                                                                 \
      out[r] = idx;                                                                     \
    }                                                                                   \
  }
DEF_NANARGEXT(float, f32, nanargmax, >, -INFINITY)
DEF_NANARGEXT(double, f64, nanargmax, >, -INFINITY)
DEF_NANARGEXT(int32_t, i32, nanargmax, >, -INFINITY)
DEF_NANARGEXT(int64_t, i64, nanargmax, >, -INFINITY)
DEF_NANARGEXT(float, f32, nanargmin, <, INFINITY)
DEF_NANARGEXT(double, f64, nanargmin, <, INFINITY)
DEF_NANARGEXT(int32_t, i32, nanargmin, <, INFINITY)
DEF_NANARGEXT(int64_t, i64, nanargmin, <, INFINITY)

/* ------------------------------------------------------------------------ */
/* matrix functions: numbagg/funcs.py:338-532, numbagg/moving_matrix.py      */
/* Pairwise-complete statistics per (i, j): an observation counts for a pair */
/* only where both variables are valid. Accumulators are the data dtype, as  */
/* numba types the np.zeros(..., dtype=a.dtype) arrays; counts are int64.   */
/* ------------------------------------------------------------------------ */

/* nancorrmatrix (numbagg/funcs.py:338-441) and nancovmatrix (:444-532):
 * x (batch, nv, nobs) -> out (batch, nv, nv); upper triangle accumulated in
 * observation order, mirrored at the end. */
#define DEF_STATIC_MATRIX(T, SFX)                                                         \
  static void orc_static_matrix_##SFX(const T* x, T* out, int64_t batch, int64_t nv,      \
                                      int64_t nobs, int corr) {                           \
    ORC_PRAGMA_OMP                                                                        \
    for (int64_t b = 0; b < batch; ++b) {                                                 \
      const T* a = x + b * nv * nobs;                                                     \
      T* o = out + b * nv * nv;                                                           \
      const int64_t P = nv * nv;                                                          \
      T* si = (T*)calloc((size_t)P * 5, sizeof(T));                                       \
      T *sj = si + P, *sqi = si + 2 * P, *sqj = si + 3 * P, *sij = si + 4 * P;             \
      int64_t* cnt = (int64_t*)calloc((size_t)P, sizeof(int64_t));                        \
      for (int64_t k = 0; k < nobs; ++k) {                                                \
        for (int64_t i = 0; i < nv; ++i) {                                                \
          const T vi = a[i * nobs + k];                                                   \
          if (isnan(vi)) continue;                                                        \
          for (int64_t j = i; j < nv; ++j) {                                              \
            const T vj = a[j * nobs + k];                                                 \
            if (isnan(vj)) continue;                                                      \
            const int64_t p = i * nv + j;                                                 \
            si[p] += vi;                                                                  \
            sj[p] += vj;                                                                  \
            sqi[p] += (T)(vi * vi);                                                       \
            sqj[p] += (T)(vj * vj);                                                       \
            sij[p] += (T)(vi * vj);                                                       \
            cnt[p] += 1;                                                                  \
          }                                                                               \
        }                                                                                 \
      }                                                                                   \
      for (int64_t i = 0; i < nv; ++i) {                                                  \
        for (int64_t j = i; j < nv; ++j) {                                                \
          const int64_t p = i * nv + j;                                                   \
          const int64_t c = cnt[p];                                                       \
          double r = NAN;                                                                 \
          if (c > 1) {                                                                    \
            const double mi = (double)si[p] / (double)c, mj = (double)sj[p] / (double)c;  \
            const double cv = ((double)sij[p] / (double)c - mi * mj) * (double)c / (double)(c - 1); \
            if (!corr) {                                                                  \
              r = cv;                                                                     \
            } else {                                                                      \
              const double vi = ((double)sqi[p] / (double)c - mi * mi) * (double)c / (double)(c - 1); \
              const double vj = ((double)sqj[p] / (double)c - mj * mj) * (double)c / (double)(c - 1); \
              if (vi > 0 && vj > 0) r = cv / sqrt(vi * vj);                               \
            }                                                                             \
          }                                                                               \
          o[i * nv + j] = (T)r;                                                           \
          o[j * nv + i] = (T)r;                                                           \
        }                                                                                 \
      }                                                                                   \
      free(si);                                                                           \
      free(cnt);                                                                          \
    }                                                                                     \
  }                                                                                       \
  void orc_nancorrmatrix_##SFX(const T* x, T* out, int64_t batch, int64_t nv, int64_t nobs) { \
    orc_static_matrix_##SFX(x, out, batch, nv, nobs, 1);                                  \
  }                                                                                       \
  void orc_nancovmatrix_##SFX(const T* x, T* out, int64_t batch, int64_t nv, int64_t nobs) { \
    orc_static_matrix_##SFX(x, out, batch, nv, nobs, 0);                                  \
  }
DEF_STATIC_MATRIX(float, f32)
DEF_STATIC_MATRIX(double, f64)

/* move_corrmatrix (numbagg/moving_matrix.py:16-114) and move_covmatrix
 * (:117-201): x (batch, nobs, nv) -> out (batch, nobs, nv, nv). Each step
 * removes the observation leaving the window, then adds the new one, for every
 * ordered pair. */
#define DEF_MOVE_MATRIX(T, SFX)                                                           \
  static void orc_move_matrix_##SFX(const T* x, T* out, int64_t batch, int64_t nobs,      \
                                    int64_t nv, int64_t window, int64_t min_count, int corr) { \
    ORC_PRAGMA_OMP                                                                        \
    for (int64_t b = 0; b < batch; ++b) {                                                 \
      const T* a = x + b * nobs * nv;                                                     \
      T* o = out + b * nobs * nv * nv;                                                    \
      const int64_t P = nv * nv;                                                          \
      const int64_t mc = min_count > 1 ? min_count : 1;                                   \
      T* si = (T*)calloc((size_t)P * 5, sizeof(T));                                       \
      T *sj = si + P, *sqi = si + 2 * P, *sqj = si + 3 * P, *sij = si + 4 * P;             \
      int64_t* cnt = (int64_t*)calloc((size_t)P, sizeof(int64_t));                        \
      for (int64_t t = 0; t < nobs; ++t) {                                                \
        if (t >= window) {                                                                \
          const T* old = a + (t - window) * nv;                                           \
          for (int64_t i = 0; i < nv; ++i) {                                              \
            if (isnan(old[i])) continue;                                                  \
            for (int64_t j = 0; j < nv; ++j) {                                            \
              if (isnan(old[j])) continue;                                                \
              const int64_t p = i * nv + j;                                               \
              si[p] -= old[i];                                                            \
              sj[p] -= old[j];                                                            \
              sqi[p] -= (T)(old[i] * old[i]);                                             \
              sqj[p] -= (T)(old[j] * old[j]);                                             \
              sij[p] -= (T)(old[i] * old[j]);                                             \
              cnt[p] -= 1;                                                                \
            }                                                                             \
          }                                                                               \
        }                                                                                 \
        const T* nw = a + t * nv;                                                         \
        for (int64_t i = 0; i < nv; ++i) {                                                \
          if (isnan(nw[i])) continue;                                                     \
          for (int64_t j = 0; j < nv; ++j) {                                              \
            if (isnan(nw[j])) continue;                                                   \
            const int64_t p = i * nv + j;                                                 \
            si[p] += nw[i];                                                               \
            sj[p] += nw[j];                                                               \
            sqi[p] += (T)(nw[i] * nw[i]);                                                 \
            sqj[p] += (T)(nw[j] * nw[j]);                                                 \
            sij[p] += (T)(nw[i] * nw[j]);                                                 \
            cnt[p] += 1;                                                                  \
          }                                                                               \
        }                                                                                 \
        T* ot = o + t * P;                                                                \
        for (int64_t p = 0; p < P; ++p) {                                                 \
          const int64_t c = cnt[p];                                                       \
          double r = NAN;                                                                 \
          if (corr) {                                                                     \
            if (c >= (mc > 2 ? mc : 2)) {                                                 \
              const double mi = (double)si[p] / (double)c, mj = (double)sj[p] / (double)c; \
              const double vi = (double)sqi[p] / (double)c - mi * mi;                     \
              const double vj = (double)sqj[p] / (double)c - mj * mj;                     \
              const double cv = (double)sij[p] / (double)c - mi * mj;                     \
              if (vi > 0 && vj > 0) r = cv / sqrt(vi * vj);                               \
            }                                                                             \
          } else if (c >= mc && c > 1) {                                                  \
            const double mi = (double)si[p] / (double)c, mj = (double)sj[p] / (double)c;  \
            r = ((double)sij[p] / (double)c - mi * mj) * (double)c / (double)(c - 1);     \
          }                                                                               \
          ot[p] = (T)r;                                                                   \
        }                                                                                 \
      }                                                                                   \
      free(si);                                                                           \
      free(cnt);                                                                          \
    }                                                                                     \
  }                                                                                       \
  void orc_move_corrmatrix_##SFX(const T* x, T* out, int64_t batch, int64_t nobs, int64_t nv, \
                                 int64_t window, int64_t min_count) {                     \
    orc_move_matrix_##SFX(x, out, batch, nobs, nv, window, min_count, 1);                 \
  }                                                                                       \
  void orc_move_covmatrix_##SFX(const T* x, T* out, int64_t batch, int64_t nobs, int64_t nv, \
                                int64_t window, int64_t min_count) {                      \
    orc_move_matrix_##SFX(x, out, batch, nobs, nv, window, min_count, 0);                 \
  }
DEF_MOVE_MATRIX(float, f32)
DEF_MOVE_MATRIX(double, f64)

/* move_exp_nancorrmatrix (numbagg/moving_matrix.py:204-328) and
 * move_exp_nancovmatrix (:331-432): x (batch, nobs, nv), alpha (nobs) per
 * batch row (ld_alpha 0: shared) in the data dtype -> out (batch, nobs, nv,
 * nv). Every statistic decays each step (decay = 1.0 - alpha, float64), the
 * squared-weight sum by decay**2; a pair adds its observation where both
 * values are valid. */
#define DEF_EXP_MATRIX(T, SFX)                                                            \
  static void orc_exp_matrix_##SFX(const T* x, const T* alpha, int64_t ld_alpha, T* out,  \
                                   int64_t batch, int64_t nobs, int64_t nv, T min_weight, int corr) { \
    ORC_PRAGMA_OMP                                                                        \
    for (int64_t b = 0; b < batch; ++b) {                                                 \
      const T* a = x + b * nobs * nv;                                                     \
      const T* al = alpha + b * ld_alpha;                                                 \
      T* o = out + b * nobs * nv * nv;                                                    \
      const int64_t P = nv * nv;                                                          \
      T* si = (T*)calloc((size_t)P * 8, sizeof(T));                                       \
      T *sj = si + P, *sqi = si + 2 * P, *sqj = si + 3 * P, *sij = si + 4 * P;             \
      T *pw = si + 5 * P, *psw = si + 6 * P, *psw2 = si + 7 * P;                          \
      for (int64_t t = 0; t < nobs; ++t) {                                                \
        const T alpha_t = al[t];                                                          \
        const double decay = 1.0 - (double)alpha_t;                                       \
        for (int64_t p = 0; p < P; ++p) {                                                 \
          si[p] = (T)(si[p] * decay);                                                     \
          sj[p] = (T)(sj[p] * decay);                                                     \
          sqi[p] = (T)(sqi[p] * decay);                                                   \
          sqj[p] = (T)(sqj[p] * decay);                                                   \
          sij[p] = (T)(sij[p] * decay);                                                   \
          pw[p] = (T)(pw[p] * decay);                                                     \
          psw[p] = (T)(psw[p] * decay);                                                   \
          psw2[p] = (T)(psw2[p] * (decay * decay));                                       \
        }                                                                                 \
        const T* nw = a + t * nv;                                                         \
        for (int64_t i = 0; i < nv; ++i) {                                                \
          if (isnan(nw[i])) continue;                                                     \
          for (int64_t j = 0; j < nv; ++j) {                                              \
            if (isnan(nw[j])) continue;                                                   \
            const int64_t p = i * nv + j;                                                 \
            si[p] += nw[i];                                                               \
            sj[p] += nw[j];                                                               \
            sqi[p] += (T)(nw[i] * nw[i]);                                                 \
            sqj[p] += (T)(nw[j] * nw[j]);                                                 \
            sij[p] += (T)(nw[i] * nw[j]);                                                 \
            pw[p] += alpha_t;                                                             \
            psw[p] = (T)(psw[p] + 1.0);                                                   \
            psw2[p] = (T)(psw2[p] + 1.0);                                                 \
          }                                                                               \
        }                                                                                 \
        T* ot = o + t * P;                                                                \
        for (int64_t p = 0; p < P; ++p) {                                                 \
          const double bias = psw[p] > 0 ? 1.0 - (double)(T)(psw2[p] / (T)(psw[p] * psw[p])) : 0.0; \
          double r = NAN;                                                                 \
          if (pw[p] >= min_weight && bias > 0) {                                          \
            const T n = psw[p];                                                           \
            const T mi = si[p] / n, mj = sj[p] / n;                                       \
            const T cb = (T)(sij[p] / n - mi * mj);                                       \
            if (!corr) {                                                                  \
              r = (double)cb / bias;                                                      \
            } else {                                                                      \
              const double vi = (double)(T)(sqi[p] / n - mi * mi) / bias;                 \
              const double vj = (double)(T)(sqj[p] / n - mj * mj) / bias;                 \
              if (vi > 0 && vj > 0) r = ((double)cb / bias) / sqrt(vi * vj);              \
            }                                                                             \
          }                                                                               \
          ot[p] = (T)r;                                                                   \
        }                                                                                 \
      }                                                                                   \
      free(si);                                                                           \
    }                                                                                     \
  }                                                                                       \
  void orc_move_exp_nancorrmatrix_##SFX(const T* x, const T* alpha, int64_t ld_alpha, T* out, \
                                        int64_t batch, int64_t nobs, int64_t nv, T min_weight) { \
    orc_exp_matrix_##SFX(x, alpha, ld_alpha, out, batch, nobs, nv, min_weight, 1);        \
  }                                                                                       \
  void orc_move_exp_nancovmatrix_##SFX(const T* x, const T* alpha, int64_t ld_alpha, T* out, \
                                       int64_t batch, int64_t nobs, int64_t nv, T min_weight) { \
    orc_exp_matrix_##SFX(x, alpha, ld_alpha, out, batch, nobs, nv, min_weight, 0);        \
  }
DEF_EXP_MATRIX(float, f32)
DEF_EXP_MATRIX(double, f64)

/* Exact static matrices: per pair, two-pass long-double mean and
 * (co)variance over the pairwise-complete observations. kind 0 cov, 1 corr.
 * x (batch, nv, nobs) float64 -> out (batch, nv, nv) float64. */
void orc_exact_matrix(int kind, const double* x, double* out, int64_t batch, int64_t nv, int64_t nobs) {
  ORC_PRAGMA_OMP
  for (int64_t bp = 0; bp < batch * nv * nv; ++bp) {
    const int64_t b = bp / (nv * nv), i = (bp / nv) % nv, j = bp % nv;
    const double* xi = x + (b * nv + i) * nobs;
    const double* xj = x + (b * nv + j) * nobs;
    long double s1 = 0, s2 = 0;
    int64_t c = 0;
    for (int64_t k = 0; k < nobs; ++k) {
      if (isnan(xi[k]) || isnan(xj[k])) continue;
      s1 += xi[k];
      s2 += xj[k];
      c += 1;
    }
    double r = NAN;
    if (c > 1) {
      const long double m1 = s1 / c, m2 = s2 / c;
      long double s11 = 0, s22 = 0, s12 = 0;
      for (int64_t k = 0; k < nobs; ++k) {
        if (isnan(xi[k]) || isnan(xj[k])) continue;
        const long double d1 = xi[k] - m1, d2 = xj[k] - m2;
        s11 += d1 * d1;
        s22 += d2 * d2;
        s12 += d1 * d2;
      }
      if (kind == 0) {
        r = (double)(s12 / (c - 1));
      } else if (s11 > 0 && s22 > 0) {
        r = (double)(s12 / sqrtl(s11 * s22));
      }
    }
    out[bp] = r;
  }
}
