// Moving-window statistics (move_sum / mean / std / var / cov / corr) for gfx950.
//
// Reference semantics (numbagg/moving.py:12-275): out[i] is a statistic of the
// valid observations in the window [i-window+1, i] (pairwise-complete for the
// two-input kernels), or NaN when their count is below the effective
// min_count (sum: min_count; mean, corr: max(min_count, 1); std, var, cov:
// max(min_count, 2)). The reference carries running float64 sums and adds /
// removes one observation per step; here every window is evaluated from
// compensated sums, so results sit closer to the exact value while the NaN
// mask (a pure function of the valid counts) is the same.
//
// Design (no cross-workgroup communication): a workgroup owns an output tile
// [t0, t0+T) of one row, T = 256 threads x K items. It stages the input
// segment [t0-window, t0+T) in LDS (padded one slot per 16 to keep the
// per-thread strided reads conflict-free). Thread t owns outputs
// [t0+tK, t0+tK+K):
//   1. delta_t = stats(head chunk) - stats(tail chunk), where the tail chunk is
//      the K inputs `window` positions earlier;
//   2. W0 = stats of the halo [t0-window, t0) (block reduction) = the window
//      ending at t0-1;
//   3. an exclusive block scan of delta_t gives each thread the window ending
//      just before its chunk; K sequential add/remove steps produce its outputs.
// For windows too long for LDS the tail and halo are read from global memory.
//
// Accumulators: first moments of float32 data for move_sum/mean are float64
// (exact differences head - tail; FP32 TwoSum pairs with
// NBG_F32_ACCUMULATOR=twosum); the moment kernels (var/std/cov/corr) work on
// data shifted by a per-tile constant with float64 first and second moments;
// float64 data uses double-double throughout (TwoProd for squares and cross
// products). Outputs of full move_sum / move_mean tiles pass through LDS so
// that the stores are contiguous across each wave (and stream past the
// caches).
#include <cstdlib>
#include <cstring>

#include "rolling.hpp"

namespace nbg {
namespace {

template <typename T, int KIND, int KP>
int move_launch_k(const T* a, const T* b, T* out, int64_t rows, int64_t n, int64_t ld_a,
                  int64_t ld_b, int64_t ld_out, int64_t window, int64_t min_count, hipStream_t s) {
  using Tr = Traits<T, KIND>;
  constexpr int TILE = kMoveThreads * KP;
  const int64_t tiles = (n + TILE - 1) / TILE;
  // Spans of consecutive tiles per workgroup: about 96 spans per CU, at most
  // 6 tiles each (measured on (256, 2M) float32 windows of 1000: spans of
  // 4-6 tiles 0.843-0.848 ms, 8 tiles 0.854, 16 tiles 0.873, 2 tiles 0.948 —
  // longer spans stage fewer halos but leave a longer tail of the grid).
  // NBG_MOVE_SPAN=<k> forces k, up to kMaxSpan (results do not depend on it;
  // tests use it to run carried halos on small inputs).
  int64_t span = rows * tiles / (256 * 96);
  span = span > 6 ? 6 : span;
  if (const long long k = knob("NBG_MOVE_SPAN", 0)) span = k;
  span = span < 1 ? 1 : (span > kMaxSpan ? kMaxSpan : span);
  if (span > tiles) span = tiles;
  const int64_t spans = (tiles + span - 1) / span;
  const int64_t blocks = rows * spans;
  if (blocks > 0x7fffffffLL || tiles > 0x7fffffffLL) return set_error(NBG_ERR_INVALID, "move: too many tiles");
  const bool tail_lds = window <= lds_window_limit<T, Tr::NARGS>() - (KP - default_items<T>()) * kMoveThreads;
  // 16-byte vector loads and stores need element alignment only, so rows at
  // any element offset or stride take the vector path.
  constexpr int MAXV = 16 / (int)sizeof(T);
  const int v_use = pick_vec_elem<T>(a, Tr::NARGS == 2 ? b : a, out);
  // a reciprocal table in LDS (float32): 1/c for mean (NaN below the
  // effective min_count), 1/max(c, 1) for the moment kinds, read instead of
  // refined reciprocals per output, which also frees registers (move_var 1.26
  // -> 1.20 ms, move_std 1.41 -> 1.31, move_cov 1.92 -> 1.88 on (256, 2M),
  // window 1000)
  constexpr bool LUT_OK = KIND != kSum && KIND < 16 && sizeof(T) == 4;
  const bool use_lut = LUT_OK && mean_uses_lut(window);
  size_t lds = (size_t)rolling_stage_bytes<T>(KP, window, tail_lds, Tr::NARGS);
  if (use_lut) lds += (size_t)(window + 1) * sizeof(double);
  dim3 grid((unsigned)blocks), block(kMoveThreads);
  // Each next tile of a span is prefetched into registers while the current
  // one is evaluated (more VGPRs, fewer resident waves; measured faster on
  // MI355X); NBG_MOVE_PREFETCH=0 turns it off.
  const bool pf = knob("NBG_MOVE_PREFETCH", 1) != 0;
#define NBG_MOVE_LAUNCH(TL, V, LU)                                                                        \
  do {                                                                                                    \
    if (pf && TL && Tr::NARGS == 1)                                                                       \
      hipLaunchKernelGGL((move_kernel<T, KIND, TL, V, KP, LU, TL>), grid, block, lds, s, a, b, out, rows, \
                         n, ld_a, ld_b, ld_out, window, min_count, tiles, span);                          \
    else                                                                                                  \
      hipLaunchKernelGGL((move_kernel<T, KIND, TL, V, KP, LU, false>), grid, block, lds, s, a, b, out,    \
                         rows, n, ld_a, ld_b, ld_out, window, min_count, tiles, span);                    \
  } while (0)
  if (use_lut) {
    if constexpr (LUT_OK) {
      if (tail_lds) {
        if (v_use == MAXV) NBG_MOVE_LAUNCH(true, MAXV, true);
        else NBG_MOVE_LAUNCH(true, 1, true);
      } else {
        if (v_use == MAXV) NBG_MOVE_LAUNCH(false, MAXV, true);
        else NBG_MOVE_LAUNCH(false, 1, true);
      }
    }
  } else if (tail_lds) {
    if (v_use == MAXV) NBG_MOVE_LAUNCH(true, MAXV, false);
    else NBG_MOVE_LAUNCH(true, 1, false);
  } else {
    if (v_use == MAXV) NBG_MOVE_LAUNCH(false, MAXV, false);
    else NBG_MOVE_LAUNCH(false, 1, false);
  }
#undef NBG_MOVE_LAUNCH
  return check_hip(hipGetLastError(), "move_kernel launch");
}


template <typename T, int KIND>
int move_launch(const T* a, const T* b, T* out, int64_t rows, int64_t n, int64_t ld_a,
                int64_t ld_b, int64_t ld_out, int64_t window, int64_t min_count, hipStream_t s) {
  if constexpr (sizeof(T) == 4 && (KIND == kSum || KIND == kMean)) {
    if (f32_sum_items() == 32)
      return move_launch_k<T, KIND, 32>(a, b, out, rows, n, ld_a, ld_b, ld_out, window, min_count, s);
  }
  return move_launch_k<T, KIND, default_items<T>()>(a, b, out, rows, n, ld_a, ld_b, ld_out, window,
                                                     min_count, s);
}

// NBG_F32_ACCUMULATOR=twosum selects FP32 TwoSum accumulators for float32
// sum/mean instead of float64 (the float64 default is both more precise and
// cheaper on CDNA4, where a float64 add issues at the float32 rate).
bool f32_twosum_requested() {
  const char* v = getenv("NBG_F32_ACCUMULATOR");
  return v != nullptr && strcmp(v, "twosum") == 0;
}

template <int KIND>
int move_entry(const char* name, int dtype, const void* a, const void* b, void* out, int64_t rows,
               int64_t n, int64_t ld_a, int64_t ld_b, int64_t ld_out, int64_t window,
               int64_t min_count, void* stream) {
  if (int rc = injected_error(name)) return rc;
  if (rows < 0 || n < 0) return set_error(NBG_ERR_INVALID, "%s: negative shape", name);
  if (min_count < 0) return set_error(NBG_ERR_INVALID, "%s: min_count must be positive: %lld", name, (long long)min_count);
  if (!(window > 0 && window <= n)) return set_error(NBG_ERR_INVALID, "%s: window not in valid range: %lld", name, (long long)window);
  if (rows == 0 || n == 0) return NBG_OK;
  constexpr int NARGS = (KIND == kCov || KIND == kCorr) ? 2 : 1;
  if (a == nullptr || out == nullptr || (NARGS == 2 && b == nullptr))
    return set_error(NBG_ERR_INVALID, "%s: null pointer", name);
  if (rows > 1 && (ld_a < n || ld_out < n || (NARGS == 2 && ld_b < n)))
    return set_error(NBG_ERR_INVALID, "%s: row stride < n", name);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case NBG_F32:
      if constexpr (KIND == kSum || KIND == kMean) {
        if (f32_twosum_requested())
          return move_launch<float, KIND + 16>((const float*)a, (const float*)b, (float*)out, rows,
                                               n, ld_a, ld_b, ld_out, window, min_count, s);
      }
      return move_launch<float, KIND>((const float*)a, (const float*)b, (float*)out, rows, n, ld_a,
                                      ld_b, ld_out, window, min_count, s);
    case NBG_F64:
      return move_launch<double, KIND>((const double*)a, (const double*)b, (double*)out, rows, n,
                                       ld_a, ld_b, ld_out, window, min_count, s);
    default:
      return set_error(NBG_ERR_INVALID, "%s: unsupported dtype code %d", name, dtype);
  }
}

}  // namespace
}  // namespace nbg

extern "C" {

int nbg_move_sum(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld_x,
                 int64_t ld_out, int64_t window, int64_t min_count, void* stream) {
  return nbg::move_entry<nbg::kSum>("nbg_move_sum", dtype, x, nullptr, out, rows, n, ld_x, 0,
                                    ld_out, window, min_count, stream);
}
int nbg_move_mean(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld_x,
                  int64_t ld_out, int64_t window, int64_t min_count, void* stream) {
  return nbg::move_entry<nbg::kMean>("nbg_move_mean", dtype, x, nullptr, out, rows, n, ld_x, 0,
                                     ld_out, window, min_count, stream);
}
int nbg_move_var(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld_x,
                 int64_t ld_out, int64_t window, int64_t min_count, void* stream) {
  return nbg::move_entry<nbg::kVar>("nbg_move_var", dtype, x, nullptr, out, rows, n, ld_x, 0,
                                    ld_out, window, min_count, stream);
}
int nbg_move_std(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld_x,
                 int64_t ld_out, int64_t window, int64_t min_count, void* stream) {
  return nbg::move_entry<nbg::kStd>("nbg_move_std", dtype, x, nullptr, out, rows, n, ld_x, 0,
                                    ld_out, window, min_count, stream);
}
int nbg_move_cov(int dtype, const void* a, const void* b, void* out, int64_t rows, int64_t n,
                 int64_t ld_a, int64_t ld_b, int64_t ld_out, int64_t window, int64_t min_count,
                 void* stream) {
  return nbg::move_entry<nbg::kCov>("nbg_move_cov", dtype, a, b, out, rows, n, ld_a, ld_b, ld_out,
                                    window, min_count, stream);
}
int nbg_move_corr(int dtype, const void* a, const void* b, void* out, int64_t rows, int64_t n,
                  int64_t ld_a, int64_t ld_b, int64_t ld_out, int64_t window, int64_t min_count,
                  void* stream) {
  return nbg::move_entry<nbg::kCorr>("nbg_move_corr", dtype, a, b, out, rows, n, ld_a, ld_b,
                                     ld_out, window, min_count, stream);
}

}  // extern "C"
