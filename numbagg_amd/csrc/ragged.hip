// Ragged batched serving path: many float32 series of uneven lengths, each
// with move_mean (window w) and group_nanmean over its own int32 labels, in one
// persistent launch. No reference counterpart: its reference semantics are
// per-series plain calls of move_mean (numbagg/moving.py:12-55) and
// group_nanmean (numbagg/grouped.py:7-27).
//
// Design:
//   * work items are (series, tile) pairs, tiles counted from the series start
//     with exactly the length and arithmetic a plain move_mean launch on that
//     series uses, so the move_mean outputs are bitwise identical to plain
//     launches;
//   * a persistent grid (one resident wave of workgroups) pulls items from an
//     atomic work-queue counter, a few items per grab; every workgroup leaves
//     the loop once the counter passes the last item. The next grab is taken
//     while the current one runs, and a grab's tile descriptors (series
//     offsets, label range, strategy) are read into LDS once, so no tile waits
//     on a chain of dependent lookups;
//   * consecutive tiles of one series within a grab are software-pipelined as
//     in the plain move kernel's spans: the next tile's inputs are loaded into
//     registers while the current tile is evaluated, and its window halo is
//     carried in LDS;
//   * series sit at any element offset of the packed buffers: values, outputs
//     and labels move 16 bytes per access at element alignment;
//   * each item evaluates its rolling tile (halo staged in LDS) and then folds
//     the tile's values into the series' label accumulators, either in
//     LDS-privatised slots (flushed with global atomics when the workgroup
//     moves to another series or its grab ends; labels beyond the slots go to
//     global atomics) or straight into global float64/uint32 accumulators — the
//     strategy is chosen per series by the host (heuristic, autotuner or cost
//     model);
//   * a finalize kernel writes group means (float64 sum / count, rounded once).
#include "rolling.hpp"

namespace nbg {
namespace {

// Work items (tiles) per queue grab. A series' LDS label slots are flushed
// once per grab it appears in, so longer grabs flush less often; shorter ones
// balance the grid's tail (measured on MI355X, C5-shaped batch of 2.1e9
// elements with 4096 label slots: grabs of 4 / 8 / 16 tiles 11.0 / 10.0 /
// 9.75 ms). Knob: NBG_RAGGED_GRAB.
constexpr int kRaggedGrab = 16;
constexpr int kRaggedMaxGrab = 64;  // a grab's tile descriptors are held in LDS
constexpr int64_t kRaggedLdsLabels = 1024;     // default LDS label slots per launch
constexpr int64_t kRaggedMaxLdsLabels = 8192;  // 96 KB of slots

// The ragged kernel evaluates move_mean tiles with exactly the variant a plain
// float32 move_mean launch uses (items per thread, 1/c table), so both produce
// bitwise-identical outputs.
struct RaggedArgs {
  const float* values;
  const int32_t* labels;
  const int64_t* offsets;        // nseries + 1
  const int64_t* label_offsets;  // nseries + 1
  const int64_t* tile_offsets;   // nseries + 1
  const int32_t* tile_series;    // total_tiles
  const uint8_t* strategy;       // nseries: 0 = LDS-privatised, 1 = global atomics
  float* move_out;
  double* gsum;
  unsigned int* gcount;
  unsigned int* queue;
  int64_t total_tiles, window, min_count;
  size_t stage_bytes;  // dynamic LDS used by the rolling tile
  size_t lut_bytes;    // 1/c table after the staged segment (0 when unused)
  int64_t lds_labels;  // LDS label slots (labels [0, lds_labels) of a series)
  int grab;            // tiles per queue grab
};

// Zeroes the label accumulators (float64 sums then uint32 counts: `words`
// 4-byte words from w) and the work queue. The 16-byte-aligned body goes out
// as streaming 16-byte stores, one per thread (the accumulators are GBs for
// a large batch); the few words before and after it as single words.
__global__ __launch_bounds__(256) void ragged_init_kernel(unsigned int* w, int64_t words, unsigned int* queue) {
  const int64_t head0 = (int64_t)(((16 - ((uintptr_t)w & 15)) & 15) / 4);
  const int64_t head = head0 < words ? head0 : words;
  const int64_t body = (words - head) / 4;  // 16-byte chunks
  const int64_t tail0 = head + body * 4;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const unsigned int zero[4] = {0u, 0u, 0u, 0u};
  for (int64_t i = tid; i < body; i += stride) store_stream16(w + head + 4 * i, zero);
  if (tid < head) w[tid] = 0u;
  if (tid < words - tail0) w[tail0 + tid] = 0u;
  if (tid == 0) *queue = 0u;
}

// A tile of the current grab, as read once from the batch description.
struct RaggedTile {
  int64_t off;   // series start in values / move_out
  int64_t n;     // series length
  int64_t lo;    // series start in the label accumulators
  int64_t K;     // series label count
  int tile;      // tile index within the series
  int series;
  int strategy;  // 0: LDS slots for labels below lds_labels, 1: global only
};

template <int VEC, int KP, bool LUT>
__device__ __forceinline__ void ragged_rolling(const RaggedArgs& A, int64_t off, int64_t n,
                                               int64_t tile, unsigned char* smem,
                                               RollingShared<Traits<float, kMean>::S>& rs,
                                               const double* lut, int staged, int full) {
  rolling_tile<float, kMean, true, VEC, KP, LUT>(A.values + off, nullptr, A.move_out + off, n,
                                                 A.window, A.min_count, tile, smem, rs, lut, staged, full);
}

// Flushes the LDS label slots [0, nl) of series `s` into its global
// accumulators (only slots that saw a value), then waits for the workgroup.
__device__ __forceinline__ void ragged_flush(const RaggedArgs& A, int64_t lo, int64_t nl,
                                             const double* l_sum, const unsigned int* l_cnt) {
  for (int64_t l = threadIdx.x; l < nl; l += kMoveThreads) {
    const unsigned int c = l_cnt[l];
    if (c) {
      unsafeAtomicAdd(&A.gsum[lo + l], l_sum[l]);
      atomicAdd(&A.gcount[lo + l], c);
    }
  }
  __syncthreads();
}

template <int KP, bool LUT>
// Held to 3 waves per SIMD (168 VGPRs): the prefetched tile stays in registers
// without spills (measured on MI355X: a C5-shaped batch of 2.1e9 elements with
// 4096 label slots takes 10.0 ms, against 12.1 ms at 4 waves with spills and
// 12.3 ms without the prefetch).
__global__ __launch_bounds__(kMoveThreads) __attribute__((amdgpu_waves_per_eu(3))) void ragged_kernel(RaggedArgs A) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ RollingShared<Traits<float, kMean>::S> rs;
  __shared__ RaggedTile s_tiles[kRaggedMaxGrab];
  __shared__ long long s_next;
  constexpr int TILE = kMoveThreads * KP;
  constexpr int NV = KP / 4;  // 16-byte vectors per thread of one tile
  const int tid = threadIdx.x;
  double* lut = reinterpret_cast<double*>(smem_raw + A.stage_bytes);
  if constexpr (LUT) fill_recip_lut(lut, A.window, A.min_count > 1 ? A.min_count : 1);  // ordered by the loop's first barrier
  double* l_sum = reinterpret_cast<double*>(smem_raw + A.stage_bytes + A.lut_bytes);
  unsigned int* l_cnt = reinterpret_cast<unsigned int*>(l_sum + A.lds_labels);
  float* la = reinterpret_cast<float*>(smem_raw);  // the staged segment
  // The segment of tile t is [floor16(t0 - window), t0 + TILE): H = ceil16(window)
  // halo elements, then the tile's own TILE inputs. The next tile's segment
  // begins with this one's last H elements, so they are carried when H <= TILE.
  const int H = (int)carried_halo(A.window);
  const bool carry = H <= TILE;

  if (tid == 0) s_next = (long long)atomicAdd(A.queue, (unsigned)A.grab);
  while (true) {
    __syncthreads();  // s_next published; the previous grab's LDS reads done
    const int64_t base = s_next;
    if (base >= A.total_tiles) break;  // uniform: every workgroup exits here
    const int cnt = (int)(base + A.grab < A.total_tiles ? A.grab : A.total_tiles - base);
    if (tid < cnt) {
      const int64_t t = base + tid;
      const int s = A.tile_series[t];
      RaggedTile d;
      d.off = A.offsets[s];
      d.n = A.offsets[s + 1] - d.off;
      d.lo = A.label_offsets[s];
      d.K = A.label_offsets[s + 1] - d.lo;
      d.tile = (int)(t - A.tile_offsets[s]);
      d.series = s;
      d.strategy = A.strategy[s];
      s_tiles[tid] = d;
    }
    __syncthreads();  // descriptors visible; every thread has read s_next
    if (tid == 0) s_next = (long long)atomicAdd(A.queue, (unsigned)A.grab);  // next grab, in flight meanwhile

    // LDS label slots stay live across consecutive tiles of one series within
    // the grab and are flushed when the series changes or the grab ends.
    int cur = -1;
    int64_t cur_lo = 0, cur_nl = 0;
    int staged = 0;  // leading segment elements already in LDS (carried halo + prefetched tile)
    float pa[KP];
    for (int i = 0; i < cnt; ++i) {
      const RaggedTile d = s_tiles[i];
      const int s = d.series;
      const int64_t n = d.n;
      const int64_t tile = d.tile;
      const int64_t lo = d.lo;
      const int64_t K = d.K;
      // Strategy 0 keeps labels below lds_labels in LDS (all of them when K
      // fits) and sends the rest to global atomics; strategy 1 is global only.
      const int64_t nl = d.strategy == 0 ? (K < A.lds_labels ? K : A.lds_labels) : 0;
      if (s != cur) {
        if (cur_nl > 0) ragged_flush(A, cur_lo, cur_nl, l_sum, l_cnt);
        for (int64_t l = tid; l < nl; l += kMoveThreads) {
          l_sum[l] = 0.0;
          l_cnt[l] = 0u;
        }
        cur = s;
        cur_lo = lo;
        cur_nl = nl;  // the zeroing is ordered before use by rolling_tile's barriers
      }
      const float* xs = A.values + d.off;
      // 16-byte vector accesses need element alignment only (Vec16), so every
      // series of a float32 batch takes the vector path.
      const bool aligned = (((uintptr_t)xs) % 4 == 0) && (((uintptr_t)(A.move_out + d.off)) % 4 == 0);
      const int64_t t0 = tile * TILE;
      const int full = t0 + TILE <= n ? 1 : 0;
      const int64_t c0 = t0 + (int64_t)tid * KP;
      // The next tile of this series, when it is the grab's next item: its
      // inputs [t0 + TILE, t0 + 2 TILE) are loaded now (segment offsets
      // [H, H + TILE); NaN past the series end), thread tid holding the
      // elements it would stage itself.
      const bool more = carry && aligned && i + 1 < cnt && s_tiles[i + 1].series == s &&
                        s_tiles[i + 1].tile == d.tile + 1;
      if (more) {
        const int64_t t0n = t0 + TILE;
        const bool nfull = t0n + TILE <= n;
#pragma unroll
        for (int m = 0; m < NV; ++m) {
          const int64_t p = t0n + ((int64_t)tid + (int64_t)m * kMoveThreads) * 4;
          if (nfull || p + 4 <= n) {
            const Vec16<float> q = load_stream(reinterpret_cast<const Vec16<float>*>(xs + p));
#pragma unroll
            for (int v = 0; v < 4; ++v) pa[m * 4 + v] = q.v[v];
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) pa[m * 4 + v] = p + v < n ? xs[p + v] : canonical_nan<float>();
          }
        }
      }
      // 1. move_mean of this tile (identical arithmetic to the plain kernel).
      if (aligned)
        ragged_rolling<4, KP, LUT>(A, d.off, n, tile, smem_raw, rs, lut, staged, full);
      else
        ragged_rolling<1, KP, LUT>(A, d.off, n, tile, smem_raw, rs, lut, staged, full);
      // Labels of this thread's KP positions (16-byte loads inside the series;
      // plain loads: the thread's four loads are strided across the wave and
      // share cache lines, which streaming loads would fetch four times).
      const int32_t* lab = A.labels + d.off;
      int32_t lj[KP];
      if (c0 + KP <= n) {
#pragma unroll
        for (int j = 0; j < KP; j += 4) {
          const Vec16<int32_t> q = *reinterpret_cast<const Vec16<int32_t>*>(lab + c0 + j);
#pragma unroll
          for (int v = 0; v < 4; ++v) lj[j + v] = q.v[v];
        }
      } else {
#pragma unroll
        for (int j = 0; j < KP; ++j) lj[j] = c0 + j < n ? lab[c0 + j] : -1;
      }
      // 2. group_nanmean accumulation of the thread's KP values, read from the
      //    staged segment (offsets [H, H + TILE); positions past the series end
      //    hold NaN and carry label -1 here).
      const int ehp = pad16((int)(c0 - ((t0 - A.window) & ~(int64_t)15)));
#pragma unroll
      for (int j = 0; j < KP; ++j) {
        const float x = la[ehp + j + (j >> 4)];
        const int32_t l = lj[j];
        if (l >= 0 && l < K && !is_nan(x)) {
          if (l < nl) {
            atomicAdd(&l_sum[l], (double)x);
            atomicAdd(&l_cnt[l], 1u);
          } else {
            unsafeAtomicAdd(&A.gsum[lo + l], (double)x);
            atomicAdd(&A.gcount[lo + l], 1u);
          }
        }
      }
      if (more) {
        // Carry the halo: segment offsets [TILE, TILE + H) -> [0, H). The
        // group reads above use offsets [H, H + TILE) only, so the copy may
        // overlap them; the barrier then orders the copy's reads and the
        // group reads before the prefetched inputs overwrite [H, H + TILE).
        for (int e = tid; e < H; e += kMoveThreads) la[pad16(e)] = la[pad16(e + TILE)];
        __syncthreads();
#pragma unroll
        for (int m = 0; m < NV; ++m) {
          const int e = H + (tid + m * kMoveThreads) * 4;
#pragma unroll
          for (int v = 0; v < 4; ++v) la[pad16(e + v)] = pa[m * 4 + v];
        }
        staged = H + TILE;  // fully staged; rolling_tile's first barrier publishes these writes
      } else {
        __syncthreads();  // staged segment reads done before the next tile stages
        staged = 0;
      }
    }
    if (cur_nl > 0) ragged_flush(A, cur_lo, cur_nl, l_sum, l_cnt);
  }
}

__global__ void ragged_finalize_kernel(const double* gsum, const unsigned int* gcount, float* out,
                                       int64_t total_labels) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total_labels; i += stride) {
    const unsigned int c = gcount[i];
    out[i] = c ? (float)(gsum[i] / (double)c) : canonical_nan<float>();
  }
}

size_t ragged_ws_bytes(int64_t total_labels) {
  return (size_t)total_labels * 8 + (size_t)total_labels * 4 + 256;
}

}  // namespace
}  // namespace nbg

extern "C" {

int64_t nbg_ragged_tile_size(void) { return (int64_t)nbg::kMoveThreads * nbg::f32_sum_items(); }

int64_t nbg_ragged_max_lds_labels(void) { return nbg::kRaggedMaxLdsLabels; }

size_t nbg_ragged_workspace_bytes(int64_t nseries, int64_t total_labels) {
  (void)nseries;
  return nbg::ragged_ws_bytes(total_labels);
}

int nbg_ragged_move_mean_group_nanmean(const float* values, const int32_t* labels,
                                       const int64_t* offsets, int64_t nseries,
                                       const int64_t* label_offsets, int64_t total_labels,
                                       const int64_t* tile_offsets, const int32_t* tile_series,
                                       int64_t total_tiles, const uint8_t* strategy,
                                       int64_t lds_labels, int64_t window, int64_t min_count, float* move_out,
                                       float* group_out, void* ws, size_t ws_bytes, void* stream) {
  using namespace nbg;
  if (int rc = injected_error("nbg_ragged_move_mean_group_nanmean")) return rc;
  if (nseries < 0 || total_labels < 0 || total_tiles < 0)
    return set_error(NBG_ERR_INVALID, "ragged: negative size");
  if (window <= 0) return set_error(NBG_ERR_INVALID, "ragged: window not in valid range: %lld", (long long)window);
  if (min_count < 0) return set_error(NBG_ERR_INVALID, "ragged: min_count must be positive: %lld", (long long)min_count);
  if (lds_labels == 0) lds_labels = kRaggedLdsLabels;
  if (lds_labels < 1 || lds_labels > kRaggedMaxLdsLabels)
    return set_error(NBG_ERR_INVALID, "ragged: lds_labels %lld not in [1, %lld]", (long long)lds_labels,
                     (long long)kRaggedMaxLdsLabels);
  if (window > lds_window_limit<float, 1>())
    return set_error(NBG_ERR_INVALID, "ragged: window %lld exceeds the staged-halo limit %lld",
                     (long long)window, (long long)lds_window_limit<float, 1>());
  if (nseries == 0) return NBG_OK;
  if (!values || !labels || !offsets || !label_offsets || !tile_offsets || !tile_series || !strategy ||
      !move_out || (!group_out && total_labels))
    return set_error(NBG_ERR_INVALID, "ragged: null pointer");
  const size_t need = ragged_ws_bytes(total_labels);
  if (ws == nullptr || ws_bytes < need)
    return set_error(NBG_ERR_INVALID, "ragged: workspace too small (%zu < %zu)", ws_bytes, need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  RaggedArgs a{};
  a.values = values;
  a.labels = labels;
  a.offsets = offsets;
  a.label_offsets = label_offsets;
  a.tile_offsets = tile_offsets;
  a.tile_series = tile_series;
  a.strategy = strategy;
  a.move_out = move_out;
  a.gsum = static_cast<double*>(ws);
  a.gcount = reinterpret_cast<unsigned int*>(a.gsum + total_labels);
  a.queue = reinterpret_cast<unsigned int*>(static_cast<char*>(ws) + need - 256);
  a.total_tiles = total_tiles;
  a.window = window;
  a.min_count = min_count;
  const int kp = f32_sum_items();
  const bool lut = mean_uses_lut(window);
  a.stage_bytes = (size_t)rolling_stage_bytes<float>(kp, window, true, 1);
  a.lut_bytes = lut ? ((size_t)(window + 1) * sizeof(double) + 15) / 16 * 16 : 0;
  a.lds_labels = lds_labels;
  const size_t lds = a.stage_bytes + a.lut_bytes + (size_t)lds_labels * (8 + 4);
  // init: one 16-byte store per thread (0.38 ms for the 2.5 GB of a 212M-label
  // batch, against 0.61 ms for a grid-stride loop of 8- and 4-byte stores)
  const int64_t words = total_labels * 3;  // float64 sums and uint32 counts, back to back
  int64_t init_blocks = (words / 4 + 1 + 255) / 256;
  if (init_blocks > 0x7fffffffLL) init_blocks = 0x7fffffffLL;
  hipLaunchKernelGGL(ragged_init_kernel, dim3((unsigned)init_blocks), dim3(256), 0, s,
                     reinterpret_cast<unsigned int*>(a.gsum), words, a.queue);
  // finalize: grid-stride over at most 4096 workgroups (measured faster
  // than one label per thread: 0.58 against 0.71 ms for 212M labels)
  int64_t fin_blocks = (total_labels + 255) / 256;
  if (fin_blocks > 4096) fin_blocks = 4096;
  if (fin_blocks < 1) fin_blocks = 1;
  NBG_HIP_TRY(hipGetLastError());
  const void* kfn = kp == 32 ? (lut ? (const void*)ragged_kernel<32, true> : (const void*)ragged_kernel<32, false>)
                             : (lut ? (const void*)ragged_kernel<16, true> : (const void*)ragged_kernel<16, false>);
  int cap = device_capacity_blocks(kfn, kMoveThreads, lds);
  if (cap <= 0)
    return set_error(NBG_ERR_INVALID, "ragged: no workgroup fits (%zu bytes of LDS for window %lld and %lld label slots)",
                     lds, (long long)window, (long long)lds_labels);
  const long long grab = knob("NBG_RAGGED_GRAB", kRaggedGrab);
  if (grab < 1 || grab > kRaggedMaxGrab)
    return set_error(NBG_ERR_INVALID, "ragged: NBG_RAGGED_GRAB %lld not in [1, %d]", grab, kRaggedMaxGrab);
  a.grab = (int)grab;
  int64_t grid = total_tiles < cap ? total_tiles : cap;
  if (grid < 1) grid = 1;
  // The 32-bit queue counter ends at most one grab per workgroup past the
  // last tile.
  if (total_tiles + grid * (int64_t)grab > 0xffffffffLL)
    return set_error(NBG_ERR_INVALID, "ragged: %lld tiles exceed the work-queue range", (long long)total_tiles);
  if (kp == 32) {
    if (lut) hipLaunchKernelGGL((ragged_kernel<32, true>), dim3((unsigned)grid), dim3(kMoveThreads), lds, s, a);
    else hipLaunchKernelGGL((ragged_kernel<32, false>), dim3((unsigned)grid), dim3(kMoveThreads), lds, s, a);
  } else {
    if (lut) hipLaunchKernelGGL((ragged_kernel<16, true>), dim3((unsigned)grid), dim3(kMoveThreads), lds, s, a);
    else hipLaunchKernelGGL((ragged_kernel<16, false>), dim3((unsigned)grid), dim3(kMoveThreads), lds, s, a);
  }
  NBG_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(ragged_finalize_kernel, dim3((unsigned)fin_blocks), dim3(256), 0, s, a.gsum,
                     a.gcount, group_out, total_labels);
  NBG_HIP_TRY(hipGetLastError());
  return NBG_OK;
}

}  // extern "C"
