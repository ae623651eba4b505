// ffill / bfill for gfx950.
//
// Reference semantics (numbagg/funcs.py:294-326): walking the row in scan order
// (ascending for ffill, descending for bfill), a valid value resets a
// `lives_remaining = limit` budget and becomes the current value; each NaN costs
// one life, and once the budget is spent the current value becomes NaN. In closed
// form: out[i] = x[j] where j is the last valid index at or before i in scan
// order, provided i - j <= limit (distance in scan order); otherwise NaN (the
// canonical quiet NaN, since the reference stores `np.nan`). Integer rows are
// copied through by the host layer (isnan is false for integers).
//
// Design: a row is cut into tiles of 256 threads x U x VEC elements. Each
// workgroup takes a ticket from its row's counter, so tiles of a row start in
// scan order and a tile only ever waits on tiles that are already running
// (decoupled look-back, deadlock-free whatever the residency). Blocks are
// interleaved over rows (by default tile t of every row before tile t + 1 of
// any), so a tile's predecessor was dispatched `rows` blocks earlier and has
// normally published its prefix already; its status word is read while the
// tile's own loads are in flight.
//
// Inside a tile, loads are wave-striped vectors (every load instruction of a wave
// covers 64*VEC contiguous elements); the carry "last valid (offset, value)" is
// found per 64*VEC row from a ballot of the lanes holding a valid value,
// across waves through LDS and across tiles through one 64-bit status word per
// tile:
//   bits 63..62  flag (1 = aggregate: tile holds no valid value;
//                      2 = inclusive prefix)
//   bits 61..0   1 + scan-order index of the last valid value (0 = none).
// For float32 rows shorter than 2**30 the word also carries the value
// (PACK): bits 61..32 = 1 + index, bits 31..0 = the value's bits, so the tile
// that finds its carry has its value without a second, dependent load.
#include <cstdlib>
#include <cstring>

#include "common.hpp"

namespace nbg {
namespace {

constexpr int kFillWaves = 4;
constexpr int kFillThreads = kFillWaves * kWave;

constexpr unsigned long long kFlagA = 1ull << 62;
constexpr unsigned long long kFlagP = 2ull << 62;
constexpr unsigned long long kIdxMask = (1ull << 62) - 1;
constexpr int64_t kPackMaxN = (1ll << 30) - 2;  // PACK: 1 + index fits 30 bits

__device__ __forceinline__ unsigned long long pack_carry(long long idx, float v) {
  return ((unsigned long long)(idx + 1) << 32) | (unsigned long long)__float_as_uint(v);
}

template <typename T>
struct alignas(16) VecT;
template <>
struct alignas(16) VecT<float> { float v[4]; };
template <>
struct alignas(16) VecT<double> { double v[2]; };

// Running "last valid value" carry: scan-order offset (or -1) plus the value.
template <typename T>
struct Last {
  int so;
  T val;
};

// Wave step of the carry scan for one slot. Each lane holds `agg`, the last
// valid value of its own elements (so < 0: none); scan order runs up the
// lanes (ffill) or down them (bfill, REV), and scan offsets are monotonic in
// the lane. A ballot of the lanes holding a valid value gives every lane its
// nearest earlier such lane (`excl`, so < 0: none) with two lane reads, and
// the wave's last one in scan order (`tail`) with two scalar reads.
template <bool REV, typename T>
__device__ __forceinline__ void wave_last_carry(const Last<T>& agg, Last<T>& excl, Last<T>& tail) {
  const int lane = lane_id();
  const unsigned long long m = __ballot(agg.so >= 0);
  unsigned long long before;
  if (REV) before = lane == kWave - 1 ? 0ull : m & (~0ull << (lane + 1));
  else before = m & ((1ull << lane) - 1ull);
  const int src = before == 0ull ? -1 : (REV ? __ffsll((long long)before) - 1 : 63 - __clzll((long long)before));
  // The lane reads run on every lane (a lane read from a lane outside the
  // active set returns 0), and the result is selected afterwards.
  const int from = src < 0 ? lane : src;
  const int oso = __shfl(agg.so, from, kWave);
  const T oval = __shfl(agg.val, from, kWave);
  excl.so = src < 0 ? -1 : oso;
  excl.val = oval;
  const int t = m == 0ull ? 0 : (REV ? __ffsll((long long)m) - 1 : 63 - __clzll((long long)m));
  const int tso = read_lane(agg.so, t);
  const T tval = read_lane(agg.val, t);
  tail.so = m == 0ull ? -1 : tso;
  tail.val = tval;
}

// Per-tile summary written by the independent first pass of the two-pass
// fill, one 16-byte store per tile: scan-order offsets within the tile.
struct alignas(16) FillSummary {
  double val;  // the tile's last valid value
  int last;    // its offset, -1: none
  int first;   // offset of the tile's first valid value, -1: none
};

// Dispatch order of a row's tiles: runs of `run` consecutive tiles of one row,
// rows taking turns (run = 1: row-interleaved, tile t of every row before tile
// t + 1 of any; run = tiles: row-major). Block b gets its row and the tile it
// will most likely draw as its ticket; the mapping is a bijection onto
// rows x tiles (the last, shorter runs are interleaved the same way).
__device__ __forceinline__ void fill_block_coords(uint32_t b, int64_t rows, int64_t tiles, int64_t run,
                                                  int64_t& row, int64_t& tile) {
  const uint32_t r = (uint32_t)rows, g = (uint32_t)run;
  const uint32_t full = (uint32_t)tiles / g * g;
  const uint32_t head = r * full;  // rows * tiles < 2**31
  if (b < head) {
    const uint32_t per = r * g;
    const uint32_t sup = b / per, rem = b % per;
    row = rem / g;
    tile = (int64_t)sup * g + rem % g;
  } else {
    const uint32_t tl = (uint32_t)tiles - full, b2 = b - head;
    row = b2 / tl;
    tile = full + b2 % tl;
  }
}

// INDEP: the first pass of the two-pass fill. Tiles take no ticket and no
// carry (a tile's leading NaN run stays NaN) and write their FillSummary;
// fill_fix_kernel then fills the leading runs.
template <typename T, int VEC, int U, bool REV, bool LOOKBACK, bool PACK = false, bool INDEP = false>
__global__ __launch_bounds__(kFillThreads) void fill_kernel(
    const T* __restrict__ x, T* __restrict__ out, int64_t rows, int64_t n, int64_t ld_x,
    int64_t ld_out, int64_t limit, int64_t tiles_per_row, unsigned int* __restrict__ counters,
    unsigned long long* __restrict__ status, unsigned int* __restrict__ timeouts, int64_t run,
    FillSummary* __restrict__ sums = nullptr) {
  static_assert(!(INDEP && LOOKBACK), "the independent pass takes no look-back");
  constexpr int TILE = kFillThreads * U * VEC;
  constexpr int kNone = 0x7fffffff;
  __shared__ int s_ticket;
  __shared__ int s_wfirst[kFillWaves];
  __shared__ int s_wso[kFillWaves];
  __shared__ T s_wval[kFillWaves];
  __shared__ long long s_carry;  // scan-order absolute index of the carry, -1 = none
  __shared__ T s_carry_val;
  __shared__ int s_first_valid;

  // INDEP (row-major): consecutive tiles run on one XCD, so the summaries of
  // neighbouring tiles, which share cache lines, are written through one L2.
  int64_t row, guess;
  fill_block_coords(INDEP ? xcd_order(blockIdx.x, gridDim.x) : blockIdx.x, rows, tiles_per_row, run, row,
                    guess);
  const T* xr = x + row * ld_x;
  T* orow = out + row * ld_out;
  const int w = wave_id();
  const int l = lane_id();
  T vals[U][VEC];
  auto load_tile = [&](int64_t tk) {
    const int64_t tl = REV ? (tiles_per_row - 1 - tk) : tk;
    const int64_t q0 = tl * TILE;
    const bool q_full = (q0 + TILE <= n);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e0 = ((w * U + u) * kWave + l) * VEC;
      if (q_full && VEC > 1) {
        VecT<T> t = load_stream(reinterpret_cast<const VecT<T>*>(xr + q0 + e0));
#pragma unroll
        for (int v = 0; v < VEC; ++v) vals[u][v] = t.v[v];
      } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const int64_t p = q0 + e0 + v;
          vals[u][v] = (p < n) ? xr[p] : canonical_nan<T>();
        }
      }
    }
  };

  // Blocks of a row are dispatched in order, so a block almost always receives
  // the ticket of its dispatch position in the row: its tile is loaded while
  // the ticket atomic is in flight, and reloaded only when the ticket differs.
  int64_t ticket = guess;
  // Thread 0 also reads the status word of the likely predecessor (tile
  // guess - 1) now: when it already holds an inclusive prefix, the look-back
  // after the scan needs no further round trip.
  unsigned long long early = 0ull;
  if (LOOKBACK) {
    if (threadIdx.x == 0) {
      s_ticket = (int)atomicAdd(&counters[row], 1u);
      if (guess > 0) early = ld_agent(&status[row * tiles_per_row + guess - 1]);
    }
    load_tile(guess);
    __syncthreads();
    ticket = s_ticket;
    if (ticket != guess) load_tile(ticket);
  } else {
    load_tile(guess);
  }
  // Tiles are aligned to the row start; bfill walks them from the last one.
  const int64_t tile = REV ? (tiles_per_row - 1 - ticket) : ticket;
  const int64_t p0 = tile * TILE;
  // Scan-order coordinate of the tile's first element (may be < 0 for the
  // partial last tile in bfill; only real elements are ever addressed).
  const int64_t s_base = REV ? (n - p0 - TILE) : p0;
  const bool full = (p0 + TILE <= n);

  // Scan-order offset inside the tile of element (u, v) of this lane.
  auto so_of = [&](int u, int v) -> int {
    const int e = ((w * U + u) * kWave + l) * VEC + v;
    return REV ? (TILE - 1 - e) : e;
  };

  // Per-slot exclusive carries (from earlier lanes / slots of this wave).
  Last<T> cin[U];
  Last<T> wcarry{-1, T(0)};
  int wfirst = kNone;  // INDEP: the wave's first valid offset in scan order
#pragma unroll
  for (int i = 0; i < U; ++i) {
    const int u = REV ? (U - 1 - i) : i;  // scan order over slots
    Last<T> agg{-1, T(0)};
    int fso = kNone;
#pragma unroll
    for (int v = 0; v < VEC; ++v) {
      if (!is_nan(vals[u][v])) {
        const int so = so_of(u, v);
        if (so > agg.so) {
          agg.so = so;
          agg.val = vals[u][v];
        }
        fso = so < fso ? so : fso;
      }
    }
    Last<T> excl, tail;
    wave_last_carry<REV>(agg, excl, tail);
    if (excl.so < 0) excl = wcarry;
    cin[u] = excl;
    if (tail.so >= 0) wcarry = tail;
    if constexpr (INDEP) {
      // slots run in scan order, so the first slot holding a valid value
      // gives the wave's first: its earliest lane in scan order
      const unsigned long long m = __ballot(fso != kNone);
      const int f = m == 0ull ? 0 : (REV ? 63 - __clzll((long long)m) : __ffsll((long long)m) - 1);
      const int ff = read_lane(fso, f);
      if (wfirst == kNone && m != 0ull) wfirst = ff;
    }
  }
  if (l == 0) {
    s_wso[w] = wcarry.so;
    s_wval[w] = wcarry.val;
    if constexpr (INDEP) s_wfirst[w] = wfirst;
  }
  // The owner of scan offset 0 reports whether that element is valid.
  if (REV ? (w == kFillWaves - 1 && l == kWave - 1) : (w == 0 && l == 0))
    s_first_valid = !is_nan(vals[REV ? U - 1 : 0][REV ? VEC - 1 : 0]);
  __syncthreads();

  // Carry from earlier waves (scan order) and the tile aggregate.
  Last<T> wave_in{-1, T(0)};
  Last<T> tile_agg{-1, T(0)};
#pragma unroll
  for (int i = 0; i < kFillWaves; ++i) {
    const int ww = REV ? (kFillWaves - 1 - i) : i;
    if (s_wso[ww] >= 0) {
      if ((REV && ww > w) || (!REV && ww < w)) {
        wave_in.so = s_wso[ww];
        wave_in.val = s_wval[ww];
      }
      tile_agg.so = s_wso[ww];
      tile_agg.val = s_wval[ww];
    }
  }

  // INDEP: the tile's summary, by thread 0.
  if (INDEP && threadIdx.x == 0) {
    int first = kNone;
#pragma unroll
    for (int i = 0; i < kFillWaves; ++i) first = s_wfirst[i] < first ? s_wfirst[i] : first;
    FillSummary sm;
    sm.val = (double)tile_agg.val;
    sm.last = tile_agg.so;
    sm.first = first != kNone ? first : -1;
    sums[row * tiles_per_row + tile] = sm;
  }

  // Cross-tile carry (decoupled look-back), by thread 0.
  if (threadIdx.x == 0) {
    long long carry = -1;
    T cval = T(0);
    bool have_val = false;
    if (LOOKBACK) {
      unsigned long long* srow = status + row * tiles_per_row;
      // Status is indexed by scan-order ticket.
      if (tile_agg.so >= 0) {
        const long long idx = s_base + tile_agg.so;
        if constexpr (PACK) st_agent(&srow[ticket], kFlagP | pack_carry(idx, tile_agg.val));
        else st_agent(&srow[ticket], kFlagP | (unsigned long long)(idx + 1));
      } else if (ticket > 0) {
        st_agent(&srow[ticket], kFlagA);
      } else {
        st_agent(&srow[ticket], kFlagP);
      }
      // The carry matters unless the tile's first element in scan order is valid.
      if (ticket > 0 && !s_first_valid) {
        int64_t j = ticket - 1;
        unsigned long long found = 0ull;  // payload of the first prefix found
        // An inclusive prefix is final once published, so the early read is
        // used as is; anything else (not yet published, or a tile without a
        // valid value) goes through the look-back loop.
        const bool early_hit = ticket == guess && (early >> 62) == 2;
        if (early_hit) found = early & kIdxMask;
        while (!early_hit) {
          unsigned long long st;
          unsigned int spins = 0;
          while (((st = ld_agent(&srow[j])) >> 62) == 0) {
            spin_pause();
            if (++spins > kSpinLimit) {
              atomicAdd(timeouts, 1u);
              st = kFlagP;
              break;
            }
          }
          if ((st >> 62) == 2) {
            found = st & kIdxMask;
            break;
          }
          if (j == 0) break;
          --j;
        }
        if constexpr (PACK) {
          carry = (long long)(found >> 32) - 1;
          if (carry >= 0) {
            cval = __uint_as_float((unsigned int)(found & 0xffffffffull));
            have_val = true;
          }
        } else {
          carry = (long long)found - 1;
        }
        if (tile_agg.so < 0) st_agent(&srow[ticket], kFlagP | found);
      }
    }
    s_carry = carry;
    if (carry >= 0) {
      if (have_val) {
        s_carry_val = cval;
      } else {
        const int64_t pc = REV ? (n - 1 - carry) : carry;
        s_carry_val = xr[pc];
      }
    }
  }
  __syncthreads();
  const long long carry = s_carry;
  const T carry_val = (carry >= 0) ? s_carry_val : T(0);

  // Outputs.
#pragma unroll
  for (int u = 0; u < U; ++u) {
    // Best carry-in for this slot: tile carry < earlier waves < earlier lanes/slots.
    long long run_abs = carry;
    T run_val = carry_val;
    if (wave_in.so >= 0) {
      run_abs = s_base + wave_in.so;
      run_val = wave_in.val;
    }
    if (cin[u].so >= 0) {
      run_abs = s_base + cin[u].so;
      run_val = cin[u].val;
    }
    T res[VEC];
#pragma unroll
    for (int i = 0; i < VEC; ++i) {
      const int v = REV ? (VEC - 1 - i) : i;
      const int64_t s_abs = s_base + so_of(u, v);
      if (!is_nan(vals[u][v])) {
        run_abs = s_abs;
        run_val = vals[u][v];
      }
      res[v] = (run_abs >= 0 && (s_abs - run_abs) <= limit) ? run_val : canonical_nan<T>();
    }
    const int e0 = ((w * U + u) * kWave + l) * VEC;
    if (full && VEC > 1) {
      VecT<T> t;
#pragma unroll
      for (int v = 0; v < VEC; ++v) t.v[v] = res[v];
      store_stream16(orow + p0 + e0, t.v);  // whole-wave contiguous runs
    } else {
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        const int64_t p = p0 + e0 + v;
        if (p < n) orow[p] = res[v];
      }
    }
  }
}

// ---- second pass of the two-pass fill -------------------------------------------
// One workgroup per row walks the row's tile summaries in scan order, 256 at
// a time (one per thread): each tile's carry (the last valid value before it)
// comes from a ballot over the lanes holding one, the earlier waves' last
// through LDS, and the running carry of the earlier chunks. Each wave then
// fills the leading NaN runs of its own tiles that need it (positions
// [tile start, first valid) within `limit` of the carry), one tile at a time
// with whole-wave stores.
template <typename T, bool REV>
__global__ __launch_bounds__(kFillThreads) void fill_fix_kernel(T* __restrict__ out, int64_t n, int64_t ld_out,
                                                                int64_t limit, int64_t tiles, int64_t tile_len,
                                                                const FillSummary* __restrict__ sums) {
  __shared__ long long s_wso[kFillWaves];
  __shared__ T s_wval[kFillWaves];
  const int64_t row = blockIdx.x;
  const FillSummary* srow = sums + row * tiles;
  T* orow = out + row * ld_out;
  const int w = wave_id();
  const int l = lane_id();
  long long run_so = -1;  // carry from the earlier chunks (block-uniform)
  T run_val = T(0);
  for (int64_t c0 = 0; c0 < tiles; c0 += kFillThreads) {
    const int64_t k = c0 + threadIdx.x;  // tile index in scan order
    long long last = -1, beg = 0, end = 0;
    T val = T(0);
    if (k < tiles) {
      const int64_t tl = REV ? tiles - 1 - k : k;
      const FillSummary sm = srow[tl];
      // scan-order position of the tile's first element (< 0 for the partial
      // last tile of bfill) and end of its span
      const int64_t p0 = tl * tile_len;
      const int64_t base = REV ? n - p0 - tile_len : p0;
      const int64_t stop = REV ? n - p0 : (p0 + tile_len < n ? p0 + tile_len : n);
      last = sm.last >= 0 ? base + sm.last : -1;
      end = sm.first >= 0 ? base + sm.first : stop;
      val = (T)sm.val;
      beg = base > 0 ? base : 0;
    }
    // nearest earlier lane (lower lane: earlier tile) holding a valid value;
    // the lane reads run on every lane, results are selected afterwards
    const unsigned long long m = __ballot(last >= 0);
    const unsigned long long before = m & ((1ull << l) - 1ull);
    const int src = before == 0ull ? -1 : 63 - __clzll((long long)before);
    const int from = src < 0 ? l : src;
    const long long oso = __shfl(last, from, kWave);
    const T oval = __shfl(val, from, kWave);
    const int t = m == 0ull ? 0 : 63 - __clzll((long long)m);
    const long long tso = read_lane(last, t);
    const T tval = read_lane(val, t);
    if (l == 0) {
      s_wso[w] = m == 0ull ? -1 : tso;
      s_wval[w] = tval;
    }
    __syncthreads();
    long long cso = run_so;
    T cval = run_val;
    for (int i = 0; i < w; ++i) {
      if (s_wso[i] >= 0) {
        cso = s_wso[i];
        cval = s_wval[i];
      }
    }
    if (src >= 0) {
      cso = oso;
      cval = oval;
    }
    for (int i = 0; i < kFillWaves; ++i) {  // the chunk's last valid value carries on
      if (s_wso[i] >= 0) {
        run_so = s_wso[i];
        run_val = s_wval[i];
      }
    }
    // leading run positions within `limit` of the carry
    const long long lim_end = cso >= 0 ? (cso + limit + 1 < end ? cso + limit + 1 : end) : beg;
    unsigned long long todo = __ballot(lim_end > beg);
    while (todo != 0ull) {  // wave-uniform
      const int i = __ffsll((long long)todo) - 1;
      todo &= todo - 1ull;
      const long long b = read_lane(beg, i);
      const long long e = read_lane(lim_end, i);
      const T v = read_lane(cval, i);
      for (long long q = b + l; q < e; q += kWave) orow[REV ? n - 1 - q : q] = v;
    }
    __syncthreads();  // before the next chunk rewrites s_wso / s_wval
  }
}

// ---- span kernel -------------------------------------------------------------
// One workgroup walks `span` consecutive tiles of a row (one ticket per span).
// The next tile is prefetched into registers while the current one is scanned;
// the running carry stays block-uniform across the span's tiles, and each tile
// with a valid value publishes its prefix at once. Only the span's leading run
// of NaNs (scan positions before its first valid value) depends on earlier
// spans: those outputs are written after the span, once one look-back has
// found the carry, so a workgroup never waits in the middle of its span.
template <typename T, int VEC, int U, bool REV, bool PACK, bool PF = false>
__global__ __launch_bounds__(kFillThreads) void fill_span_kernel(
    const T* __restrict__ x, T* __restrict__ out, int64_t rows, int64_t n, int64_t ld_x,
    int64_t ld_out, int64_t limit, int64_t tiles_per_row, int64_t span,
    unsigned int* __restrict__ counters, unsigned long long* __restrict__ status,
    unsigned int* __restrict__ timeouts) {
  constexpr int TILE = kFillThreads * U * VEC;
  __shared__ int s_ticket;
  __shared__ int s_wso[kFillWaves];
  __shared__ T s_wval[kFillWaves];
  __shared__ int s_wfirst[kFillWaves];
  __shared__ long long s_carry;
  __shared__ T s_carry_val;

  const int64_t row = (int64_t)(blockIdx.x % (uint64_t)rows);
  const T* xr = x + row * ld_x;
  T* orow = out + row * ld_out;
  unsigned long long* srow = status + row * tiles_per_row;
  const int w = wave_id();
  const int l = lane_id();
  if (threadIdx.x == 0) s_ticket = (int)atomicAdd(&counters[row], 1u);
  __syncthreads();
  const int64_t tk0 = (int64_t)s_ticket * span;
  const int64_t tk1 = tk0 + span < tiles_per_row ? tk0 + span : tiles_per_row;

  auto s_base_of = [&](int64_t tk) -> int64_t {
    const int64_t tile = REV ? (tiles_per_row - 1 - tk) : tk;
    return REV ? (n - tile * TILE - TILE) : tile * TILE;
  };
  auto so_of = [&](int u, int v) -> int {
    const int e = ((w * U + u) * kWave + l) * VEC + v;
    return REV ? (TILE - 1 - e) : e;
  };
  auto load_tile = [&](int64_t tk, T (&vals)[U][VEC]) {
    const int64_t tile = REV ? (tiles_per_row - 1 - tk) : tk;
    const int64_t q0 = tile * TILE;
    const bool q_full = (q0 + TILE <= n);
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int e0 = ((w * U + u) * kWave + l) * VEC;
      if (q_full && VEC > 1) {
        VecT<T> t = load_stream(reinterpret_cast<const VecT<T>*>(xr + q0 + e0));
#pragma unroll
        for (int v = 0; v < VEC; ++v) vals[u][v] = t.v[v];
      } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const int64_t p = q0 + e0 + v;
          vals[u][v] = (p < n) ? xr[p] : canonical_nan<T>();
        }
      }
    }
  };

  // Block-uniform state across the span's tiles.
  bool known = (tk0 == 0);  // carry into the current tile known (none before the row)
  long long run_idx = -1;   // scan-order index of the last valid value so far (-1 none)
  T run_val = T(0);
  long long lead_end = -1;  // scan position of the span's first valid value, once seen

  // PF: the next tile is loaded into registers while this one is scanned.
  T cur[U][VEC];
  if constexpr (PF) load_tile(tk0, cur);
  for (int64_t tk = tk0; tk < tk1; ++tk) {
    T nxt[U][VEC];
    if constexpr (PF) {
      if (tk + 1 < tk1) load_tile(tk + 1, nxt);
    } else {
      load_tile(tk, cur);
    }
    const int64_t s_base = s_base_of(tk);
    const int64_t tile = REV ? (tiles_per_row - 1 - tk) : tk;
    const int64_t p0 = tile * TILE;
    const bool full = (p0 + TILE <= n);

    // In-tile scan: per-slot exclusive carries, wave aggregates, first valid.
    Last<T> cin[U];
    Last<T> wcarry{-1, T(0)};
    int wfirst = 0x7fffffff;
#pragma unroll
    for (int i = 0; i < U; ++i) {
      const int u = REV ? (U - 1 - i) : i;
      Last<T> agg{-1, T(0)};
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        if (!is_nan(cur[u][v])) {
          const int so = so_of(u, v);
          if (so > agg.so) {
            agg.so = so;
            agg.val = cur[u][v];
          }
          wfirst = so < wfirst ? so : wfirst;
        }
      }
      Last<T> excl, tail;
      wave_last_carry<REV>(agg, excl, tail);
      if (excl.so < 0) excl = wcarry;
      cin[u] = excl;
      if (tail.so >= 0) wcarry = tail;
    }
#pragma unroll
    for (int d = kWave / 2; d > 0; d >>= 1) {
      const int o = shfl_down(wfirst, d);
      wfirst = o < wfirst ? o : wfirst;
    }
    if (l == 0) {
      s_wso[w] = wcarry.so;
      s_wval[w] = wcarry.val;
      s_wfirst[w] = wfirst;
    }
    __syncthreads();
    Last<T> wave_in{-1, T(0)};
    Last<T> tile_agg{-1, T(0)};
    int tile_first = 0x7fffffff;
#pragma unroll
    for (int i = 0; i < kFillWaves; ++i) {
      const int ww = REV ? (kFillWaves - 1 - i) : i;
      if (s_wso[ww] >= 0) {
        if ((REV && ww > w) || (!REV && ww < w)) {
          wave_in.so = s_wso[ww];
          wave_in.val = s_wval[ww];
        }
        tile_agg.so = s_wso[ww];
        tile_agg.val = s_wval[ww];
      }
      tile_first = s_wfirst[ww] < tile_first ? s_wfirst[ww] : tile_first;
    }

    // Publish this tile's status (never waits).
    if (threadIdx.x == 0) {
      if (tile_agg.so >= 0) {
        const long long idx = s_base + tile_agg.so;
        if constexpr (PACK) st_agent(&srow[tk], kFlagP | pack_carry(idx, tile_agg.val));
        else st_agent(&srow[tk], kFlagP | (unsigned long long)(idx + 1));
      } else if (known) {
        if constexpr (PACK) st_agent(&srow[tk], kFlagP | (run_idx >= 0 ? pack_carry(run_idx, run_val) : 0ull));
        else st_agent(&srow[tk], kFlagP | (unsigned long long)(run_idx + 1));
      } else {
        st_agent(&srow[tk], kFlagA);
      }
    }

    // Outputs; positions with no valid value before them in the span while
    // the carry is pending are left for the end of the span.
#pragma unroll
    for (int u = 0; u < U; ++u) {
      long long r_abs = known ? run_idx : -2;  // -2: carry not known yet
      T r_val = run_val;
      if (wave_in.so >= 0) {
        r_abs = s_base + wave_in.so;
        r_val = wave_in.val;
      }
      if (cin[u].so >= 0) {
        r_abs = s_base + cin[u].so;
        r_val = cin[u].val;
      }
      T res[VEC];
      bool keep[VEC];
#pragma unroll
      for (int i = 0; i < VEC; ++i) {
        const int v = REV ? (VEC - 1 - i) : i;
        const int64_t s_abs = s_base + so_of(u, v);
        if (!is_nan(cur[u][v])) {
          r_abs = s_abs;
          r_val = cur[u][v];
        }
        keep[v] = r_abs != -2;
        res[v] = (r_abs >= 0 && (s_abs - r_abs) <= limit) ? r_val : canonical_nan<T>();
      }
      const int e0 = ((w * U + u) * kWave + l) * VEC;
      bool all_keep = true;
#pragma unroll
      for (int v = 0; v < VEC; ++v) all_keep = all_keep && keep[v];
      if (full && VEC > 1 && all_keep) {
        VecT<T> t;
#pragma unroll
        for (int v = 0; v < VEC; ++v) t.v[v] = res[v];
        *reinterpret_cast<VecT<T>*>(orow + p0 + e0) = t;
      } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const int64_t p = p0 + e0 + v;
          if (p < n && keep[v]) orow[p] = res[v];
        }
      }
    }

    // Block-uniform updates for the next tile.
    if (tile_agg.so >= 0) {
      if (!known) lead_end = s_base + tile_first;
      run_idx = s_base + tile_agg.so;
      run_val = tile_agg.val;
      known = true;
    }
    if constexpr (PF) {
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int v = 0; v < VEC; ++v) cur[u][v] = nxt[u][v];
    }
    __syncthreads();  // LDS aggregates reused by the next tile
  }

  // The span's leading run [span start, first valid value) in scan order:
  // one look-back for the carry into the span, then its outputs.
  if (tk0 > 0) {
    const int64_t sb0 = s_base_of(tk0);
    const int64_t span_s0 = sb0 > 0 ? sb0 : 0;
    const int64_t sb1 = s_base_of(tk1 - 1) + TILE;
    const int64_t span_s1 = sb1 < n ? sb1 : n;
    const int64_t lead = lead_end >= 0 ? lead_end : span_s1;
    if (lead > span_s0) {
      if (threadIdx.x == 0) {
        unsigned long long found = 0ull;
        for (int64_t j = tk0 - 1;; --j) {
          unsigned long long st;
          unsigned int spins = 0;
          while (((st = ld_agent(&srow[j])) >> 62) == 0) {
            spin_pause();
            if (++spins > kSpinLimit) {
              atomicAdd(timeouts, 1u);
              st = kFlagP;
              break;
            }
          }
          if ((st >> 62) == 2) {
            found = st & kIdxMask;
            break;
          }
          if (j == 0) break;
        }
        long long carry;
        if constexpr (PACK) {
          carry = (long long)(found >> 32) - 1;
          if (carry >= 0) s_carry_val = __uint_as_float((unsigned int)(found & 0xffffffffull));
        } else {
          carry = (long long)found - 1;
          if (carry >= 0) s_carry_val = xr[REV ? (n - 1 - carry) : carry];
        }
        s_carry = carry;
        // a span without a valid value now knows its prefix
        if (!known) st_agent(&srow[tk1 - 1], kFlagP | found);
      }
      __syncthreads();
      const long long carry = s_carry;
      const T cval = carry >= 0 ? s_carry_val : T(0);
      for (int64_t s = span_s0 + threadIdx.x; s < lead; s += kFillThreads) {
        const int64_t p = REV ? (n - 1 - s) : s;
        orow[p] = (carry >= 0 && (s - carry) <= limit) ? cval : canonical_nan<T>();
      }
    }
  }
}

// Input bytes per thread of a tile: 64 by default (16 float32 / 8 float64
// elements); NBG_FILL_BYTES=16, 32 or 128 selects other tiles (1, 2 or 8
// 16-byte vectors per thread: more or fewer workgroups and look-back tiles).
// The workspace query reads the same setting.
int fill_bytes_per_thread() {
  const long long v = knob("NBG_FILL_BYTES", 64);
  return (v == 16 || v == 32 || v == 128) ? (int)v : 64;
}

template <typename T, int BYTES>
int fill_impl_b(bool rev, const void* xv, void* ov, int64_t rows, int64_t n, int64_t ld_x,
                int64_t ld_out, int64_t limit, void* ws, size_t ws_bytes, hipStream_t s) {
  const T* x = static_cast<const T*>(xv);
  T* out = static_cast<T*>(ov);
  constexpr int TILE = kFillThreads * BYTES / (int)sizeof(T);
  const int64_t tiles = (n + TILE - 1) / TILE;
  const int64_t blocks = rows * tiles;
  if (blocks > 0x7fffffffLL) return set_error(NBG_ERR_INVALID, "fill: too many tiles (%lld)", (long long)blocks);
  const bool lookback = tiles > 1;
  unsigned int* counters = nullptr;
  unsigned long long* status = nullptr;
  if (lookback) {
    size_t need = nbg_fill_workspace_bytes(rows, n, sizeof(T) == 4 ? NBG_F32 : NBG_F64);
    if (ws == nullptr || ws_bytes < need)
      return set_error(NBG_ERR_INVALID, "fill: workspace too small (%zu < %zu)", ws_bytes, need);
    status = static_cast<unsigned long long*>(ws);
    counters = reinterpret_cast<unsigned int*>(status + rows * tiles);
    int rc = launch_zero(ws, need, s);
    if (rc) return rc;
  }
  unsigned int* to = spin_counter_ptr();
  if (to == nullptr) return set_error(NBG_ERR_HIP, "fill: cannot resolve device counter");
  const int vx = pick_vec<T>(x, ld_x, rows), vo = pick_vec<T>(out, ld_out, rows);
  const int vec = vx < vo ? vx : vo;
  constexpr int MAXV = 16 / (int)sizeof(T);
  constexpr int UV = BYTES / (int)sizeof(T) / MAXV;  // slots per thread at full vector width
  dim3 grid((unsigned)blocks), block(kFillThreads);
  // PACK: the status word carries the float32 carry value (rows < 2**30;
  // NBG_FILL_PACK=0 turns it off, so tests reach the other path on short rows).
  constexpr bool PACK_OK = sizeof(T) == 4;
  const bool pack = PACK_OK && n <= kPackMaxN && knob("NBG_FILL_PACK", 1) != 0;
  // Tiles per dispatch run of one row: NBG_FILL_RUN (default 1, row-interleaved).
  int64_t run = knob("NBG_FILL_RUN", 1);
  run = run < 1 ? 1 : (run > tiles ? tiles : run);
#define NBG_FILL_LAUNCH_P(V, U, R, LB, P)                                                          \
  hipLaunchKernelGGL((fill_kernel<T, V, U, R, LB, P>), grid, block, 0, s, x, out, rows, n, ld_x, \
                     ld_out, limit, tiles, counters, status, to, run)
  // Spans: NBG_FILL_SPAN=<k> tiles per workgroup (1 = one tile each).
  int64_t span = knob("NBG_FILL_SPAN", 1);
  span = span < 1 ? 1 : (span > 16 ? 16 : span);
  if (span > tiles) span = tiles;
  const int64_t spans = (tiles + span - 1) / span;
  dim3 sgrid((unsigned)(rows * spans));
  const bool fill_pf = knob("NBG_FILL_PREFETCH", 0) == 1;
#define NBG_FILL_SPAN_P(V, U, R, P)                                                                       \
  do {                                                                                                    \
    if (fill_pf)                                                                                          \
      hipLaunchKernelGGL((fill_span_kernel<T, V, U, R, P, true>), sgrid, block, 0, s, x, out, rows, n,   \
                         ld_x, ld_out, limit, tiles, span, counters, status, to);                         \
    else                                                                                                  \
      hipLaunchKernelGGL((fill_span_kernel<T, V, U, R, P, false>), sgrid, block, 0, s, x, out, rows, n,  \
                         ld_x, ld_out, limit, tiles, span, counters, status, to);                         \
  } while (0)
#define NBG_FILL_LAUNCH(V, U)                         \
  do {                                                \
    if (!lookback) {                                  \
      if (rev) NBG_FILL_LAUNCH_P(V, U, true, false, false); \
      else NBG_FILL_LAUNCH_P(V, U, false, false, false);    \
    } else if (span > 1) {                            \
      if (pack) {                                     \
        if (rev) NBG_FILL_SPAN_P(V, U, true, PACK_OK); \
        else NBG_FILL_SPAN_P(V, U, false, PACK_OK);    \
      } else {                                        \
        if (rev) NBG_FILL_SPAN_P(V, U, true, false);  \
        else NBG_FILL_SPAN_P(V, U, false, false);     \
      }                                               \
    } else if (pack) {                                \
      if (rev) NBG_FILL_LAUNCH_P(V, U, true, true, PACK_OK); \
      else NBG_FILL_LAUNCH_P(V, U, false, true, PACK_OK);    \
    } else {                                          \
      if (rev) NBG_FILL_LAUNCH_P(V, U, true, true, false);  \
      else NBG_FILL_LAUNCH_P(V, U, false, true, false);     \
    }                                                 \
  } while (0)
  if (vec == MAXV)
    NBG_FILL_LAUNCH(MAXV, UV);
  else
    NBG_FILL_LAUNCH(1, UV * MAXV);
#undef NBG_FILL_LAUNCH
#undef NBG_FILL_LAUNCH_P
#undef NBG_FILL_SPAN_P
  return check_hip(hipGetLastError(), "fill_kernel launch");
}

// Two-pass fill for inputs with many rows: independent tiles of 32 B per
// thread in row-major order (no tickets, no waiting: the pattern of a plain
// copy), then one fix-up workgroup per row. NBG_FILL_TWO_PASS=0 / 1 forces
// the choice; the workspace query asks here too. (Measured on MI355X for the
// first pass over float32 (1024, 1M): 16 B per thread 1.67 ms, 32 B per thread
// 1.33 ms — with 4 KB tiles the per-workgroup cost shows.) Two passes from 8
// rows up: float32 ffill of 4.096e9 elements as rows x n = 1024 x 1M / 256 x
// 4M / 64 x 16M / 16 x 64M / 4 x 256M took 1.42 / 1.36 / 1.44 / 1.73 / 2.99
// ms in two passes against 1.64 / 1.65 / 1.80 / 2.97 / 3.11 ms in one (the
// fix-up's per-row scan of tile summaries grows with fewer, longer rows).
constexpr int64_t kFillTwoPassRows = 8;
constexpr int kFillPass1Vecs = 2;  // 16-byte vectors per thread and tile
inline int64_t fill_pass1_tile(int elem) { return (int64_t)kFillThreads * 16 * kFillPass1Vecs / elem; }
bool fill_two_pass(int64_t rows, int64_t n, int elem) {
  const int64_t t = fill_pass1_tile(elem);
  if (rows <= 0 || (n + t - 1) / t <= 1) return false;
  const long long v = knob("NBG_FILL_TWO_PASS", -1);
  if (v == 0) return false;
  if (v == 1) return true;
  return rows >= kFillTwoPassRows;
}

template <typename T>
int fill_two_pass_impl(bool rev, const void* xv, void* ov, int64_t rows, int64_t n, int64_t ld_x,
                       int64_t ld_out, int64_t limit, void* ws, size_t ws_bytes, hipStream_t s) {
  const T* x = static_cast<const T*>(xv);
  T* out = static_cast<T*>(ov);
  constexpr int MAXV = 16 / (int)sizeof(T);
  constexpr int TILE = kFillThreads * MAXV * kFillPass1Vecs;
  static_assert(TILE == kFillThreads * 16 * kFillPass1Vecs / (int)sizeof(T), "tile as in fill_pass1_tile");
  const int64_t tiles = (n + TILE - 1) / TILE;
  const int64_t blocks = rows * tiles;
  if (blocks > 0x7fffffffLL) return set_error(NBG_ERR_INVALID, "fill: too many tiles (%lld)", (long long)blocks);
  const size_t need = (size_t)blocks * sizeof(FillSummary);
  if (ws == nullptr || ws_bytes < need)
    return set_error(NBG_ERR_INVALID, "fill: workspace too small (%zu < %zu)", ws_bytes, need);
  FillSummary* sums = static_cast<FillSummary*>(ws);
  unsigned int* to = spin_counter_ptr();
  if (to == nullptr) return set_error(NBG_ERR_HIP, "fill: cannot resolve device counter");
  const int vx = pick_vec<T>(x, ld_x, rows), vo = pick_vec<T>(out, ld_out, rows);
  const int vec = vx < vo ? vx : vo;
  dim3 grid((unsigned)blocks), block(kFillThreads);
  // run = tiles: row-major tile order, consecutive workgroups on consecutive memory
#define NBG_FILL_PASS1(V, U, R)                                                                          \
  hipLaunchKernelGGL((fill_kernel<T, V, U, R, false, false, true>), grid, block, 0, s, x, out, rows, n, \
                     ld_x, ld_out, limit, tiles, nullptr, nullptr, to, tiles, sums)
  if (vec == MAXV) {
    if (rev) NBG_FILL_PASS1(MAXV, kFillPass1Vecs, true);
    else NBG_FILL_PASS1(MAXV, kFillPass1Vecs, false);
  } else {
    if (rev) NBG_FILL_PASS1(1, MAXV * kFillPass1Vecs, true);
    else NBG_FILL_PASS1(1, MAXV * kFillPass1Vecs, false);
  }
#undef NBG_FILL_PASS1
  if (int rc = check_hip(hipGetLastError(), "fill_kernel (independent pass) launch")) return rc;
  if (rev)
    hipLaunchKernelGGL((fill_fix_kernel<T, true>), dim3((unsigned)rows), block, 0, s, out, n, ld_out, limit, tiles,
                       (int64_t)TILE, sums);
  else
    hipLaunchKernelGGL((fill_fix_kernel<T, false>), dim3((unsigned)rows), block, 0, s, out, n, ld_out, limit, tiles,
                       (int64_t)TILE, sums);
  return check_hip(hipGetLastError(), "fill_fix_kernel launch");
}

template <typename T>
int fill_impl(bool rev, const void* x, void* out, int64_t rows, int64_t n, int64_t ld_x,
              int64_t ld_out, int64_t limit, void* ws, size_t ws_bytes, hipStream_t s) {
  if (fill_two_pass(rows, n, (int)sizeof(T)))
    return fill_two_pass_impl<T>(rev, x, out, rows, n, ld_x, ld_out, limit, ws, ws_bytes, s);
  switch (fill_bytes_per_thread()) {
    case 16: return fill_impl_b<T, 16>(rev, x, out, rows, n, ld_x, ld_out, limit, ws, ws_bytes, s);
    case 32: return fill_impl_b<T, 32>(rev, x, out, rows, n, ld_x, ld_out, limit, ws, ws_bytes, s);
    case 128: return fill_impl_b<T, 128>(rev, x, out, rows, n, ld_x, ld_out, limit, ws, ws_bytes, s);
    default: return fill_impl_b<T, 64>(rev, x, out, rows, n, ld_x, ld_out, limit, ws, ws_bytes, s);
  }
}

int fill_entry(const char* name, bool rev, int dtype, const void* x, void* out, int64_t rows,
               int64_t n, int64_t ld_x, int64_t ld_out, int64_t limit, void* ws, size_t ws_bytes,
               void* stream) {
  if (int rc = injected_error(name)) return rc;
  if (rows < 0 || n < 0) return set_error(NBG_ERR_INVALID, "%s: negative shape", name);
  if (limit < 0) return set_error(NBG_ERR_INVALID, "%s: `limit` must be positive: %lld", name, (long long)limit);
  if (rows == 0 || n == 0) return NBG_OK;
  if (x == nullptr || out == nullptr) return set_error(NBG_ERR_INVALID, "%s: null pointer", name);
  if (rows > 1 && (ld_x < n || ld_out < n)) return set_error(NBG_ERR_INVALID, "%s: row stride < n", name);
  // A gap within a row is at most n - 1 positions, so every limit >= n fills
  // like limit = n; clamping keeps `carry + limit + 1` from overflowing for
  // limits near INT64_MAX (the reference accepts any non-negative limit).
  if (limit > n) limit = n;
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case NBG_F32: return fill_impl<float>(rev, x, out, rows, n, ld_x, ld_out, limit, ws, ws_bytes, s);
    case NBG_F64: return fill_impl<double>(rev, x, out, rows, n, ld_x, ld_out, limit, ws, ws_bytes, s);
    default: return set_error(NBG_ERR_INVALID, "%s: unsupported dtype code %d", name, dtype);
  }
}

}  // namespace
}  // namespace nbg

extern "C" {

size_t nbg_fill_workspace_bytes(int64_t rows, int64_t n, int dtype) {
  const int elem = dtype == NBG_F64 ? 8 : 4;
  if (nbg::fill_two_pass(rows, n, elem)) {
    const int64_t t = nbg::fill_pass1_tile(elem);
    return (size_t)(rows * ((n + t - 1) / t)) * sizeof(nbg::FillSummary);
  }
  const int64_t tile = (int64_t)nbg::kFillThreads * nbg::fill_bytes_per_thread() / elem;
  const int64_t tiles = (n + tile - 1) / tile;
  if (rows <= 0 || tiles <= 1) return 0;
  return (size_t)(rows * tiles) * 8 + (size_t)((rows * 4 + 7) / 8) * 8;
}

int nbg_ffill(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld_x,
              int64_t ld_out, int64_t limit, void* ws, size_t ws_bytes, void* stream) {
  return nbg::fill_entry("nbg_ffill", false, dtype, x, out, rows, n, ld_x, ld_out, limit, ws,
                         ws_bytes, stream);
}

int nbg_bfill(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld_x,
              int64_t ld_out, int64_t limit, void* ws, size_t ws_bytes, void* stream) {
  return nbg::fill_entry("nbg_bfill", true, dtype, x, out, rows, n, ld_x, ld_out, limit, ws,
                         ws_bytes, stream);
}

}  // extern "C"
