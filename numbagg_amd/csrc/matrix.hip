// Correlation / covariance matrices for gfx950 (numbagg/funcs.py:338-532,
// numbagg/moving_matrix.py:16-432).
//
// Every statistic is pairwise complete: an observation counts for the pair
// (i, j) only where both variables are valid. All sums are float64 over values
// shifted by a per-variable constant (the first valid value of the variable in
// the span a kernel covers), so a covariance is formed from sums of O(spread)
// terms rather than by cancelling sums of O(level^2) terms; every statistic is
// shift invariant, and the shift is exact to apply for float32 data.
//
// Static matrices (x: (batch, vars, obs) -> (batch, vars, vars)) are six
// pairwise contractions over the observations, with y = valid ? x - shift : 0
// and m = valid ? 1 : 0 per element:
//   P = Y Y^T, Si = Y M^T, Sj = M Y^T, N = M M^T, Qi = Y^2 M^T, Qj = M (Y^2)^T
// (covariance needs the first four). They run on the float64 matrix cores
// (V_MFMA_F64_16X16X4_F64): a wave owns a 16 TW x 16 TW tile of variable pairs
// over a range of observations; each lane loads four consecutive observations
// of one row of the i block and of the j block, and the four MFMA steps of a
// 16-observation group take element e of every lane's quadruple, so that the
// k slot (lane / 16) of step e is observation 4 (lane / 16) + e for both
// operands. Only tiles on or above the diagonal are computed. Observation
// ranges are split over waves when there are few tiles; splits write float64
// partial sums that a second kernel adds in split order (deterministic), and a
// single split finalises in the main kernel.
//
// Moving matrices (x: (batch, obs, vars) -> (batch, obs, vars, vars)) assign a
// lane to one ordered pair (i, j): a wave covers 64 consecutive pairs of the
// output matrix (or, for fewer than 64 pairs, several consecutive time ranges
// of all pairs), so every store writes consecutive output elements. Each lane
// walks its time range with running float64 window sums (add the new
// observation, remove the one leaving the window), after warming up over the
// window - 1 observations before the range; shifts are taken per time tile.
//
// Exponentially weighted matrices carry the same per-pair state as
// move_exp_nancov/nancorr (numbagg/moving_exp.py:227-335) across the whole
// series: the series is cut into chunks, a first kernel forms each chunk's
// state from zero, a second carries the chunk states along the series (state
// at a chunk's start = decay product of the previous chunk x state at its
// start + its local state), and a third replays each chunk from its carried
// state and writes the outputs. Shifts are per variable over the whole series.
#include <initializer_list>
#include <type_traits>

#include "common.hpp"

namespace nbg {
namespace {

typedef double mx_d4 __attribute__((ext_vector_type(4)));

constexpr int kMxThreads = 256;
constexpr int kMxWaves = kMxThreads / kWave;

template <typename T>
__device__ __forceinline__ T mx_nan() {
  return canonical_nan<T>();
}

// This wave's index in its workgroup, as a scalar: work-item indices derived
// from it (and loop bounds derived from those) are then known to be uniform
// across the wave, so loops branch on scalars and the MFMA accumulators stay
// in their registers between iterations.
__device__ __forceinline__ int mx_wave_id() { return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)); }

// 1/sqrt(v) from the hardware estimate and one Newton step (about 1 ULP of
// float64), for v within [2^-1000, 2^1000].
__device__ __forceinline__ double mx_rsqrt(double v) {
  const double y = __builtin_amdgcn_rsq(v);
  return y * fma(-(0.5 * v) * y, y, 1.5);
}

// cov / sqrt(vi vj) for vi, vj > 0.
__device__ __forceinline__ double mx_corr(double cov, double vv) {
  return (vv >= 0x1p-1000 && vv <= 0x1p1000) ? cov * mx_rsqrt(vv) : cov / sqrt(vv);
}

// ---------------------------------------------------------------------------
// Shifts: the first valid value of each (outer, variable) series within
// [lo, hi) of its time axis (0 when none is valid). One wave per series.
// ---------------------------------------------------------------------------
struct MxShiftArgs {
  int64_t outer;      // series groups: batch x tiles
  int64_t nv;         // variables
  int64_t tiles;      // time tiles per batch row (1: whole series)
  int64_t tile_len;   // steps per tile
  int64_t halo;       // steps before a tile that its windows reach
  int64_t len;        // time axis length
  int64_t ld_b;       // batch stride (elements)
  int64_t vstride;    // variable stride
  int64_t tstride;    // time stride
};

template <typename T>
__global__ __launch_bounds__(kMxThreads) void mx_shift_kernel(const T* __restrict__ x, double* __restrict__ shift,
                                                             MxShiftArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t s = (int64_t)blockIdx.x * kMxWaves + mx_wave_id();
  if (s >= a.outer * a.nv) return;  // whole wave
  const int64_t o = s / a.nv, v = s % a.nv;
  const int64_t b = o / a.tiles, tile = o % a.tiles;
  int64_t lo = tile * a.tile_len - a.halo;
  if (lo < 0) lo = 0;
  int64_t hi = tile * a.tile_len + a.tile_len;
  if (hi > a.len) hi = a.len;
  const T* p = x + b * a.ld_b + v * a.vstride;
  double found = 0.0;
  for (int64_t t0 = lo; t0 < hi; t0 += kWave) {
    const int64_t t = t0 + lane;
    const T val = t < hi ? p[t * a.tstride] : mx_nan<T>();
    const unsigned long long m = __ballot(val == val);
    if (m) {
      const int first = __ffsll((long long)m) - 1;
      found = __shfl((double)val, first);
      break;
    }
  }
  if (lane == 0) shift[s] = found;
}

// ---------------------------------------------------------------------------
// Static matrices on the float64 matrix cores.
// ---------------------------------------------------------------------------
struct MxStaticArgs {
  const void* x;
  void* out;
  const double* shift;  // (batch, nv)
  double* part;         // partial sums when splits > 1
  int64_t batch, nv, nobs;
  int64_t ntile;        // tiles per side
  int64_t tiles;        // tiles on or above the diagonal
  int64_t splits;       // observation ranges per tile
  int64_t span;         // observations per split (a multiple of 16)
};

template <bool CORR>
struct StaticSpec {
  static constexpr int NPROD = CORR ? 6 : 4;  // P, Si, Sj, N[, Qi, Qj]
};

// The output of one pair from its sums (i == j: the variance itself, so the
// diagonal of a correlation matrix is exactly 1).
template <bool CORR>
__device__ __forceinline__ double static_value(double P, double Si, double Sj, double N, double Qi, double Qj,
                                               bool diag) {
  const double nan = __builtin_nan("");
  if (!(N > 1.0)) return nan;
  const double cv = P - Si * Sj / N;
  if (!CORR) return cv / (N - 1.0);
  const double vi = Qi - Si * Si / N, vj = Qj - Sj * Sj / N;
  if (!(vi > 0.0 && vj > 0.0)) return nan;
  return (diag ? vi : cv) / sqrt(vi * vj);
}

__device__ __forceinline__ void tile_of(int64_t t, int64_t ntile, int64_t& I, int64_t& J) {
  int64_t i = 0;
  while (t >= ntile - i) {
    t -= ntile - i;
    ++i;
  }
  I = i;
  J = i + t;
}

template <typename T, bool VEC>
__device__ __forceinline__ void load4(const T* __restrict__ row, int64_t k, int64_t k_hi, bool ok, T (&v)[4]) {
  if (VEC && ok && k + 4 <= k_hi) {
    if constexpr (sizeof(T) == 4) {
      const float4 f = *reinterpret_cast<const float4*>(row + k);
      v[0] = f.x;
      v[1] = f.y;
      v[2] = f.z;
      v[3] = f.w;
    } else {
      const double2 f0 = *reinterpret_cast<const double2*>(row + k);
      const double2 f1 = *reinterpret_cast<const double2*>(row + k + 2);
      v[0] = f0.x;
      v[1] = f0.y;
      v[2] = f1.x;
      v[3] = f1.y;
    }
  } else {
#pragma unroll
    for (int e = 0; e < 4; ++e) v[e] = (ok && k + e < k_hi) ? row[k + e] : mx_nan<T>();
  }
}

template <typename T, bool CORR, int TW, bool VEC>
__global__ __launch_bounds__(kMxThreads) void mx_static_kernel(MxStaticArgs a) {
  constexpr int NP = StaticSpec<CORR>::NPROD;
  constexpr int TS = 16 * TW;
  const int64_t item = (int64_t)blockIdx.x * kMxWaves + mx_wave_id();
  if (item >= a.batch * a.tiles * a.splits) return;  // whole wave
  const int lane = threadIdx.x & (kWave - 1);
  const int r = lane & 15, q = lane >> 4;
  const int64_t split = item % a.splits;
  const int64_t tile = (item / a.splits) % a.tiles;
  const int64_t b = item / (a.splits * a.tiles);
  int64_t I, J;
  tile_of(tile, a.ntile, I, J);
  const T* x = static_cast<const T*>(a.x) + b * a.nv * a.nobs;
  const T* irow[TW];
  const T* jrow[TW];
  bool iok[TW], jok[TW];
  double ci[TW], cj[TW];
#pragma unroll
  for (int t = 0; t < TW; ++t) {
    const int64_t i = I * TS + t * 16 + r, j = J * TS + t * 16 + r;
    iok[t] = i < a.nv;
    jok[t] = j < a.nv;
    irow[t] = x + (iok[t] ? i : 0) * a.nobs;
    jrow[t] = x + (jok[t] ? j : 0) * a.nobs;
    ci[t] = iok[t] ? a.shift[b * a.nv + i] : 0.0;
    cj[t] = jok[t] ? a.shift[b * a.nv + j] : 0.0;
  }
  mx_d4 acc[NP][TW][TW];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int u = 0; u < TW; ++u)
#pragma unroll
      for (int w = 0; w < TW; ++w) acc[p][u][w] = mx_d4{0.0, 0.0, 0.0, 0.0};

  const int64_t k_lo = split * a.span;
  const int64_t k_hi = k_lo + a.span < a.nobs ? k_lo + a.span : a.nobs;
  // The next group's loads are issued before this group's MFMAs.
  T vi[TW][4], vj[TW][4];
#pragma unroll
  for (int t = 0; t < TW; ++t) {
    load4<T, VEC>(irow[t], k_lo + 4 * q, k_hi, iok[t], vi[t]);
    load4<T, VEC>(jrow[t], k_lo + 4 * q, k_hi, jok[t], vj[t]);
  }
  for (int64_t k0 = k_lo; k0 < k_hi; k0 += 16) {
    T ni[TW][4], nj[TW][4];
    const bool more = k0 + 16 < k_hi;
    if (more) {
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        load4<T, VEC>(irow[t], k0 + 16 + 4 * q, k_hi, iok[t], ni[t]);
        load4<T, VEC>(jrow[t], k0 + 16 + 4 * q, k_hi, jok[t], nj[t]);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      double yi[TW], mi[TW], yj[TW], mj[TW], qi[TW], qj[TW];
#pragma unroll
      for (int t = 0; t < TW; ++t) {
        const bool vai = vi[t][e] == vi[t][e], vaj = vj[t][e] == vj[t][e];
        yi[t] = vai ? (double)vi[t][e] - ci[t] : 0.0;
        yj[t] = vaj ? (double)vj[t][e] - cj[t] : 0.0;
        mi[t] = vai ? 1.0 : 0.0;
        mj[t] = vaj ? 1.0 : 0.0;
        qi[t] = yi[t] * yi[t];
        qj[t] = yj[t] * yj[t];
      }
#pragma unroll
      for (int u = 0; u < TW; ++u)
#pragma unroll
        for (int w = 0; w < TW; ++w) {
          acc[0][u][w] = __builtin_amdgcn_mfma_f64_16x16x4f64(yi[u], yj[w], acc[0][u][w], 0, 0, 0);
          acc[1][u][w] = __builtin_amdgcn_mfma_f64_16x16x4f64(yi[u], mj[w], acc[1][u][w], 0, 0, 0);
          acc[2][u][w] = __builtin_amdgcn_mfma_f64_16x16x4f64(mi[u], yj[w], acc[2][u][w], 0, 0, 0);
          acc[3][u][w] = __builtin_amdgcn_mfma_f64_16x16x4f64(mi[u], mj[w], acc[3][u][w], 0, 0, 0);
          if constexpr (CORR) {
            acc[4][u][w] = __builtin_amdgcn_mfma_f64_16x16x4f64(qi[u], mj[w], acc[4][u][w], 0, 0, 0);
            acc[5][u][w] = __builtin_amdgcn_mfma_f64_16x16x4f64(mi[u], qj[w], acc[5][u][w], 0, 0, 0);
          }
        }
    }
    if (more) {
#pragma unroll
      for (int t = 0; t < TW; ++t)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          vi[t][e] = ni[t][e];
          vj[t][e] = nj[t][e];
        }
    }
  }

  // Lane l holds rows 4 c + l / 16 (c = 0..3), column l % 16 of each 16 x 16
  // block (the float64 MFMA result layout).
  if (a.splits == 1) {
    T* out = static_cast<T*>(a.out) + b * a.nv * a.nv;
#pragma unroll
    for (int u = 0; u < TW; ++u)
#pragma unroll
      for (int w = 0; w < TW; ++w)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int64_t i = I * TS + u * 16 + 4 * c + q, j = J * TS + w * 16 + r;
          if (i >= a.nv || j >= a.nv || (I == J && i > j)) continue;
          const double Qi = CORR ? acc[4 % NP][u][w][c] : 0.0;
          const double Qj = CORR ? acc[5 % NP][u][w][c] : 0.0;
          const T v = (T)static_value<CORR>(acc[0][u][w][c], acc[1][u][w][c], acc[2][u][w][c], acc[3][u][w][c],
                                            Qi, Qj, i == j);
          out[i * a.nv + j] = v;
          out[j * a.nv + i] = v;
        }
  } else {
    double* part = a.part + item * (int64_t)(NP * TS * TS);
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int u = 0; u < TW; ++u)
#pragma unroll
        for (int w = 0; w < TW; ++w)
#pragma unroll
          for (int c = 0; c < 4; ++c) part[(p * TS + u * 16 + 4 * c + q) * TS + w * 16 + r] = acc[p][u][w][c];
  }
}

// Adds the splits' partial sums. One workgroup per (batch row, tile, 16 tile
// positions): 16 lanes per position sum interleaved splits, then lane 0 adds
// the 16 sums in a fixed order (deterministic), forms the output and writes it
// to both triangles.
constexpr int kFinishPos = 16;

template <typename T, bool CORR, int TW>
__global__ __launch_bounds__(kMxThreads) void mx_static_finish_kernel(MxStaticArgs a) {
  constexpr int NP = StaticSpec<CORR>::NPROD;
  constexpr int TS = 16 * TW;
  constexpr int POS = TS * TS;
  constexpr int SL = kMxThreads / kFinishPos;  // split lanes per position
  __shared__ double red[NP][SL][kFinishPos + 1];
  const int64_t per_tile = POS / kFinishPos;
  const int64_t bt = blockIdx.x / per_tile;  // batch row x tiles + tile
  const int pl = threadIdx.x % kFinishPos, sl = threadIdx.x / kFinishPos;
  const int pos = (int)(blockIdx.x % per_tile) * kFinishPos + pl;
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = 0.0;
  const double* part = a.part + bt * a.splits * (int64_t)(NP * POS) + pos;
  for (int64_t sp = sl; sp < a.splits; sp += SL) {
#pragma unroll
    for (int p = 0; p < NP; ++p) acc[p] += part[(sp * NP + p) * POS];
  }
#pragma unroll
  for (int p = 0; p < NP; ++p) red[p][sl][pl] = acc[p];
  __syncthreads();
  if (sl != 0) return;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    double t = 0.0;
    for (int k = 0; k < SL; ++k) t += red[p][k][pl];
    acc[p] = t;
  }
  const int64_t b = bt / a.tiles, tile = bt % a.tiles;
  int64_t I, J;
  tile_of(tile, a.ntile, I, J);
  const int64_t i = I * TS + pos / TS, j = J * TS + pos % TS;
  if (i >= a.nv || j >= a.nv || (I == J && i > j)) return;
  const double Qi = CORR ? acc[4 % NP] : 0.0, Qj = CORR ? acc[5 % NP] : 0.0;
  const T v = (T)static_value<CORR>(acc[0], acc[1], acc[2], acc[3], Qi, Qj, i == j);
  T* out = static_cast<T*>(a.out) + b * a.nv * a.nv;
  out[i * a.nv + j] = v;
  out[j * a.nv + i] = v;
}

// ---------------------------------------------------------------------------
// Static matrices of a few variables (V <= kSmallV) over a large batch: one
// workgroup per batch row on the vector units (a 16 x 16 matrix-core tile
// would leave at least three quarters of its products unused). Thread t
// takes pair t % PP (i <= j, row-major; PP = pairs rounded up to a power of
// two) and every K-th observation of each staged chunk (K = 256 / PP, from
// t / PP); the chunk's shifted values sit in LDS as float64 (NaN where
// missing), a row per variable with one spare slot so that the rows fall in
// different banks. The K partial sums of a pair are added in a fixed order
// (deterministic) and the output formed as static_value does. The LDS (a
// chunk's rows, then the partial sums in the same space) stays near 12 KB, so
// many workgroups share a CU: each one's work is short and latency-bound.
// ---------------------------------------------------------------------------
constexpr int kSmallV = 8;
constexpr int kSmallChunk = kMxThreads;  // observations staged per round (one per thread)
constexpr int64_t kSmallMinBatch = 512;

template <typename T, bool CORR>
__global__ __launch_bounds__(kMxThreads) void mx_small_kernel(MxStaticArgs a) {
  constexpr int NP = StaticSpec<CORR>::NPROD;
  constexpr int RS = kSmallChunk + 1;  // LDS row stride (doubles)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* ys = reinterpret_cast<double*>(smem_raw);   // nv rows of RS
  double* red = reinterpret_cast<double*>(smem_raw);  // NP x kMxThreads, after the last chunk
  const int64_t b = blockIdx.x;
  const int nv = (int)a.nv;
  const int npairs = nv * (nv + 1) / 2;
  int pp = 1;
  while (pp < npairs) pp <<= 1;
  const int K = kMxThreads / pp;
  const int tid = threadIdx.x;
  const int pair = tid % pp, k = tid / pp;
  // pair -> (i, j), i <= j, row-major over the upper triangle
  int i = 0, rem = pair < npairs ? pair : 0;
  while (rem >= nv - i) {
    rem -= nv - i;
    ++i;
  }
  const int j = i + rem;
  const T* x = static_cast<const T*>(a.x) + b * a.nv * a.nobs;
  const double* shift = a.shift + b * a.nv;
  double acc[NP];
#pragma unroll
  for (int q = 0; q < NP; ++q) acc[q] = 0.0;
  // Thread t stages observation t of every variable of a chunk; the next
  // chunk's values are loaded into registers while the current one is summed.
  T pre[kSmallV];
  auto load_chunk = [&](int64_t o0) {
    const bool in = o0 + tid < a.nobs;
#pragma unroll
    for (int v = 0; v < kSmallV; ++v)
      if (v < nv) pre[v] = in ? x[(int64_t)v * a.nobs + o0 + tid] : mx_nan<T>();
  };
  load_chunk(0);
  for (int64_t o0 = 0; o0 < a.nobs; o0 += kSmallChunk) {
    const int cnt = (int)(a.nobs - o0 < kSmallChunk ? a.nobs - o0 : kSmallChunk);
    __syncthreads();  // the previous chunk's reads are done
#pragma unroll
    for (int v = 0; v < kSmallV; ++v)
      if (v < nv) ys[v * RS + tid] = pre[v] == pre[v] ? (double)pre[v] - shift[v] : __builtin_nan("");
    if (o0 + kSmallChunk < a.nobs) load_chunk(o0 + kSmallChunk);
    __syncthreads();
    if (pair < npairs) {
      const double* yi = ys + i * RS;
      const double* yj = ys + j * RS;
      for (int o = k; o < cnt; o += K) {
        const double a_ = yi[o], b_ = yj[o];
        const bool both = a_ == a_ && b_ == b_;
        const double u = both ? a_ : 0.0, w = both ? b_ : 0.0;
        acc[0] += u * w;
        acc[1] += u;
        acc[2] += w;
        acc[3] += both ? 1.0 : 0.0;
        if constexpr (CORR) {
          acc[4 % NP] += u * u;
          acc[5 % NP] += w * w;
        }
      }
    }
  }
  __syncthreads();  // the last chunk's reads are done: its space takes the partial sums
#pragma unroll
  for (int q = 0; q < NP; ++q) red[q * kMxThreads + tid] = acc[q];
  __syncthreads();
  if (tid < npairs) {
    double t[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      t[q] = 0.0;
      for (int kk = 0; kk < K; ++kk) t[q] += red[q * kMxThreads + kk * pp + tid];
    }
    const double Qi = CORR ? t[4 % NP] : 0.0, Qj = CORR ? t[5 % NP] : 0.0;
    const T v = (T)static_value<CORR>(t[0], t[1], t[2], t[3], Qi, Qj, i == j);
    T* out = static_cast<T*>(a.out) + b * a.nv * a.nv;
    out[i * a.nv + j] = v;
    out[j * a.nv + i] = v;
  }
}

// Launch plan: wave tile, observation splits, workspace layout.
struct StaticPlan {
  int tw;
  int64_t ntile, tiles, splits, span;
  size_t shift_bytes, part_bytes;
};

// Waves the static kernel aims for (several per SIMD of 256 CUs).
constexpr int64_t kStaticTargetWaves = 4096;
// Fewest observations a split covers (keeps partials small next to the input).
constexpr int64_t kStaticMinSpan = 4096;

StaticPlan static_plan(int64_t batch, int64_t nv, int64_t nobs, bool corr) {
  StaticPlan p{};
  // 16 x 16 wave tiles by default: a 32 x 32 tile's 96 (corr) float64
  // accumulators do not fit beside its operands without spills (measured on
  // MI355X, float32 (256, 1M) corr: 10.9 ms at 32 x 32). Knob NBG_MATRIX_TW.
  p.tw = knob("NBG_MATRIX_TW", 1) == 2 && nv > 16 ? 2 : 1;
  const int64_t ts = 16 * p.tw;
  p.ntile = (nv + ts - 1) / ts;
  p.tiles = p.ntile * (p.ntile + 1) / 2;
  const int64_t items = batch * p.tiles;
  // Whole rounds of waves: at most the target, so no last round runs a few.
  int64_t splits = knob("NBG_MATRIX_WAVES", kStaticTargetWaves) / (items > 0 ? items : 1);
  const int64_t max_splits = (nobs + kStaticMinSpan - 1) / kStaticMinSpan;
  if (splits > max_splits) splits = max_splits;
  if (splits < 1) splits = 1;
  int64_t span = (nobs + splits - 1) / splits;
  span = (span + 15) / 16 * 16;
  if (span < 16) span = 16;
  p.splits = (nobs + span - 1) / span;
  if (p.splits < 1) p.splits = 1;
  p.span = span;
  p.shift_bytes = (size_t)(batch * nv) * 8;
  const int np = corr ? 6 : 4;
  p.part_bytes = p.splits > 1 ? (size_t)(items * p.splits) * np * ts * ts * 8 : 0;
  return p;
}

// ---------------------------------------------------------------------------
// Moving-window matrices: a lane per ordered pair and time range.
// ---------------------------------------------------------------------------
struct MxMoveArgs {
  const void* x;
  void* out;
  const double* shift;  // (batch, tiles, nv)
  int64_t batch, nobs, nv;
  int64_t window, min_count;  // min_count as the reference adjusts it (>= 1)
  int64_t tile_len, tiles;
  int64_t pair_waves;         // waves per (batch row, tile)
  int64_t groups;             // time ranges per wave (1 when nv * nv >= 64)
  int64_t sub_len;            // steps per time range
};

template <bool CORR>
struct MoveState {
  double si, sj, p, qi, qj;
  int64_t n;
};

// Window sums of one pair: add or remove the observation (xi, xj).
template <typename T, bool CORR>
struct PairSums {
  double si = 0.0, sj = 0.0, sp = 0.0, qi = 0.0, qj = 0.0;
  int64_t n = 0;
  __device__ __forceinline__ void add(T xi, T xj, double ci, double cj, double sign) {
    if (xi == xi && xj == xj) {
      const double yi = (double)xi - ci, yj = (double)xj - cj;
      si += sign * yi;
      sj += sign * yj;
      sp += sign * (yi * yj);
      if (CORR) {
        qi += sign * (yi * yi);
        qj += sign * (yj * yj);
      }
      n += sign > 0.0 ? 1 : -1;
    }
  }
  // The output for the current window, in forms scaled by n so that neither
  // needs a division by n: corr = (n sp - si sj) / sqrt((n qi - si^2)(n qj -
  // sj^2)); cov = (n sp - si sj) / (n (n - 1)).
  __device__ __forceinline__ double value(int64_t min_count, int64_t mcorr, bool diag) const {
    const double dn = (double)n;
    if (CORR) {
      if (n < mcorr) return __builtin_nan("");
      const double vi = dn * qi - si * si, vj = dn * qj - sj * sj;
      if (!(vi > 0.0 && vj > 0.0)) return __builtin_nan("");
      return diag ? 1.0 : mx_corr(dn * sp - si * sj, vi * vj);
    }
    if (n < min_count || n < 2) return __builtin_nan("");
    return (dn * sp - si * sj) / (dn * (dn - 1.0));
  }
};

// IT: the index type of element offsets within one batch row's input and
// output (32-bit when both fit, so addresses are a scalar base plus a 32-bit
// lane offset).
template <typename T, bool CORR, typename IT>
__global__ __launch_bounds__(kMxThreads) void mx_move_kernel(MxMoveArgs a) {
  const int64_t wv = (int64_t)blockIdx.x * kMxWaves + mx_wave_id();
  if (wv >= a.batch * a.tiles * a.pair_waves) return;  // whole wave
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t pw = wv % a.pair_waves;
  const int64_t tile = (wv / a.pair_waves) % a.tiles;
  const int64_t b = wv / (a.pair_waves * a.tiles);
  const int64_t np = a.nv * a.nv;
  int64_t qq, g;
  if (a.groups == 1) {
    qq = pw * kWave + lane;
    g = 0;
  } else {
    qq = lane % np;
    g = lane / np;
  }
  if (qq >= np || g >= a.groups) return;
  const int64_t i = qq / a.nv, j = qq % a.nv;
  const int64_t tile_lo = tile * a.tile_len;
  int64_t tile_hi = tile_lo + a.tile_len;
  if (tile_hi > a.nobs) tile_hi = a.nobs;
  const int64_t t_start = tile_lo + g * a.sub_len;
  int64_t t_end = t_start + a.sub_len;
  if (t_end > tile_hi) t_end = tile_hi;
  if (t_start >= t_end) return;
  const int64_t t_w = t_start - a.window + 1 > 0 ? t_start - a.window + 1 : 0;
  // removals start once a whole window lies after t_w
  const int64_t t_r = t_w + a.window > t_start ? t_w + a.window : t_start;
  const T* x = static_cast<const T*>(a.x) + b * a.nobs * a.nv;
  T* out = static_cast<T*>(a.out) + b * a.nobs * np;
  const double ci = a.shift[(b * a.tiles + tile) * a.nv + i];
  const double cj = a.shift[(b * a.tiles + tile) * a.nv + j];
  const bool diag = i == j;
  const int64_t mcorr = a.min_count > 2 ? a.min_count : 2;
  const IT step = (IT)a.nv, ostep = (IT)np, back = (IT)(a.window * a.nv);
  PairSums<T, CORR> st;
  IT oi = (IT)(t_w * a.nv + i), oj = (IT)(t_w * a.nv + j);
  int64_t t = t_w;
  for (; t < t_start; ++t, oi += step, oj += step) st.add(x[oi], x[oj], ci, cj, 1.0);
  // Outputs (V^2 per step, most of the traffic) by streaming stores: each
  // wave writes 64 consecutive elements per step, and the lines need not stay
  // in the caches (measured on MI355X, (200000, 32) float32, window 20:
  // move_covmatrix 0.528 -> 0.448 ms, move_corrmatrix 0.60 -> 0.56 ms).
  IT oo = (IT)(t_start * np + qq);
  for (; t < t_r && t < t_end; ++t, oi += step, oj += step, oo += ostep) {
    st.add(x[oi], x[oj], ci, cj, 1.0);
    __builtin_nontemporal_store((T)st.value(a.min_count, mcorr, diag), &out[oo]);
  }
  for (; t < t_end; ++t, oi += step, oj += step, oo += ostep) {
    st.add(x[oi], x[oj], ci, cj, 1.0);
    st.add(x[oi - back], x[oj - back], ci, cj, -1.0);
    __builtin_nontemporal_store((T)st.value(a.min_count, mcorr, diag), &out[oo]);
  }
}

// Waves the moving and exponential kernels aim for (8 per SIMD).
constexpr int64_t kMoveTargetWaves = 8192;

struct MovePlan {
  int64_t tile_len, tiles, pair_waves, groups, sub_len;
  size_t shift_bytes;
};

MovePlan move_plan(int64_t batch, int64_t nobs, int64_t nv, int64_t window) {
  MovePlan p{};
  const int64_t np = nv * nv;
  if (np >= kWave) {
    p.groups = 1;
    p.pair_waves = (np + kWave - 1) / kWave;
  } else {
    p.groups = kWave / np;
    p.pair_waves = 1;
  }
  // Time tiles: enough waves to fill the device with several per SIMD, but
  // every time range at least twice the window (the warm-up is repeated per
  // range) and 64 steps.
  const int64_t waves_per_tile = batch * p.pair_waves;
  const int64_t target = knob("NBG_MATRIX_MOVE_WAVES", kMoveTargetWaves);
  int64_t tiles = (target + waves_per_tile - 1) / (waves_per_tile > 0 ? waves_per_tile : 1);
  int64_t min_range = 2 * window > 64 ? 2 * window : 64;
  int64_t max_tiles = nobs / (min_range * p.groups);
  if (max_tiles < 1) max_tiles = 1;
  if (tiles > max_tiles) tiles = max_tiles;
  if (tiles < 1) tiles = 1;
  p.tile_len = (nobs + tiles - 1) / tiles;
  if (p.tile_len < 1) p.tile_len = 1;
  p.tiles = (nobs + p.tile_len - 1) / p.tile_len;
  if (p.tiles < 1) p.tiles = 1;
  p.sub_len = (p.tile_len + p.groups - 1) / p.groups;
  p.shift_bytes = (size_t)(batch * p.tiles * nv) * 8;
  return p;
}

// ---------------------------------------------------------------------------
// Exponentially weighted matrices: chunk states, carries, replay.
// State order (as move_exp_nancov / nancorr): s1, s2, s12, sw, sw2, weight,
// s11, s22; `seen` counts valid pair observations, saturating at 2.
// ---------------------------------------------------------------------------
template <bool CORR>
struct ExpM {
  static constexpr int NS = CORR ? 8 : 6;
  static constexpr int REC = NS + 1;  // + seen
};

struct MxExpArgs {
  const void* x;
  const void* alpha;   // loop dtype, row stride ld_alpha (0: shared)
  void* out;
  const double* shift; // (batch, nv)
  double* state;       // (batch, chunks, nv * nv, REC)
  double* decay;       // (batch, chunks, 2): product of d and of d^2 over the chunk
  int64_t ld_alpha;
  double min_weight;
  int64_t batch, nobs, nv;
  int64_t chunk, chunks;
  int64_t pair_waves, groups;
};

template <bool CORR>
__device__ __forceinline__ void exp_step(double (&s)[ExpM<CORR>::NS], int& seen, double d, double al, bool valid,
                                         double yi, double yj) {
  constexpr int NS = ExpM<CORR>::NS;
#pragma unroll
  for (int c = 0; c < NS; ++c) s[c] *= (c == 4 ? d * d : d);
  if (valid) {
    s[0] += yi;
    s[1] += yj;
    s[2] += yi * yj;
    s[3] += 1.0;
    s[4] += 1.0;
    s[5] += al;
    if (CORR) {
      s[6 % NS] += yi * yi;
      s[7 % NS] += yj * yj;
    }
    seen += seen < 2;
  }
}

// Work item of lane `lane` in wave `wv`: (batch row, chunk, pair); false when
// the lane has none.
__device__ __forceinline__ bool exp_item(const MxExpArgs& a, int64_t wv, int lane, int64_t& b, int64_t& c,
                                         int64_t& qq) {
  const int64_t np = a.nv * a.nv;
  if (a.groups == 1) {
    const int64_t pw = wv % a.pair_waves;
    c = (wv / a.pair_waves) % a.chunks;
    b = wv / (a.pair_waves * a.chunks);
    qq = pw * kWave + lane;
  } else {
    const int64_t cgs = (a.chunks + a.groups - 1) / a.groups;
    const int64_t cg = wv % cgs;
    b = wv / cgs;
    qq = lane % np;
    const int64_t g = lane / np;
    if (g >= a.groups) return false;
    c = cg * a.groups + g;
  }
  return b < a.batch && c < a.chunks && qq < np;
}

// IT: the index type of element offsets within one batch row (as for
// mx_move_kernel).
template <typename TX, typename TA, bool CORR, bool REPLAY, typename IT>
__global__ __launch_bounds__(kMxThreads) void mx_exp_chunk_kernel(MxExpArgs a) {
  using TO = TA;
  constexpr int NS = ExpM<CORR>::NS, REC = ExpM<CORR>::REC;
  const int64_t wv = (int64_t)blockIdx.x * kMxWaves + mx_wave_id();
  const int lane = threadIdx.x & (kWave - 1);
  int64_t b, c, qq;
  if (!exp_item(a, wv, lane, b, c, qq)) return;
  const int64_t np = a.nv * a.nv;
  const int64_t i = qq / a.nv, j = qq % a.nv;
  const int64_t t0 = c * a.chunk;
  const int64_t t1 = t0 + a.chunk < a.nobs ? t0 + a.chunk : a.nobs;
  const TX* x = static_cast<const TX*>(a.x) + b * a.nobs * a.nv;
  const double ci = a.shift[b * a.nv + i], cj = a.shift[b * a.nv + j];
  double* rec = a.state + ((b * a.chunks + c) * np + qq) * REC;
  double s[NS];
  int seen = 0;
  if (REPLAY) {
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = rec[k];
    seen = (int)rec[NS];
  } else {
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = 0.0;
  }
  double dp = 1.0, dp2 = 1.0;
  const double mw = a.min_weight;
  const bool diag = i == j;
  TO* out = static_cast<TO*>(a.out) + b * a.nobs * np;
  const TA* alpha = static_cast<const TA*>(a.alpha) + b * a.ld_alpha;
  const IT step = (IT)a.nv, ostep = (IT)np;
  IT oi = (IT)(t0 * a.nv + i), oj = (IT)(t0 * a.nv + j), oo = (IT)(t0 * np + qq);
  for (int64_t t = t0; t < t1; ++t, oi += step, oj += step, oo += ostep) {
    const double al = (double)alpha[t];
    const double d = 1.0 - al;
    const TX xi = x[oi], xj = x[oj];
    const bool valid = xi == xi && xj == xj;
    exp_step<CORR>(s, seen, d, al, valid, valid ? (double)xi - ci : 0.0, valid ? (double)xj - cj : 0.0);
    if (!REPLAY) {
      dp *= d;
      dp2 *= d * d;
      continue;
    }
    // Forms scaled by sw^2 (sw = s[3]): the bias correction 1 - sw2 / sw^2 is
    // positive where sw^2 - sw2 is, cov = (s12 sw - s1 s2) / (sw^2 - sw2),
    // and corr = (s12 sw - s1 s2) / sqrt((s11 sw - s1^2)(s22 sw - s2^2)).
    double v = __builtin_nan("");
    const double sws = s[3] * s[3];
    if (s[5] >= mw && seen >= 2 && s[4] < sws) {
      const double cv = s[2] * s[3] - s[0] * s[1];
      if (!CORR) {
        v = cv / (sws - s[4]);
      } else {
        const double vi = s[6 % NS] * s[3] - s[0] * s[0], vj = s[7 % NS] * s[3] - s[1] * s[1];
        if (vi > 0.0 && vj > 0.0) v = diag ? 1.0 : mx_corr(cv, vi * vj);
      }
    }
    __builtin_nontemporal_store((TO)v, &out[oo]);  // streaming, as the moving kernel's
  }
  if (!REPLAY) {
#pragma unroll
    for (int k = 0; k < NS; ++k) rec[k] = s[k];
    rec[NS] = (double)seen;
    if (qq == 0) {
      a.decay[(b * a.chunks + c) * 2] = dp;
      a.decay[(b * a.chunks + c) * 2 + 1] = dp2;
    }
  }
}

// One wave per (batch row, pair): replaces each chunk's local state by the
// state at the chunk's start. A chunk's effect on the state is the affine map
// s -> D s + L (D the chunk's decay product, its square for sw2). Each lane
// composes the maps of a contiguous segment of chunks, a shuffle scan over the
// wave composes the segments, and each lane then replays its segment from
// the state its predecessors leave.
template <bool CORR>
__global__ __launch_bounds__(kMxThreads) void mx_exp_carry_kernel(MxExpArgs a) {
  constexpr int NS = ExpM<CORR>::NS, REC = ExpM<CORR>::REC;
  const int64_t np = a.nv * a.nv;
  const int64_t wv = (int64_t)blockIdx.x * kMxWaves + mx_wave_id();
  if (wv >= a.batch * np) return;  // whole wave
  const int lane = threadIdx.x & (kWave - 1);
  const int64_t b = wv / np, qq = wv % np;
  const int64_t seg = (a.chunks + kWave - 1) / kWave;
  const int64_t c0 = lane * seg < a.chunks ? lane * seg : a.chunks;
  const int64_t c1 = c0 + seg < a.chunks ? c0 + seg : a.chunks;
  double D = 1.0, D2 = 1.0, L[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) L[k] = 0.0;
  int seen = 0;
  for (int64_t c = c0; c < c1; ++c) {
    const double* rec = a.state + ((b * a.chunks + c) * np + qq) * REC;
    const double d = a.decay[(b * a.chunks + c) * 2], d2 = a.decay[(b * a.chunks + c) * 2 + 1];
#pragma unroll
    for (int k = 0; k < NS; ++k) L[k] = (k == 4 ? d2 : d) * L[k] + rec[k];
    D *= d;
    D2 *= d2;
    const int ls = (int)rec[NS];
    seen = seen + ls < 2 ? seen + ls : 2;
  }
  // Inclusive scan: lane i ends with the composition of lanes 0..i.
  for (int off = 1; off < kWave; off <<= 1) {
    const double pD = __shfl_up(D, off), pD2 = __shfl_up(D2, off);
    double pL[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) pL[k] = __shfl_up(L[k], off);
    const int ps = __shfl_up(seen, off);
    if (lane >= off) {
#pragma unroll
      for (int k = 0; k < NS; ++k) L[k] = (k == 4 ? D2 : D) * pL[k] + L[k];
      D *= pD;
      D2 *= pD2;
      seen = seen + ps < 2 ? seen + ps : 2;
    }
  }
  // Exclusive: the state entering this lane's segment.
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) {
    const double v = __shfl_up(L[k], 1);
    s[k] = lane > 0 ? v : 0.0;
  }
  const int vs = __shfl_up(seen, 1);
  seen = lane > 0 ? vs : 0;
  for (int64_t c = c0; c < c1; ++c) {
    double* rec = a.state + ((b * a.chunks + c) * np + qq) * REC;
    const double d = a.decay[(b * a.chunks + c) * 2], d2 = a.decay[(b * a.chunks + c) * 2 + 1];
    double loc[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) loc[k] = rec[k];
    const int ls = (int)rec[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) rec[k] = s[k];
    rec[NS] = (double)seen;
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = (k == 4 ? d2 : d) * s[k] + loc[k];
    seen = seen + ls < 2 ? seen + ls : 2;
  }
}

struct ExpPlan {
  int64_t chunk, chunks, pair_waves, groups, waves;
  size_t shift_bytes, state_bytes, decay_bytes;
};

ExpPlan exp_plan(int64_t batch, int64_t nobs, int64_t nv, bool corr) {
  ExpPlan p{};
  const int64_t np = nv * nv;
  if (np >= kWave) {
    p.groups = 1;
    p.pair_waves = (np + kWave - 1) / kWave;
  } else {
    p.groups = kWave / np;
    p.pair_waves = 1;
  }
  // Chunks: enough lanes busy to fill the device with several waves per SIMD,
  // at least 64 steps each.
  const int64_t lanes_per_chunk = batch * (p.groups == 1 ? p.pair_waves * kWave : np);
  const int64_t target = knob("NBG_MATRIX_MOVE_WAVES", kMoveTargetWaves) * kWave;
  int64_t chunks = (target + lanes_per_chunk - 1) / (lanes_per_chunk > 0 ? lanes_per_chunk : 1);
  // At most 64 chunks per lane of the carry scan's wave: its segments are
  // walked one chunk after another.
  int64_t max_chunks = nobs / 64;
  if (max_chunks > kWave * 64) max_chunks = kWave * 64;
  if (max_chunks < 1) max_chunks = 1;
  if (chunks > max_chunks) chunks = max_chunks;
  if (chunks < 1) chunks = 1;
  p.chunk = (nobs + chunks - 1) / chunks;
  if (p.chunk < 1) p.chunk = 1;
  p.chunks = (nobs + p.chunk - 1) / p.chunk;
  if (p.chunks < 1) p.chunks = 1;
  p.waves = p.groups == 1 ? batch * p.chunks * p.pair_waves : batch * ((p.chunks + p.groups - 1) / p.groups);
  p.shift_bytes = (size_t)(batch * nv) * 8;
  p.state_bytes = (size_t)(batch * p.chunks * np) * (corr ? ExpM<true>::REC : ExpM<false>::REC) * 8;
  p.decay_bytes = (size_t)(batch * p.chunks) * 2 * 8;
  return p;
}

size_t align256(size_t v) { return (v + 255) / 256 * 256; }

int mx_grid(int64_t waves) {
  const int64_t blocks = (waves + kMxWaves - 1) / kMxWaves;
  return (int)(blocks < 1 ? 1 : blocks);
}

// Every grid of a launch fits a 32-bit block count.
bool grids_fit(std::initializer_list<int64_t> waves) {
  for (int64_t w : waves)
    if (w < 0 || (w + kMxWaves - 1) / kMxWaves > 0x7fffffffLL) return false;
  return true;
}

template <typename T>
int static_launch(bool corr, const void* x, void* out, int64_t batch, int64_t nv, int64_t nobs, void* ws,
                  hipStream_t s) {
  const StaticPlan p = static_plan(batch, nv, nobs, corr);
  if (!grids_fit({batch * nv, batch * p.tiles * p.splits, batch * p.tiles * 64}))
    return set_error(NBG_ERR_INVALID, "matrix: batch too large for one launch");
  char* w = static_cast<char*>(ws);
  double* shift = reinterpret_cast<double*>(w);
  double* part = reinterpret_cast<double*>(w + align256(p.shift_bytes));
  MxShiftArgs sa{batch, nv, 1, nobs, 0, nobs, nv * nobs, nobs, 1};
  hipLaunchKernelGGL(mx_shift_kernel<T>, dim3(mx_grid(batch * nv)), dim3(kMxThreads), 0, s,
                     static_cast<const T*>(x), shift, sa);
  MxStaticArgs a{x, out, shift, part, batch, nv, nobs, p.ntile, p.tiles, p.splits, p.span};
  if (nv <= kSmallV && batch >= kSmallMinBatch && batch <= 0x7fffffffLL &&
      knob("NBG_MATRIX_SMALL", 1) != 0) {  // a workgroup per batch row on the vector units
    const int64_t rows_d = nv * (kSmallChunk + 1), red_d = (corr ? 6 : 4) * (int64_t)kMxThreads;
    const size_t lds = (size_t)(rows_d > red_d ? rows_d : red_d) * sizeof(double);
    if (corr) hipLaunchKernelGGL((mx_small_kernel<T, true>), dim3((unsigned)batch), dim3(kMxThreads), lds, s, a);
    else hipLaunchKernelGGL((mx_small_kernel<T, false>), dim3((unsigned)batch), dim3(kMxThreads), lds, s, a);
    return check_hip(hipGetLastError(), "matrix launch");
  }
  const bool vec = ((uintptr_t)x % 16) == 0 && (nobs % (16 / (int64_t)sizeof(T))) == 0;
  const int grid = mx_grid(batch * p.tiles * p.splits);
#define NBG_MX_STATIC(C, TW)                                                                                   \
  do {                                                                                                         \
    if (vec) hipLaunchKernelGGL((mx_static_kernel<T, C, TW, true>), dim3(grid), dim3(kMxThreads), 0, s, a);   \
    else hipLaunchKernelGGL((mx_static_kernel<T, C, TW, false>), dim3(grid), dim3(kMxThreads), 0, s, a);      \
    if (p.splits > 1)                                                                                          \
      hipLaunchKernelGGL((mx_static_finish_kernel<T, C, TW>),                                                  \
                         dim3((unsigned)(batch * p.tiles * (256 * TW * TW / kFinishPos))), dim3(kMxThreads), 0, s,  \
                         a);                                                                                   \
  } while (0)
  if (corr) {
    if (p.tw == 2) NBG_MX_STATIC(true, 2);
    else NBG_MX_STATIC(true, 1);
  } else {
    if (p.tw == 2) NBG_MX_STATIC(false, 2);
    else NBG_MX_STATIC(false, 1);
  }
#undef NBG_MX_STATIC
  return check_hip(hipGetLastError(), "matrix launch");
}

size_t static_ws(int64_t batch, int64_t nv, int64_t nobs, bool corr) {
  const StaticPlan p = static_plan(batch, nv, nobs, corr);
  return align256(p.shift_bytes) + p.part_bytes;
}

template <typename T>
int move_launch(bool corr, const void* x, void* out, int64_t batch, int64_t nobs, int64_t nv, int64_t window,
                int64_t min_count, void* ws, hipStream_t s) {
  const MovePlan p = move_plan(batch, nobs, nv, window);
  if (!grids_fit({batch * p.tiles * nv, batch * p.tiles * p.pair_waves}))
    return set_error(NBG_ERR_INVALID, "matrix: batch too large for one launch");
  double* shift = static_cast<double*>(ws);
  MxShiftArgs sa{batch * p.tiles, nv, p.tiles, p.tile_len, window - 1, nobs, nobs * nv, 1, nv};
  hipLaunchKernelGGL(mx_shift_kernel<T>, dim3(mx_grid(batch * p.tiles * nv)), dim3(kMxThreads), 0, s,
                     static_cast<const T*>(x), shift, sa);
  MxMoveArgs a{x, out, shift, batch, nobs, nv, window, min_count > 1 ? min_count : 1,
               p.tile_len, p.tiles, p.pair_waves, p.groups, p.sub_len};
  const int grid = mx_grid(batch * p.tiles * p.pair_waves);
  const bool small = nobs * nv * nv < ((int64_t)1 << 31);  // output offsets (the larger) fit 32 bits
#define NBG_MX_MOVE(C, IT) hipLaunchKernelGGL((mx_move_kernel<T, C, IT>), dim3(grid), dim3(kMxThreads), 0, s, a)
  if (corr) {
    if (small) NBG_MX_MOVE(true, uint32_t);
    else NBG_MX_MOVE(true, int64_t);
  } else {
    if (small) NBG_MX_MOVE(false, uint32_t);
    else NBG_MX_MOVE(false, int64_t);
  }
#undef NBG_MX_MOVE
  return check_hip(hipGetLastError(), "matrix launch");
}

template <typename TX, typename TA>
int exp_launch(bool corr, const void* x, const void* alpha, int64_t ld_alpha, double min_weight, void* out,
               int64_t batch, int64_t nobs, int64_t nv, void* ws, hipStream_t s) {
  const ExpPlan p = exp_plan(batch, nobs, nv, corr);
  if (!grids_fit({batch * nv, p.waves, batch * nv * nv}))
    return set_error(NBG_ERR_INVALID, "matrix: batch too large for one launch");
  char* w = static_cast<char*>(ws);
  double* shift = reinterpret_cast<double*>(w);
  double* state = reinterpret_cast<double*>(w + align256(p.shift_bytes));
  double* decay = reinterpret_cast<double*>(w + align256(p.shift_bytes) + align256(p.state_bytes));
  MxShiftArgs sa{batch, nv, 1, nobs, 0, nobs, nobs * nv, 1, nv};
  hipLaunchKernelGGL(mx_shift_kernel<TX>, dim3(mx_grid(batch * nv)), dim3(kMxThreads), 0, s,
                     static_cast<const TX*>(x), shift, sa);
  MxExpArgs a{x, alpha, out, shift, state, decay, ld_alpha, min_weight, batch, nobs, nv,
              p.chunk, p.chunks, p.pair_waves, p.groups};
  const int grid = mx_grid(p.waves);
  const int carry_grid = mx_grid(batch * nv * nv);
  const bool small = nobs * nv * nv < ((int64_t)1 << 31);  // output offsets (the larger) fit 32 bits
#define NBG_MX_EXP(C, IT)                                                                                      \
  do {                                                                                                         \
    hipLaunchKernelGGL((mx_exp_chunk_kernel<TX, TA, C, false, IT>), dim3(grid), dim3(kMxThreads), 0, s, a);   \
    hipLaunchKernelGGL((mx_exp_carry_kernel<C>), dim3(carry_grid), dim3(kMxThreads), 0, s, a);                \
    hipLaunchKernelGGL((mx_exp_chunk_kernel<TX, TA, C, true, IT>), dim3(grid), dim3(kMxThreads), 0, s, a);    \
  } while (0)
  if (corr) {
    if (small) NBG_MX_EXP(true, uint32_t);
    else NBG_MX_EXP(true, int64_t);
  } else {
    if (small) NBG_MX_EXP(false, uint32_t);
    else NBG_MX_EXP(false, int64_t);
  }
#undef NBG_MX_EXP
  return check_hip(hipGetLastError(), "matrix launch");
}

size_t exp_ws(int64_t batch, int64_t nobs, int64_t nv, bool corr) {
  const ExpPlan p = exp_plan(batch, nobs, nv, corr);
  return align256(p.shift_bytes) + align256(p.state_bytes) + p.decay_bytes;
}

bool sizes_ok(int64_t a, int64_t b, int64_t c) { return a >= 0 && b >= 0 && c >= 0; }

}  // namespace
}  // namespace nbg

extern "C" {

size_t nbg_matrix_workspace_bytes(int64_t batch, int64_t nvars, int64_t nobs) {
  if (!nbg::sizes_ok(batch, nvars, nobs) || batch * nvars * nobs == 0) return 0;
  return nbg::static_ws(batch, nvars, nobs, true);
}

static int static_entry(const char* name, bool corr, int dtype, const void* x, void* out, int64_t batch,
                        int64_t nvars, int64_t nobs, void* ws, size_t ws_bytes, void* stream) {
  if (int rc = nbg::injected_error(name)) return rc;
  if (!nbg::sizes_ok(batch, nvars, nobs)) return nbg::set_error(NBG_ERR_INVALID, "%s: negative size", name);
  if (batch == 0 || nvars == 0) return NBG_OK;
  if (x == nullptr && nobs > 0) return nbg::set_error(NBG_ERR_INVALID, "%s: null input", name);
  if (out == nullptr) return nbg::set_error(NBG_ERR_INVALID, "%s: null output", name);
  if (nobs == 0) {  // every pair has no observation: NaN
    return nbg::set_error(NBG_ERR_INVALID, "%s: no observations (the host fills NaN)", name);
  }
  if (ws == nullptr || ws_bytes < nbg::static_ws(batch, nvars, nobs, corr))
    return nbg::set_error(NBG_ERR_INVALID, "%s: workspace too small", name);
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case NBG_F32: return nbg::static_launch<float>(corr, x, out, batch, nvars, nobs, ws, s);
    case NBG_F64: return nbg::static_launch<double>(corr, x, out, batch, nvars, nobs, ws, s);
    default: return nbg::set_error(NBG_ERR_INVALID, "%s: unsupported dtype code %d", name, dtype);
  }
}

int nbg_nancorrmatrix(int dtype, const void* x, void* out, int64_t batch, int64_t nvars, int64_t nobs, void* ws,
                      size_t ws_bytes, void* stream) {
  return static_entry("nbg_nancorrmatrix", true, dtype, x, out, batch, nvars, nobs, ws, ws_bytes, stream);
}

int nbg_nancovmatrix(int dtype, const void* x, void* out, int64_t batch, int64_t nvars, int64_t nobs, void* ws,
                     size_t ws_bytes, void* stream) {
  return static_entry("nbg_nancovmatrix", false, dtype, x, out, batch, nvars, nobs, ws, ws_bytes, stream);
}

size_t nbg_move_matrix_workspace_bytes(int64_t batch, int64_t nobs, int64_t nvars, int64_t window) {
  if (!nbg::sizes_ok(batch, nobs, nvars) || window < 1 || batch * nobs * nvars == 0) return 0;
  return nbg::move_plan(batch, nobs, nvars, window).shift_bytes;
}

static int move_entry(const char* name, bool corr, int dtype, const void* x, void* out, int64_t batch, int64_t nobs,
                      int64_t nvars, int64_t window, int64_t min_count, void* ws, size_t ws_bytes, void* stream) {
  if (int rc = nbg::injected_error(name)) return rc;
  if (!nbg::sizes_ok(batch, nobs, nvars)) return nbg::set_error(NBG_ERR_INVALID, "%s: negative size", name);
  if (window < 1 || window > nobs) return nbg::set_error(NBG_ERR_INVALID, "%s: window not in valid range: %lld", name,
                                                         (long long)window);
  if (min_count < 0) return nbg::set_error(NBG_ERR_INVALID, "%s: min_count must be positive", name);
  if (batch == 0 || nobs == 0 || nvars == 0) return NBG_OK;
  if (x == nullptr || out == nullptr) return nbg::set_error(NBG_ERR_INVALID, "%s: null pointer", name);
  if (ws == nullptr || ws_bytes < nbg::move_plan(batch, nobs, nvars, window).shift_bytes)
    return nbg::set_error(NBG_ERR_INVALID, "%s: workspace too small", name);
  hipStream_t s = (hipStream_t)stream;
  switch (dtype) {
    case NBG_F32: return nbg::move_launch<float>(corr, x, out, batch, nobs, nvars, window, min_count, ws, s);
    case NBG_F64: return nbg::move_launch<double>(corr, x, out, batch, nobs, nvars, window, min_count, ws, s);
    default: return nbg::set_error(NBG_ERR_INVALID, "%s: unsupported dtype code %d", name, dtype);
  }
}

int nbg_move_corrmatrix(int dtype, const void* x, void* out, int64_t batch, int64_t nobs, int64_t nvars,
                        int64_t window, int64_t min_count, void* ws, size_t ws_bytes, void* stream) {
  return move_entry("nbg_move_corrmatrix", true, dtype, x, out, batch, nobs, nvars, window, min_count, ws, ws_bytes,
                    stream);
}

int nbg_move_covmatrix(int dtype, const void* x, void* out, int64_t batch, int64_t nobs, int64_t nvars,
                       int64_t window, int64_t min_count, void* ws, size_t ws_bytes, void* stream) {
  return move_entry("nbg_move_covmatrix", false, dtype, x, out, batch, nobs, nvars, window, min_count, ws, ws_bytes,
                    stream);
}

size_t nbg_move_exp_matrix_workspace_bytes(int64_t batch, int64_t nobs, int64_t nvars) {
  if (!nbg::sizes_ok(batch, nobs, nvars) || batch * nobs * nvars == 0) return 0;
  return nbg::exp_ws(batch, nobs, nvars, true);
}

static int exp_entry(const char* name, bool corr, int x_dtype, int dtype, const void* x, const void* alpha,
                     int64_t ld_alpha, double min_weight, void* out, int64_t batch, int64_t nobs, int64_t nvars,
                     void* ws, size_t ws_bytes, void* stream) {
  if (int rc = nbg::injected_error(name)) return rc;
  if (!nbg::sizes_ok(batch, nobs, nvars) || ld_alpha < 0)
    return nbg::set_error(NBG_ERR_INVALID, "%s: negative size", name);
  if (batch == 0 || nobs == 0 || nvars == 0) return NBG_OK;
  if (x == nullptr || out == nullptr || alpha == nullptr) return nbg::set_error(NBG_ERR_INVALID, "%s: null pointer", name);
  if (ws == nullptr || ws_bytes < nbg::exp_ws(batch, nobs, nvars, corr))
    return nbg::set_error(NBG_ERR_INVALID, "%s: workspace too small", name);
  hipStream_t s = (hipStream_t)stream;
  if (x_dtype == NBG_F32 && dtype == NBG_F32)
    return nbg::exp_launch<float, float>(corr, x, alpha, ld_alpha, min_weight, out, batch, nobs, nvars, ws, s);
  if (x_dtype == NBG_F32 && dtype == NBG_F64)
    return nbg::exp_launch<float, double>(corr, x, alpha, ld_alpha, min_weight, out, batch, nobs, nvars, ws, s);
  if (x_dtype == NBG_F64 && dtype == NBG_F64)
    return nbg::exp_launch<double, double>(corr, x, alpha, ld_alpha, min_weight, out, batch, nobs, nvars, ws, s);
  return nbg::set_error(NBG_ERR_INVALID, "%s: unsupported dtype pair (%d, %d)", name, x_dtype, dtype);
}

int nbg_move_exp_nancorrmatrix(int x_dtype, int dtype, const void* x, const void* alpha, int64_t ld_alpha,
                               double min_weight, void* out, int64_t batch, int64_t nobs, int64_t nvars, void* ws,
                               size_t ws_bytes, void* stream) {
  return exp_entry("nbg_move_exp_nancorrmatrix", true, x_dtype, dtype, x, alpha, ld_alpha, min_weight, out, batch,
                   nobs, nvars, ws, ws_bytes, stream);
}

int nbg_move_exp_nancovmatrix(int x_dtype, int dtype, const void* x, const void* alpha, int64_t ld_alpha,
                              double min_weight, void* out, int64_t batch, int64_t nobs, int64_t nvars, void* ws,
                              size_t ws_bytes, void* stream) {
  return exp_entry("nbg_move_exp_nancovmatrix", false, x_dtype, dtype, x, alpha, ld_alpha, min_weight, out, batch,
                   nobs, nvars, ws, ws_bytes, stream);
}

}  // extern "C"
