// Exponentially weighted moving statistics (move_exp_nan{count,mean,sum,var,std,
// cov,corr}) for gfx950.
//
// Reference semantics (numbagg/moving_exp.py:12-335): per position i, with
// decay d_i = 1 - alpha_i, every float64 state component is multiplied by d_i
// (sum_weight_2 by d_i**2), then a valid observation adds its contribution
// (x, 1, alpha_i, x*x, ...); the output is a function of the state, NaN when the
// decayed weight is below min_weight (and for var/cov/corr when the bias
// correction is not positive).
//
// The recurrence is affine with a diagonal linear part, s_i = D_i*s_{i-1} + b_i,
// so a chunk of positions acts on the state as one map (P, P2, B) with
// P = prod d, P2 = prod d*d, B = the chunk's state started from zero. Design:
//   * a workgroup owns a tile of 256 threads x K positions of one row, staged in
//     LDS; each thread builds the map of its K positions;
//   * an exclusive block scan composes the maps of earlier threads;
//   * the tile's incoming state comes from a decoupled look-back over earlier
//     tiles of the row (tile order is fixed by per-row tickets, so a tile only
//     waits on tiles that are already running), which publish their aggregate
//     map first and their inclusive state as soon as it is known; one wave
//     examines 64 predecessors per pass;
//   * each thread re-runs its K positions from its exact incoming state, one
//     fused multiply-add per state component (state = decay * state +
//     contribution);
//   * on request (NBG_EXP_TWO_PASS=1) the same tile code runs in two passes
//     instead: independent tiles write their
//     aggregate maps, one workgroup per row scans them into each tile's
//     incoming state, and independent tiles re-run their positions from it
//     (the input is read twice, but no tile waits on another).
// Squares and cross products of float32 inputs are formed in float64 (exact),
// where the reference rounds them to float32 first; the moment kinds shift
// each row by its first finite value, so data far from zero (a random walk
// offset by 1e4) does not lose its variance to the cancellation of
// sum(w x^2)/sum(w) against the squared mean.
#include "common.hpp"

namespace nbg {
namespace {

constexpr int kExpThreads = 256;
constexpr int kExpWaves = kExpThreads / kWave;
constexpr int kMaxNS = 8;
constexpr int kPayload = 2 * (kMaxNS + 3);  // doubles per tile: aggregate map + inclusive state

enum ExpKind { kECount = 0, kEMean = 1, kESum = 2, kEVar = 3, kEStd = 4, kECov = 5, kECorr = 6 };

template <int KIND>
struct ExpSpec;
// Component layout per kind. SW2: index of sum_weight_2 (decays by d*d), or -1.
template <>
struct ExpSpec<kECount> {  // count, weight
  static constexpr int NS = 2, SW2 = -1, NARGS = 1;
};
template <>
struct ExpSpec<kEMean> {  // numer, denom, weight
  static constexpr int NS = 3, SW2 = -1, NARGS = 1;
};
template <>
struct ExpSpec<kESum> {  // numer, weight (+ seen flag)
  static constexpr int NS = 2, SW2 = -1, NARGS = 1;
};
template <>
struct ExpSpec<kEVar> {  // sum_x2, sum_x, sum_w, sum_w2, weight
  static constexpr int NS = 5, SW2 = 3, NARGS = 1;
};
template <>
struct ExpSpec<kEStd> : ExpSpec<kEVar> {};
template <>
struct ExpSpec<kECov> {  // s1, s2, s12, sw, sw2, weight
  static constexpr int NS = 6, SW2 = 4, NARGS = 2;
};
template <>
struct ExpSpec<kECorr> {  // s1, s2, s12, sw, sw2, weight, s11, s22
  static constexpr int NS = 8, SW2 = 4, NARGS = 2;
};

// `seen` counts valid observations, saturating at 2: nansum needs one (NaN before
// the first value), var/std/cov/corr need two (with a single observation the
// bias correction 1 - sw2/sw^2 is exactly zero, though rounding can leave it a
// few ULPs either side of zero).
template <int NS>
struct Map {
  double p, p2;
  double b[NS];
  int seen;
  __device__ __forceinline__ static Map identity() {
    Map m;
    m.p = 1.0;
    m.p2 = 1.0;
#pragma unroll
    for (int c = 0; c < NS; ++c) m.b[c] = 0.0;
    m.seen = 0;
    return m;
  }
};

// Apply `later` after `earlier`. Kinds without a d*d component (SW2 < 0)
// leave p2 at 1.
template <int NS, int SW2>
__device__ __forceinline__ Map<NS> compose(const Map<NS>& earlier, const Map<NS>& later) {
  Map<NS> r;
  r.p = later.p * earlier.p;
  r.p2 = SW2 >= 0 ? later.p2 * earlier.p2 : 1.0;
#pragma unroll
  for (int c = 0; c < NS; ++c) r.b[c] = (c == SW2 ? later.p2 : later.p) * earlier.b[c] + later.b[c];
  r.seen = min(2, earlier.seen + later.seen);
  return r;
}

// A map moved across lanes by DPP (CTRL / RM as for dpp_move); lanes the
// move does not reach get the identity map.
template <int CTRL, int RM, int NS, int SW2>
__device__ __forceinline__ Map<NS> dpp_map(const Map<NS>& m) {
  Map<NS> r;
  r.p = dpp_move<CTRL, RM>(1.0, m.p);
  r.p2 = SW2 >= 0 ? dpp_move<CTRL, RM>(1.0, m.p2) : 1.0;
#pragma unroll
  for (int c = 0; c < NS; ++c) r.b[c] = dpp_move<CTRL, RM>(0.0, m.b[c]);
  r.seen = dpp_move<CTRL, RM>(0, m.seen);
  return r;
}

// A map as held by one lane (wave-uniform index).
template <int NS>
__device__ __forceinline__ Map<NS> read_lane_map(const Map<NS>& m, int lane) {
  Map<NS> r;
  r.p = read_lane(m.p, lane);
  r.p2 = read_lane(m.p2, lane);
#pragma unroll
  for (int c = 0; c < NS; ++c) r.b[c] = read_lane(m.b[c], lane);
  r.seen = read_lane(m.seen, lane);
  return r;
}

// Inclusive wave scan of maps in lane order (DPP row shifts by 1, 2, 4, 8,
// then broadcasts of lanes 15 and 31 into the later rows), and the exclusive
// scan by one whole-wave shift (lane 0 gets the identity).
template <int NS, int SW2>
__device__ __forceinline__ void wave_scan_maps(const Map<NS>& m, Map<NS>& incl, Map<NS>& excl) {
  incl = m;
  incl = compose<NS, SW2>(dpp_map<0x111, 0xf, NS, SW2>(incl), incl);
  incl = compose<NS, SW2>(dpp_map<0x112, 0xf, NS, SW2>(incl), incl);
  incl = compose<NS, SW2>(dpp_map<0x114, 0xf, NS, SW2>(incl), incl);
  incl = compose<NS, SW2>(dpp_map<0x118, 0xf, NS, SW2>(incl), incl);
  incl = compose<NS, SW2>(dpp_map<0x142, 0xa, NS, SW2>(incl), incl);
  incl = compose<NS, SW2>(dpp_map<0x143, 0xc, NS, SW2>(incl), incl);
  excl = dpp_map<0x138, 0xf, NS, SW2>(incl);  // wave_shr:1
}

// One step of the reference recurrence on state s (components per ExpSpec):
// every component decays (by d, sum_weight_2 by d*d) and, when the
// observation is valid, gains its contribution. Written branch-free as one
// fused multiply-add per component, s = decay * s + contribution, with
// invalid observations contributing 0: one rounding per step where the
// reference's `s *= d; s += x` has two, so the state stays at least as close
// to the exact recurrence.
template <int KIND>
__device__ __forceinline__ void exp_step(double* s, int& seen, double a, double b, double alpha) {
  using Sp = ExpSpec<KIND>;
  const double d = 1.0 - alpha;
  const double d2 = d * d;
  const bool valid = (Sp::NARGS == 2) ? !(is_nan(a) || is_nan(b)) : !is_nan(a);
  const double v = valid ? 1.0 : 0.0;
  const double a0 = valid ? a : 0.0;
  const double b0 = (Sp::NARGS == 2 && valid) ? b : 0.0;
  const double w = v * alpha;
  seen = min(2, seen + (int)valid);
  if (KIND == kECount) {
    s[0] = fma(d, s[0], v);
    s[1] = fma(d, s[1], w);
  } else if (KIND == kEMean) {
    s[0] = fma(d, s[0], a0);
    s[1] = fma(d, s[1], v);
    s[2] = fma(d, s[2], w);
  } else if (KIND == kESum) {
    s[0] = fma(d, s[0], a0);
    s[1] = fma(d, s[1], w);
  } else if (KIND == kEVar || KIND == kEStd) {
    s[0] = fma(d, s[0], a0 * a0);
    s[1] = fma(d, s[1], a0);
    s[2] = fma(d, s[2], v);
    s[3] = fma(d2, s[3], v);
    s[4] = fma(d, s[4], w);
  } else {
    s[0] = fma(d, s[0], a0);
    s[1] = fma(d, s[1], b0);
    s[2] = fma(d, s[2], a0 * b0);
    s[3] = fma(d, s[3], v);
    s[4] = fma(d2, s[4], v);
    s[5] = fma(d, s[5], w);
    if (KIND == kECorr) {
      s[6] = fma(d, s[6], a0 * a0);
      s[7] = fma(d, s[7], b0 * b0);
    }
  }
}

// Division by a shared reciprocal. r: 1/b from the hardware estimate
// refined by one Newton step (about 1 ULP of float64); div_rn then corrects
// a * r with the exact remainder a - b * (a * r) from one fused multiply-add,
// which gives the correctly rounded quotient a / b (the reference's division)
// except in rare near-tie cases, where it is 1 ULP off. Valid while
// |a|, |b| and |1/b| stay within 2**500 (no overflow, no subnormals).
__device__ __forceinline__ double recip_refined(double b) {
  const double r = __builtin_amdgcn_rcp(b);
  return fma(r, fma(-b, r, 1.0), r);
}
__device__ __forceinline__ double div_rn(double a, double b, double r) {
  const double q = a * r;
  return fma(fma(-b, q, a), r, q);
}
__device__ __forceinline__ bool div_range(double x) { return fabs(x) <= 0x1p500; }
__device__ __forceinline__ bool recip_range(double b) { return fabs(b) >= 0x1p-500 && fabs(b) <= 0x1p500; }
// 1/sqrt(v) for v in [2**-1000, 2**1000]: estimate plus one Newton step.
__device__ __forceinline__ double rsqrt_refined(double v) {
  const double y = __builtin_amdgcn_rsq(v);
  return y * fma(-(0.5 * v) * y, y, 1.5);
}

// Outputs in the reference's formulas and operation order. corr, the
// costliest, divides by the weight sum through one shared reciprocal with
// corrected quotients (div_rn) when every operand is in range — the common
// case — and forms cov / sqrt(v1 v2) with a refined 1/sqrt; the other kinds
// keep ordinary divisions (a shared reciprocal measured no faster there).
template <int KIND>
__device__ __forceinline__ double exp_out(const double* s, int seen, double mw) {
  const double nan = canonical_nan<double>();
  if (KIND == kECount) return s[1] >= mw ? s[0] : nan;
  if (KIND == kEMean) return s[2] >= mw ? s[0] / s[1] : nan;
  if (KIND == kESum) return (s[1] >= mw && seen) ? s[0] : nan;
  if (KIND == kEVar || KIND == kEStd) {
    const double m = s[1] / s[2];
    const double var_biased = s[0] / s[2] - m * m;
    const double bias = 1.0 - s[3] / (s[2] * s[2]);
    if (!(s[4] >= mw && bias > 0 && seen >= 2)) return nan;
    const double v = var_biased / bias;
    return KIND == kEStd ? sqrt(v) : v;
  }
  const double bias = 1.0 - s[4] / (s[3] * s[3]);
  if (KIND == kECov) {
    const double cov_biased = (s[2] - s[0] * s[1] / s[3]) / s[3];
    return (s[5] >= mw && bias > 0 && seen >= 2) ? cov_biased / bias : nan;
  }
  // kECorr
  const double p01 = s[0] * s[1], p00 = s[0] * s[0], p11 = s[1] * s[1];
  double cov, v1, v2;
  if (recip_range(s[3]) && div_range(p01) && div_range(p00) && div_range(p11)) {
    const double r = recip_refined(s[3]);
    cov = s[2] - div_rn(p01, s[3], r);
    v1 = s[6] - div_rn(p00, s[3], r);
    v2 = s[7] - div_rn(p11, s[3], r);
  } else {
    cov = s[2] - p01 / s[3];
    v1 = s[6] - p00 / s[3];
    v2 = s[7] - p11 / s[3];
  }
  if (!(s[5] >= mw && bias > 0 && seen >= 2)) return nan;
  const double vv = v1 * v2;
  if (!(vv > 0)) return nan;
  return (vv >= 0x1p-1000 && vv <= 0x1p1000) ? cov * rsqrt_refined(vv) : cov / sqrt(vv);
}


// Positions per thread: 64 bytes of input; 32 float32 positions for the
// kinds with small maps (count, mean, sum: twice as many positions per tile,
// half as many look-back hops per row). Measured on MI355X for float32
// (256, 2M), alpha 0.01, 16 -> 32 positions: mean 2.06 -> 1.77 ms, sum /
// count 1.99 / 2.00 -> 1.62 / 1.61 ms; var 2.51 -> 3.06 ms and corr 3.16 ->
// 4.37 ms, so the moment kinds keep 16.
template <typename TX, int KIND>
constexpr int exp_items() {
  return (sizeof(TX) == 4 && ExpSpec<KIND>::NS <= 3) ? 32 : 64 / (int)sizeof(TX);
}

struct ExpArgs {
  const void* xa;
  const void* xb;
  const void* alpha;  // nullptr: scalar alpha
  int64_t ld_alpha;   // 0: one alpha row shared by all rows
  double alpha_scalar;
  void* out;
  int64_t rows, n, ld_a, ld_b, ld_out;
  double min_weight;
  int64_t tiles_per_row;
  unsigned int* flags;
  unsigned int* counters;
  double* payload;
  unsigned int* timeouts;
};

// MODE 0: one pass, incoming states by decoupled look-back. MODE 1: aggregate
// pass (each tile's aggregate map to the payload, no outputs). MODE 2: apply
// pass (each tile's incoming state read from the payload, where
// exp_carry_kernel put it). Modes 1 and 2 process tiles in row-major order on
// XCD-consecutive workgroups and never wait.
template <typename TX, typename TA, int KIND, bool SCALAR, int MODE = 0>
__global__ __launch_bounds__(kExpThreads) void exp_kernel(ExpArgs A) {
  using Sp = ExpSpec<KIND>;
  constexpr int NS = Sp::NS;
  constexpr int K = exp_items<TX, KIND>();
  constexpr int TILE = kExpThreads * K;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ Map<NS> s_wave[kExpWaves];
  __shared__ double s_in[NS];
  __shared__ int s_in_seen;
  __shared__ int s_ticket;

  const TX* xa = static_cast<const TX*>(A.xa);
  const TX* xb = static_cast<const TX*>(A.xb);
  const TA* al = static_cast<const TA*>(A.alpha);
  TA* out = static_cast<TA*>(A.out);

  const uint32_t v_id = MODE == 0 ? 0u : xcd_order(blockIdx.x, gridDim.x);
  const int64_t row = MODE == 0 ? (int64_t)(blockIdx.x % (uint64_t)A.rows)
                                : (int64_t)(v_id / (uint64_t)A.tiles_per_row);
  const TX* ra = xa + row * A.ld_a;
  const TX* rb = (Sp::NARGS == 2) ? xb + row * A.ld_b : nullptr;
  const TA* rl = al ? al + row * A.ld_alpha : nullptr;
  TA* ro = out + row * A.ld_out;
  const int tid = threadIdx.x;
  const int64_t n = A.n;
  TX* la = reinterpret_cast<TX*>(smem_raw);
  TX* lb = la + pad16(TILE) + 16;
  constexpr int VEC = 16 / (int)sizeof(TX);

  // Stage tile tk in LDS (coalesced across the block; NaN past the row end).
  auto stage = [&](int64_t tk) {
    const int64_t q0 = tk * TILE;
    const bool q_vec = (q0 + TILE <= n) && (((uintptr_t)(ra + q0)) % 16 == 0) &&
                       (Sp::NARGS == 1 || ((uintptr_t)(rb + q0)) % 16 == 0);
    if (q_vec) {
      struct alignas(16) V { TX v[VEC]; };
      for (int e = tid * VEC; e < TILE; e += kExpThreads * VEC) {
        V va = load_stream(reinterpret_cast<const V*>(ra + q0 + e));
#pragma unroll
        for (int v = 0; v < VEC; ++v) la[pad16(e + v)] = va.v[v];
        if (Sp::NARGS == 2) {
          V vb = load_stream(reinterpret_cast<const V*>(rb + q0 + e));
#pragma unroll
          for (int v = 0; v < VEC; ++v) lb[pad16(e + v)] = vb.v[v];
        }
      }
    } else {
      for (int e = tid; e < TILE; e += kExpThreads) {
        const int64_t p = q0 + e;
        la[pad16(e)] = p < n ? ra[p] : canonical_nan<TX>();
        if (Sp::NARGS == 2) lb[pad16(e)] = p < n ? rb[p] : canonical_nan<TX>();
      }
    }
  };
  // Blocks of a row are dispatched in order, so the ticket is almost always
  // blockIdx / rows: that tile is staged while the ticket atomic is in flight
  // and restaged only when the ticket differs.
  int64_t tile;
  if constexpr (MODE == 0) {
    const int64_t guess = (int64_t)(blockIdx.x / (uint64_t)A.rows);
    if (threadIdx.x == 0) s_ticket = (int)atomicAdd(&A.counters[row], 1u);
    stage(guess);
    __syncthreads();
    tile = s_ticket;
    if (tile != guess) {
      __syncthreads();
      stage(tile);
      __syncthreads();
    }
  } else {
    tile = (int64_t)(v_id % (uint64_t)A.tiles_per_row);
    stage(tile);
    __syncthreads();
  }
  const int64_t t0 = tile * TILE;
  const bool full = t0 + TILE <= n;

  const int64_t e0 = (int64_t)tid * K;
  // The moment kinds (var, std, cov, corr) run the recurrence on x - ka (and
  // y - kb) with ka the row's first finite value among its first 64
  // positions (0 if there is none). Every wave of every tile of the row finds
  // the same shift from the same positions, so no tile waits on another for
  // it. The squares and cross products in the state then scale with the
  // data's spread around the shift rather than with its offset from zero,
  // and the outputs (variances, covariance, correlation) are unchanged by a
  // common shift. For float32 data x - ka is exact in float64.
  double ka = 0.0, kb = 0.0;
  if constexpr (KIND == kEVar || KIND == kEStd || KIND == kECov || KIND == kECorr) {
    const int ln = lane_id();
    const double va = ln < n ? (double)ra[ln] : 0.0;
    const unsigned long long ma = __ballot(ln < n && isfinite(va));
    if (ma != 0ull) ka = read_lane(va, __ffsll((long long)ma) - 1);
    if constexpr (Sp::NARGS == 2) {
      const double vb = ln < n ? (double)rb[ln] : 0.0;
      const unsigned long long mb = __ballot(ln < n && isfinite(vb));
      if (mb != 0ull) kb = read_lane(vb, __ffsll((long long)mb) - 1);
    }
  }
  // SCALAR: one alpha for every position (kernel-uniform decay factors).
  const double alpha_s = (double)(TA)A.alpha_scalar;
  auto alpha_at = [&](int64_t p) -> double {
    if constexpr (SCALAR) return alpha_s;
    else return (double)rl[p];
  };

  // 1. this thread's map.
  Map<NS> m;
  {
    double s[NS];
#pragma unroll
    for (int c = 0; c < NS; ++c) s[c] = 0.0;
    int seen = 0;
    double p = 1.0, p2 = 1.0;
    auto step = [&](int j) {
      const int64_t pos = t0 + e0 + j;
      const double alpha = alpha_at(pos);
      const double d = 1.0 - alpha;
      p *= d;
      if (Sp::SW2 >= 0) p2 *= d * d;
      const double a = (double)la[pad16(e0 + j)] - ka;
      const double b = (Sp::NARGS == 2) ? (double)lb[pad16(e0 + j)] - kb : 0.0;
      exp_step<KIND>(s, seen, a, b, alpha);
    };
    if (full) {
#pragma unroll
      for (int j = 0; j < K; ++j) step(j);  // no per-position bounds checks
    } else {
      for (int j = 0; j < K && t0 + e0 + j < n; ++j) step(j);
    }
    m.p = p;
    m.p2 = p2;
#pragma unroll
    for (int c = 0; c < NS; ++c) m.b[c] = s[c];
    m.seen = seen;
  }

  // 2. exclusive block scan of maps (composition in position order).
  const int lane = lane_id(), wave = wave_id();
  Map<NS> incl, excl;
  wave_scan_maps<NS, Sp::SW2>(m, incl, excl);
  if (lane == kWave - 1) s_wave[wave] = incl;
  __syncthreads();
  if constexpr (MODE == 1) {
    if (tid == 0) {  // the tile's aggregate map
      Map<NS> agg = Map<NS>::identity();
      for (int w = 0; w < kExpWaves; ++w) agg = compose<NS, Sp::SW2>(agg, s_wave[w]);
      double* pl = A.payload + (row * A.tiles_per_row + tile) * kPayload;
      pl[0] = agg.p;
      pl[1] = agg.p2;
#pragma unroll
      for (int c = 0; c < NS; ++c) pl[2 + c] = agg.b[c];
      pl[2 + NS] = (double)agg.seen;
    }
    return;
  }
  Map<NS> wprefix = Map<NS>::identity();
  for (int w = 0; w < wave; ++w) wprefix = compose<NS, Sp::SW2>(wprefix, s_wave[w]);
  excl = compose<NS, Sp::SW2>(wprefix, excl);

  // 3. decoupled look-back for the tile's incoming state, by wave 0: each
  //    pass reads the flags of the 64 tiles before the latest unconsumed one
  //    at once (later tiles in higher lanes), composes the run of published
  //    aggregates above the highest inclusive state or unpublished tile with
  //    one wave scan, and stops at the inclusive state (a zero state before
  //    the row start). Used for the small maps of count/mean/sum; the
  //    moment kinds' larger maps make the wave scan cost more registers and
  //    time than the walk it saves.
  constexpr bool WAVE_LOOKBACK = NS <= 3;
  if (MODE == 0 && WAVE_LOOKBACK && wave == 0) {
    Map<NS> agg = Map<NS>::identity();
    for (int w = 0; w < kExpWaves; ++w) agg = compose<NS, Sp::SW2>(agg, s_wave[w]);
    unsigned int* frow = A.flags + row * A.tiles_per_row;
    double* prow = A.payload + row * A.tiles_per_row * kPayload;
    double sin[NS];
#pragma unroll
    for (int c = 0; c < NS; ++c) sin[c] = 0.0;
    int sin_seen = 0;
    if (tile > 0) {
      if (lane == 0) {  // publish the aggregate map
        double* pl = prow + tile * kPayload;
        st_agent_f64(pl + 0, agg.p);
        st_agent_f64(pl + 1, agg.p2);
#pragma unroll
        for (int c = 0; c < NS; ++c) st_agent_f64(pl + 2 + c, agg.b[c]);
        st_agent_f64(pl + 2 + NS, (double)agg.seen);
        st_flag(&frow[tile], 1u);
      }
      Map<NS> acc = Map<NS>::identity();  // maps of the tiles (top, tile) composed
      int64_t top = tile - 1;              // latest tile not yet composed
      unsigned int spins = 0;
      while (true) {
        const int64_t j = top - (kWave - 1) + lane;
        const unsigned int f = j >= 0 ? ld_flag(&frow[j]) : 2u;
        const unsigned long long ready = __ballot(f != 0u);
        const unsigned long long incl = __ballot(f == 2u);
        const int hnr = ready == ~0ull ? -1 : 63 - __clzll((long long)~ready);  // highest unpublished
        const int hp = incl == 0ull ? -1 : 63 - __clzll((long long)incl);      // highest inclusive
        if (hnr == kWave - 1) {  // the latest tile has not published yet
          spin_pause();
          if (++spins > kSpinLimit) {
            if (lane == 0) atomicAdd(A.timeouts, 1u);
            acc = Map<NS>::identity();  // expired wait (counted): continue from zero state
            break;
          }
          continue;
        }
        // lanes above lo hold published aggregates: compose them in tile order
        const int lo = hp > hnr ? hp : hnr;
        Map<NS> mj = Map<NS>::identity();
        if (lane > lo) {
          const double* pj = prow + j * kPayload;
          mj.p = ld_agent_f64(pj + 0);
          mj.p2 = ld_agent_f64(pj + 1);
#pragma unroll
          for (int c = 0; c < NS; ++c) mj.b[c] = ld_agent_f64(pj + 2 + c);
          mj.seen = (int)ld_agent_f64(pj + 2 + NS);
        }
        Map<NS> run, unused;
        wave_scan_maps<NS, Sp::SW2>(mj, run, unused);
        acc = compose<NS, Sp::SW2>(read_lane_map(run, kWave - 1), acc);
        if (hp > hnr) {  // inclusive state of tile top - 63 + hp (zero before the row)
          const int64_t jp = top - (kWave - 1) + hp;
          double sj[NS];
          int seen_j = 0;
#pragma unroll
          for (int c = 0; c < NS; ++c) sj[c] = 0.0;
          if (jp >= 0) {
            const double* pj = prow + jp * kPayload;
#pragma unroll
            for (int c = 0; c < NS; ++c) sj[c] = ld_agent_f64(pj + kMaxNS + 3 + c);
            seen_j = (int)ld_agent_f64(pj + kMaxNS + 3 + NS);
          }
#pragma unroll
          for (int c = 0; c < NS; ++c) sin[c] = (c == Sp::SW2 ? acc.p2 : acc.p) * sj[c] + acc.b[c];
          sin_seen = min(2, seen_j + acc.seen);
          break;
        }
        top -= (kWave - 1) - hnr;  // lanes (hnr, 63] composed
      }
    }
    if (lane == 0) {  // publish the inclusive state
      double* pl = prow + tile * kPayload;
#pragma unroll
      for (int c = 0; c < NS; ++c)
        st_agent_f64(pl + kMaxNS + 3 + c, (c == Sp::SW2 ? agg.p2 : agg.p) * sin[c] + agg.b[c]);
      st_agent_f64(pl + kMaxNS + 3 + NS, (double)min(2, sin_seen + agg.seen));
      st_flag(&frow[tile], 2u);
#pragma unroll
      for (int c = 0; c < NS; ++c) s_in[c] = sin[c];
      s_in_seen = sin_seen;
    }
  }
  // 3b. the moment kinds (larger maps): serial look-back by thread 0, one
  //     predecessor at a time.
  if (MODE == 0 && !WAVE_LOOKBACK && tid == 0) {
    Map<NS> agg = Map<NS>::identity();
    for (int w = 0; w < kExpWaves; ++w) agg = compose<NS, Sp::SW2>(agg, s_wave[w]);
    unsigned int* frow = A.flags + row * A.tiles_per_row;
    double* prow = A.payload + row * A.tiles_per_row * kPayload;
    double sin[NS];
#pragma unroll
    for (int c = 0; c < NS; ++c) sin[c] = 0.0;
    int sin_seen = 0;
    if (tile > 0) {
      // publish the aggregate map
      double* pl = prow + tile * kPayload;
      st_agent_f64(pl + 0, agg.p);
      st_agent_f64(pl + 1, agg.p2);
#pragma unroll
      for (int c = 0; c < NS; ++c) st_agent_f64(pl + 2 + c, agg.b[c]);
      st_agent_f64(pl + 2 + NS, (double)agg.seen);
      st_flag(&frow[tile], 1u);
      // walk back, composing aggregates until an inclusive state is found
      Map<NS> acc = Map<NS>::identity();  // maps of tiles (j, tile) composed
      int64_t j = tile - 1;
      while (true) {
        unsigned int f;
        unsigned int spins = 0;
        while ((f = ld_flag(&frow[j])) == 0u) {
          spin_pause();
          if (++spins > kSpinLimit) {
            atomicAdd(A.timeouts, 1u);
            break;
          }
        }
        const double* pj = prow + j * kPayload;
        if (f == 2u) {
          double sj[NS];
#pragma unroll
          for (int c = 0; c < NS; ++c) sj[c] = ld_agent_f64(pj + kMaxNS + 3 + c);
          const int seen_j = (int)ld_agent_f64(pj + kMaxNS + 3 + NS);
#pragma unroll
          for (int c = 0; c < NS; ++c) sin[c] = (c == Sp::SW2 ? acc.p2 : acc.p) * sj[c] + acc.b[c];
          sin_seen = min(2, seen_j + acc.seen);
          break;
        }
        if (f == 0u) break;  // expired wait (counted above): continue from zero state
        Map<NS> mj;
        mj.p = ld_agent_f64(pj + 0);
        mj.p2 = ld_agent_f64(pj + 1);
#pragma unroll
        for (int c = 0; c < NS; ++c) mj.b[c] = ld_agent_f64(pj + 2 + c);
        mj.seen = (int)ld_agent_f64(pj + 2 + NS);
        acc = compose<NS, Sp::SW2>(mj, acc);
        if (j == 0) {
#pragma unroll
          for (int c = 0; c < NS; ++c) sin[c] = acc.b[c];
          sin_seen = acc.seen;
          break;
        }
        --j;
      }
    }
    // publish the inclusive state
    double* pl = prow + tile * kPayload;
#pragma unroll
    for (int c = 0; c < NS; ++c)
      st_agent_f64(pl + kMaxNS + 3 + c, (c == Sp::SW2 ? agg.p2 : agg.p) * sin[c] + agg.b[c]);
    st_agent_f64(pl + kMaxNS + 3 + NS, (double)min(2, sin_seen + agg.seen));
    st_flag(&frow[tile], 2u);
#pragma unroll
    for (int c = 0; c < NS; ++c) s_in[c] = sin[c];
    s_in_seen = sin_seen;
  }
  __syncthreads();

  // 4. re-run this thread's positions from its incoming state.
  double sin_t[NS];
  int sin_seen_t;
  if constexpr (MODE == 2) {  // computed by exp_carry_kernel
    const double* pl = A.payload + (row * A.tiles_per_row + tile) * kPayload + kMaxNS + 3;
#pragma unroll
    for (int c = 0; c < NS; ++c) sin_t[c] = pl[c];
    sin_seen_t = (int)pl[NS];
  } else {
#pragma unroll
    for (int c = 0; c < NS; ++c) sin_t[c] = s_in[c];
    sin_seen_t = s_in_seen;
  }
  double s[NS];
#pragma unroll
  for (int c = 0; c < NS; ++c) s[c] = (c == Sp::SW2 ? excl.p2 : excl.p) * sin_t[c] + excl.b[c];
  int seen = min(2, sin_seen_t + excl.seen);
  TA res[K];
  auto rerun = [&](int j) {
    const int64_t pos = t0 + e0 + j;
    const double alpha = alpha_at(pos);
    const double a = (double)la[pad16(e0 + j)] - ka;
    const double b = (Sp::NARGS == 2) ? (double)lb[pad16(e0 + j)] - kb : 0.0;
    exp_step<KIND>(s, seen, a, b, alpha);
    res[j] = (TA)exp_out<KIND>(s, seen, A.min_weight);
  };
  if (full) {
#pragma unroll
    for (int j = 0; j < K; ++j) rerun(j);
  } else {
    for (int j = 0; j < K && t0 + e0 + j < n; ++j) rerun(j);
  }
  constexpr int OV = 16 / (int)sizeof(TA);  // outputs per 16-byte store
  // A full tile's outputs go through LDS (the staged inputs' region, free once
  // every thread has re-run its positions), so that each store instruction
  // writes consecutive 16-byte pieces across the wave instead of one piece
  // per K * sizeof(TA) bytes. Rounds of XR threads' outputs fit the region;
  // a thread's K outputs sit at stride K + 1 (conflict-free LDS writes).
  constexpr int64_t kRegion = ((int64_t)TILE + TILE / 16 + 16) * (int64_t)sizeof(TX) * Sp::NARGS;  // pad16(TILE) + 16 slots
  constexpr int kFit = (int)(kRegion / ((K + 1) * (int64_t)sizeof(TA)));
  constexpr int XR = kFit >= kExpThreads ? kExpThreads : (kFit >= kExpThreads / 2 ? kExpThreads / 2 : kExpThreads / 4);
  static_assert((int64_t)XR * (K + 1) * (int64_t)sizeof(TA) <= kRegion, "output round fits the staged region");
  static_assert(K % OV == 0, "whole 16-byte pieces per thread");
  if (full && ((uintptr_t)(ro + t0)) % 16 == 0) {
    TA* lo = reinterpret_cast<TA*>(smem_raw);
    char* rot = reinterpret_cast<char*>(ro + t0);
    struct alignas(16) V { TA v[OV]; };
    __syncthreads();  // every re-run read of the staged inputs is done
#pragma unroll
    for (int r = 0; r < kExpThreads / XR; ++r) {
      if (r > 0) __syncthreads();  // the previous round's reads are done
      if (tid / XR == r) {
        const int b = (tid - r * XR) * (K + 1);
#pragma unroll
        for (int j = 0; j < K; ++j) lo[b + j] = res[j];
      }
      __syncthreads();
      // pieces of OV outputs: piece q of the round is thread q / (K / OV)'s
      // outputs (q % (K / OV)) * OV ..
      constexpr int kPieces = XR * K / OV;
#pragma unroll
      for (int q = tid; q < kPieces; q += kExpThreads) {
        const int th = q / (K / OV), jo = (q % (K / OV)) * OV;
        V vo;
#pragma unroll
        for (int v = 0; v < OV; ++v) vo.v[v] = lo[th * (K + 1) + jo + v];
        store_stream16(reinterpret_cast<TA*>(rot + ((uint32_t)(r * XR * K + q * OV) * (uint32_t)sizeof(TA))), vo.v);
      }
    }
  } else if (full && ((uintptr_t)(ro + t0 + e0)) % 16 == 0) {
    struct alignas(16) V { TA v[OV]; };
#pragma unroll
    for (int j = 0; j < K; j += OV) {
      V vo;
#pragma unroll
      for (int v = 0; v < OV; ++v) vo.v[v] = res[j + v];
      *reinterpret_cast<V*>(ro + t0 + e0 + j) = vo;
    }
  } else if (full) {
#pragma unroll
    for (int j = 0; j < K; ++j) ro[t0 + e0 + j] = res[j];
  } else {
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (t0 + e0 + j < n) ro[t0 + e0 + j] = res[j];
  }
}

// Two-pass middle step: one workgroup per row scans the aggregate maps of the
// row's tiles (256 at a time, DPP wave scans and a running carry) and stores
// each tile's incoming state, the composition of all earlier tiles' maps
// applied to the zero state before the row start.
template <int KIND>
__global__ __launch_bounds__(kExpThreads) void exp_carry_kernel(ExpArgs A) {
  using Sp = ExpSpec<KIND>;
  constexpr int NS = Sp::NS;
  __shared__ Map<NS> s_wave[kExpWaves];
  const int64_t row = blockIdx.x;
  const int64_t T = A.tiles_per_row;
  double* prow = A.payload + row * T * kPayload;
  const int tid = threadIdx.x, lane = lane_id(), wave = wave_id();
  Map<NS> carry = Map<NS>::identity();
  for (int64_t c0 = 0; c0 < T; c0 += kExpThreads) {
    const int64_t t = c0 + tid;
    Map<NS> m = Map<NS>::identity();
    if (t < T) {
      const double* pl = prow + t * kPayload;
      m.p = pl[0];
      m.p2 = pl[1];
#pragma unroll
      for (int c = 0; c < NS; ++c) m.b[c] = pl[2 + c];
      m.seen = (int)pl[2 + NS];
    }
    Map<NS> incl, excl;
    wave_scan_maps<NS, Sp::SW2>(m, incl, excl);
    if (lane == kWave - 1) s_wave[wave] = incl;
    __syncthreads();
    Map<NS> wprefix = Map<NS>::identity(), total = Map<NS>::identity();
    for (int w = 0; w < kExpWaves; ++w) {
      if (w < wave) wprefix = compose<NS, Sp::SW2>(wprefix, s_wave[w]);
      total = compose<NS, Sp::SW2>(total, s_wave[w]);
    }
    const Map<NS> in = compose<NS, Sp::SW2>(carry, compose<NS, Sp::SW2>(wprefix, excl));
    if (t < T) {
      double* pl = prow + t * kPayload + kMaxNS + 3;
#pragma unroll
      for (int c = 0; c < NS; ++c) pl[c] = in.b[c];
      pl[NS] = (double)in.seen;
    }
    carry = compose<NS, Sp::SW2>(carry, total);
    __syncthreads();  // s_wave is rewritten by the next chunk
  }
}

template <typename TX, int KIND>
constexpr int64_t exp_tile() {
  return kExpThreads * (int64_t)exp_items<TX, KIND>();
}
// The smallest tile of any kind (the workspace is sized for it).
template <typename TX>
constexpr int64_t exp_min_tile() {
  return kExpThreads * (64 / (int64_t)sizeof(TX));
}

// Two passes on request only (NBG_EXP_TWO_PASS=1; =0 or unset: one pass).
// Measured on MI355X for float32 (256, 2M), alpha 0.01: with 16 positions per
// thread sum / count took 1.99 / 2.00 ms in one pass and 1.78 / 1.75 ms in
// two, but with 32 positions one pass takes 1.62 / 1.61 ms against 1.89 /
// 1.85 ms in two; mean 1.77 ms in one pass, 2.12 in two; the moment kinds are
// bound by their per-tile arithmetic, which two passes repeat (cov / corr
// 2.73 / 3.17 ms in one pass, 3.12 / 3.38 in two).
template <int KIND>
bool exp_two_pass(int64_t tiles) {
  (void)tiles;
  return knob("NBG_EXP_TWO_PASS", 0) == 1;
}

template <typename TX, typename TA, int KIND>
int exp_launch(ExpArgs a, void* ws, size_t ws_bytes, hipStream_t s) {
  const int64_t tiles = (a.n + exp_tile<TX, KIND>() - 1) / exp_tile<TX, KIND>();
  const int64_t blocks = a.rows * tiles;
  if (blocks > 0x7fffffffLL) return set_error(NBG_ERR_INVALID, "move_exp: too many tiles");
  const size_t need = nbg_move_exp_workspace_bytes(a.rows, a.n, sizeof(TX) == 4 ? NBG_F32 : NBG_F64);
  if (ws == nullptr || ws_bytes < need)
    return set_error(NBG_ERR_INVALID, "move_exp: workspace too small (%zu < %zu)", ws_bytes, need);
  // layout: flags[rows*tiles] u32 | counters[rows] u32 | (8-aligned) payload
  const size_t flag_bytes = (size_t)(a.rows * tiles + a.rows) * 4;
  const size_t zero_bytes = (flag_bytes + 255) / 256 * 256;
  a.flags = static_cast<unsigned int*>(ws);
  a.counters = a.flags + a.rows * tiles;
  a.payload = reinterpret_cast<double*>(static_cast<char*>(ws) + zero_bytes);
  a.tiles_per_row = tiles;
  a.timeouts = spin_counter_ptr();
  if (a.timeouts == nullptr) return set_error(NBG_ERR_HIP, "move_exp: cannot resolve device counter");
  constexpr int NARGS = ExpSpec<KIND>::NARGS;
  const size_t lds = (size_t)(pad16(exp_tile<TX, KIND>()) + 16) * sizeof(TX) * NARGS;
  if (exp_two_pass<KIND>(tiles)) {
    if (a.alpha == nullptr) {
      hipLaunchKernelGGL((exp_kernel<TX, TA, KIND, true, 1>), dim3((unsigned)blocks), dim3(kExpThreads), lds, s, a);
      hipLaunchKernelGGL((exp_carry_kernel<KIND>), dim3((unsigned)a.rows), dim3(kExpThreads), 0, s, a);
      hipLaunchKernelGGL((exp_kernel<TX, TA, KIND, true, 2>), dim3((unsigned)blocks), dim3(kExpThreads), lds, s, a);
    } else {
      hipLaunchKernelGGL((exp_kernel<TX, TA, KIND, false, 1>), dim3((unsigned)blocks), dim3(kExpThreads), lds, s, a);
      hipLaunchKernelGGL((exp_carry_kernel<KIND>), dim3((unsigned)a.rows), dim3(kExpThreads), 0, s, a);
      hipLaunchKernelGGL((exp_kernel<TX, TA, KIND, false, 2>), dim3((unsigned)blocks), dim3(kExpThreads), lds, s, a);
    }
    return check_hip(hipGetLastError(), "exp_kernel launch");
  }
  if (int rc = launch_zero(ws, zero_bytes, s)) return rc;
  if (a.alpha == nullptr)
    hipLaunchKernelGGL((exp_kernel<TX, TA, KIND, true>), dim3((unsigned)blocks), dim3(kExpThreads), lds, s, a);
  else
    hipLaunchKernelGGL((exp_kernel<TX, TA, KIND, false>), dim3((unsigned)blocks), dim3(kExpThreads), lds, s, a);
  return check_hip(hipGetLastError(), "exp_kernel launch");
}

template <int KIND>
int exp_entry(const char* name, int x_dtype, int dtype, const void* xa, const void* xb,
              const void* alpha, int64_t ld_alpha, double alpha_scalar, void* out, int64_t rows,
              int64_t n, int64_t ld_a, int64_t ld_b, int64_t ld_out, double min_weight, void* ws,
              size_t ws_bytes, void* stream) {
  if (int rc = injected_error(name)) return rc;
  if (rows < 0 || n < 0) return set_error(NBG_ERR_INVALID, "%s: negative shape", name);
  if (rows == 0 || n == 0) return NBG_OK;
  constexpr int NARGS = ExpSpec<KIND>::NARGS;
  if (xa == nullptr || out == nullptr || (NARGS == 2 && xb == nullptr))
    return set_error(NBG_ERR_INVALID, "%s: null pointer", name);
  if (rows > 1 && (ld_a < n || ld_out < n || (NARGS == 2 && ld_b < n)))
    return set_error(NBG_ERR_INVALID, "%s: row stride < n", name);
  if (alpha != nullptr && ld_alpha != 0 && ld_alpha < n)
    return set_error(NBG_ERR_INVALID, "%s: alpha row stride must be 0 or >= n", name);
  ExpArgs a{};
  a.xa = xa;
  a.xb = xb;
  a.alpha = alpha;
  a.ld_alpha = ld_alpha;
  a.alpha_scalar = alpha_scalar;
  a.out = out;
  a.rows = rows;
  a.n = n;
  a.ld_a = ld_a;
  a.ld_b = ld_b;
  a.ld_out = ld_out;
  a.min_weight = min_weight;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (x_dtype == NBG_F32 && dtype == NBG_F32) return exp_launch<float, float, KIND>(a, ws, ws_bytes, s);
  if (x_dtype == NBG_F64 && dtype == NBG_F64) return exp_launch<double, double, KIND>(a, ws, ws_bytes, s);
  if (x_dtype == NBG_F32 && dtype == NBG_F64) return exp_launch<float, double, KIND>(a, ws, ws_bytes, s);
  return set_error(NBG_ERR_INVALID, "%s: unsupported dtype pair (%d, %d)", name, x_dtype, dtype);
}

}  // namespace
}  // namespace nbg

extern "C" {

size_t nbg_move_exp_workspace_bytes(int64_t rows, int64_t n, int x_dtype) {
  const int64_t tile = (x_dtype == NBG_F64) ? nbg::exp_min_tile<double>() : nbg::exp_min_tile<float>();
  const int64_t tiles = (n + tile - 1) / tile;
  if (rows <= 0 || tiles <= 0) return 0;
  const size_t flag_bytes = (size_t)(rows * tiles + rows) * 4;
  const size_t zero_bytes = (flag_bytes + 255) / 256 * 256;
  return zero_bytes + (size_t)(rows * tiles) * nbg::kPayload * sizeof(double);
}

#define NBG_EXP1(NAME, KIND)                                                                      \
  int NAME(int x_dtype, int dtype, const void* x, const void* alpha, int64_t ld_alpha,           \
           double alpha_scalar, void* out, int64_t rows, int64_t n, int64_t ld_x, int64_t ld_out, \
           double min_weight, void* ws, size_t ws_bytes, void* stream) {                          \
    return nbg::exp_entry<KIND>(#NAME, x_dtype, dtype, x, nullptr, alpha, ld_alpha, alpha_scalar, \
                                out, rows, n, ld_x, 0, ld_out, min_weight, ws, ws_bytes, stream); \
  }
#define NBG_EXP2(NAME, KIND)                                                                     \
  int NAME(int x_dtype, int dtype, const void* a, const void* b, const void* alpha,             \
           int64_t ld_alpha, double alpha_scalar, void* out, int64_t rows, int64_t n,           \
           int64_t ld_a, int64_t ld_b, int64_t ld_out, double min_weight, void* ws,             \
           size_t ws_bytes, void* stream) {                                                      \
    return nbg::exp_entry<KIND>(#NAME, x_dtype, dtype, a, b, alpha, ld_alpha, alpha_scalar, out, \
                                rows, n, ld_a, ld_b, ld_out, min_weight, ws, ws_bytes, stream);  \
  }

NBG_EXP1(nbg_move_exp_nancount, nbg::kECount)
NBG_EXP1(nbg_move_exp_nanmean, nbg::kEMean)
NBG_EXP1(nbg_move_exp_nansum, nbg::kESum)
NBG_EXP1(nbg_move_exp_nanvar, nbg::kEVar)
NBG_EXP1(nbg_move_exp_nanstd, nbg::kEStd)
NBG_EXP2(nbg_move_exp_nancov, nbg::kECov)
NBG_EXP2(nbg_move_exp_nancorr, nbg::kECorr)

}  // extern "C"
