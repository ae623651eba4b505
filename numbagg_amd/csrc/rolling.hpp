// Rolling-window tile machinery shared by the plain (moving.hip) and the
// ragged batched (ragged.hip) kernels: window statistics accumulators and the
// per-tile evaluation `rolling_tile`.
#pragma once

#include <cstdlib>
#include <type_traits>

#include "common.hpp"

namespace nbg {
namespace {

// Threads per rolling tile (a build-time choice: NBG_MOVE_THREADS).
#ifndef NBG_MOVE_THREADS
#define NBG_MOVE_THREADS 256
#endif
constexpr int kMoveThreads = NBG_MOVE_THREADS;
static_assert(kMoveThreads % 64 == 0 && kMoveThreads >= 64 && kMoveThreads <= 1024, "whole wavefronts");

enum MoveKind { kSum = 0, kMean = 1, kVar = 2, kStd = 3, kCov = 4, kCorr = 5 };

// ---- scalar accumulators ---------------------------------------------------
// Plain float64 accumulator with the same interface as Comp<F>.
struct D {
  double v;
  __device__ __forceinline__ static D zero() { return D{0.0}; }
  __device__ __forceinline__ void add(double x) { v += x; }
  __device__ __forceinline__ void add(const D& o) { v += o.v; }
  __device__ __forceinline__ void sub(const D& o) { v -= o.v; }
  __device__ __forceinline__ double to_double() const { return v; }
};

// An accumulator moved across lanes by DPP (zero where the move writes nothing).
template <int CTRL, int RM>
__device__ __forceinline__ D dpp_acc(const D& a) {
  return D{dpp_move<CTRL, RM>(0.0, a.v)};
}
template <int CTRL, int RM, typename F>
__device__ __forceinline__ Comp<F> dpp_acc(const Comp<F>& a) {
  return Comp<F>{dpp_move<CTRL, RM>(F(0), a.hi), dpp_move<CTRL, RM>(F(0), a.lo)};
}

// a*b added to an accumulator: exact product via fma for double-double.
__device__ __forceinline__ void add_prod(D& acc, double a, double b) { acc.v += a * b; }
__device__ __forceinline__ void add_prod(Comp<double>& acc, double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  acc.add(p);
  acc.add(e);
}
__device__ __forceinline__ void sub_prod(D& acc, double a, double b) { acc.v -= a * b; }
__device__ __forceinline__ void sub_prod(Comp<double>& acc, double a, double b) {
  double p = a * b;
  double e = fma(a, b, -p);
  acc.add(-p);
  acc.add(-e);
}

// A float32 value with NaN replaced by 0, selected in float32 and kept opaque
// to the optimiser, so that a later widening to float64 is not pushed ahead of
// the select (which would then act on both halves of the float64 value).
__device__ __forceinline__ float nan_to_zero(float v) {
  float r = is_nan(v) ? 0.0f : v;
  __asm__("" : "+v"(r));
  return r;
}

// ---- per-kind window statistics ---------------------------------------------
template <typename T, int KIND>
struct Traits;

// Exact difference head - tail as an accumulator value: in float64 the
// difference of two float32 values is exact; TwoSum pairs carry it exactly.
template <typename A, typename T>
__device__ __forceinline__ A exact_diff(T h, T t);
template <>
__device__ __forceinline__ D exact_diff<D, float>(float h, float t) {
  return D{(double)h - (double)t};
}
template <>
__device__ __forceinline__ Comp<float> exact_diff<Comp<float>, float>(float h, float t) {
  Comp<float> r;
  two_sum(h, -t, r.hi, r.lo);
  return r;
}
template <>
__device__ __forceinline__ Comp<double> exact_diff<Comp<double>, double>(double h, double t) {
  Comp<double> r;
  two_sum(h, -t, r.hi, r.lo);
  return r;
}

// First-moment-only kinds (sum, mean). A: float64 (float32 data, default),
// FP32 TwoSum pair (float32 data, selectable), double-double (float64 data).
template <typename T, typename A>
struct SumStats {
  using Acc = A;
  int c;
  A s;
  __device__ __forceinline__ static SumStats zero() { return SumStats{0, A::zero()}; }
  __device__ __forceinline__ void add_obs(T a, T, T, T) {
    c += 1;
    s.add(a);
  }
  __device__ __forceinline__ void sub_obs(T a, T, T, T) {
    c -= 1;
    s.add(-a);
  }
  __device__ __forceinline__ void merge(const SumStats& o) {
    c += o.c;
    s.add(o.s);
  }
  __device__ __forceinline__ void unmerge(const SumStats& o) {
    c -= o.c;
    s.sub(o.s);
  }
  template <int CTRL, int RM>
  __device__ __forceinline__ SumStats dpp() const {
    return SumStats{dpp_move<CTRL, RM>(0, c), dpp_acc<CTRL, RM>(s)};
  }
};

// Moment kinds on shifted data: y = x - k. NM selects the sums kept:
// 1 = (s1, s11) for var/std; 2 = (s1, s2, s12, s11, s22) for corr;
// 3 = (s1, s2, s12) for cov.
template <typename T, int NM>
struct MomStats {
  using A = typename std::conditional<sizeof(T) == 4, D, Comp<double>>::type;
  using Acc = void;  // (no first-moment fast path)
  static constexpr bool B = NM >= 2;     // second input
  static constexpr bool SQA = NM <= 2;   // sum of squares of the first
  static constexpr bool SQB = NM == 2;   // sum of squares of the second
  int c;
  A s1, s2, s11, s22, s12;
  __device__ __forceinline__ static MomStats zero() {
    return MomStats{0, A::zero(), A::zero(), A::zero(), A::zero(), A::zero()};
  }
  // a, b: observations; ka, kb: shifts.
  __device__ __forceinline__ void add_obs(T a, T b, T ka, T kb) {
    const double ya = (double)a - (double)ka;
    c += 1;
    s1.add(ya);
    if (SQA) add_prod(s11, ya, ya);
    if (B) {
      const double yb = (double)b - (double)kb;
      s2.add(yb);
      add_prod(s12, ya, yb);
      if (SQB) add_prod(s22, yb, yb);
    }
  }
  __device__ __forceinline__ void sub_obs(T a, T b, T ka, T kb) {
    const double ya = (double)a - (double)ka;
    c -= 1;
    s1.add(-ya);
    if (SQA) sub_prod(s11, ya, ya);
    if (B) {
      const double yb = (double)b - (double)kb;
      s2.add(-yb);
      sub_prod(s12, ya, yb);
      if (SQB) sub_prod(s22, yb, yb);
    }
  }
  __device__ __forceinline__ void merge(const MomStats& o) {
    c += o.c;
    s1.add(o.s1);
    if (SQA) s11.add(o.s11);
    if (B) {
      s2.add(o.s2);
      s12.add(o.s12);
      if (SQB) s22.add(o.s22);
    }
  }
  __device__ __forceinline__ void unmerge(const MomStats& o) {
    c -= o.c;
    s1.sub(o.s1);
    if (SQA) s11.sub(o.s11);
    if (B) {
      s2.sub(o.s2);
      s12.sub(o.s12);
      if (SQB) s22.sub(o.s22);
    }
  }
  template <int CTRL, int RM>
  __device__ __forceinline__ MomStats dpp() const {
    MomStats r = zero();
    r.c = dpp_move<CTRL, RM>(0, c);
    r.s1 = dpp_acc<CTRL, RM>(s1);
    if (SQA) r.s11 = dpp_acc<CTRL, RM>(s11);
    if (B) {
      r.s2 = dpp_acc<CTRL, RM>(s2);
      r.s12 = dpp_acc<CTRL, RM>(s12);
      if (SQB) r.s22 = dpp_acc<CTRL, RM>(s22);
    }
    return r;
  }
};

// 1/c for a positive count: hardware reciprocal plus one Newton step (about
// 1 ULP of float64, far below the float32 rounding of the result).
__device__ __forceinline__ double recip(int c) {
  const double dc = (double)c;
  double r = __builtin_amdgcn_rcp(dc);
  const double e = fma(-dc, r, 1.0);
  return fma(r, e, r);
}

__device__ __forceinline__ float mean_of(const D& s, int c) { return (float)(s.v * recip(c)); }
__device__ __forceinline__ float mean_of(const Comp<float>& s, int c) {
  return (float)(s.to_double() * recip(c));
}
__device__ __forceinline__ double mean_of(const Comp<double>& s, int c) {
  const double dc = (double)c;
  const double q = s.hi / dc;
  const double r = fma(-q, dc, s.hi) + s.lo;  // exact remainder of hi, plus lo
  return q + r / dc;
}

// ACC32: float32 data accumulates in FP32 TwoSum pairs instead of float64.
template <typename T, bool ACC32>
using SumAccT = typename std::conditional<
    sizeof(T) == 8, Comp<double>, typename std::conditional<ACC32, Comp<float>, D>::type>::type;

template <typename T, int KIND, bool ACC32 = false>
struct SumTraits {
  using S = SumStats<T, SumAccT<T, ACC32>>;
  static constexpr int NARGS = 1;
  static constexpr bool SHIFT = false;
  __device__ static T out(const S& s, int64_t mc, const double* lut = nullptr) {
    if (KIND == kSum) {
      if (s.c < mc) return canonical_nan<T>();
      return (T)s.s.to_double();
    }
    if (s.c < (mc > 1 ? mc : 1)) return canonical_nan<T>();
    if constexpr (std::is_same<typename S::Acc, D>::value) {
      if (lut != nullptr) return (T)(s.s.v * lut[s.c]);  // correctly rounded 1/c from LDS
    }
    return (T)mean_of(s.s, s.c);
  }
};
template <typename T>
struct Traits<T, kSum> : SumTraits<T, kSum> {};
template <typename T>
struct Traits<T, kMean> : SumTraits<T, kMean> {};
// Variants with FP32 TwoSum accumulators for float32 data.
constexpr int kSum32 = 16 + kSum;
constexpr int kMean32 = 16 + kMean;
template <typename T>
struct Traits<T, kSum32> : SumTraits<T, kSum, true> {};
template <typename T>
struct Traits<T, kMean32> : SumTraits<T, kMean, true> {};
// Reciprocals 1/c and 1/(c - 1) of a window count for the float32 moment
// outputs (c >= 0; the count-1 guards keep both finite for c < 2, where the
// caller masks the output anyway).
struct CountRecips {
  double r, r1;
};
__device__ __forceinline__ CountRecips count_recips(int c) {
  return CountRecips{recip(c > 0 ? c : 1), recip(c > 1 ? c - 1 : 1)};
}

// 1/sqrt(v) for v in [2**-1000, 2**1000]: the hardware estimate y refined by
// one Newton step y * (1.5 - v y^2 / 2), which squares its relative error.
__device__ __forceinline__ double rsqrt_refined(double v) {
  const double y = __builtin_amdgcn_rsq(v);
  return y * fma(-(0.5 * v) * y, y, 1.5);
}

// sqrt(v) for a result rounded to float32, without the float64 square-root
// expansion: s = v * rsqrt(v) plus one Newton correction (relative error
// about the square of the estimate's). Above 2**300 the float32 result is
// infinite and below 2**-300 it is zero either way, so v is clamped there;
// NaN and negative v give NaN, as sqrt does.
__device__ __forceinline__ double sqrt_for_f32(double v) {
  const double vc = v > 0x1p300 ? 0x1p300 : v;
  const bool zero = vc >= 0.0 && vc < 0x1p-300;
  const double vz = zero ? 1.0 : vc;
  const double y = __builtin_amdgcn_rsq(vz);
  const double sq = vz * y;
  const double res = fma(fma(-sq, sq, vz), 0.5 * y, sq);
  return zero ? 0.0 : res;
}

// float32 moment outputs are formed in float64 with 1/c and 1/(c - 1) from
// refined reciprocals (about 1 ULP of float64 each, far below the float32
// rounding of the result) and masked by a select, so they compile without a
// branch or a float64 division; float64 outputs keep the reference's divisions.
template <typename T, int KIND>
struct VarTraits {
  using S = MomStats<T, 1>;
  static constexpr int NARGS = 1;
  static constexpr bool SHIFT = true;
  __device__ static T out(const S& s, int64_t mc, const double* = nullptr) {
    const int64_t mce = mc > 2 ? mc : 2;
    const double m1 = s.s1.to_double();
    if constexpr (sizeof(T) == 4) {
      const CountRecips q = count_recips(s.c);
      double var = (s.s11.to_double() - m1 * m1 * q.r) * q.r1;
      var = s.c >= mce ? var : canonical_nan<double>();
      return (T)(KIND == kStd ? sqrt_for_f32(var) : var);
    } else {
      if (s.c < mce) return canonical_nan<T>();
      const double c = (double)s.c;
      const double var = (s.s11.to_double() - m1 * m1 / c) / (c - 1.0);
      return (T)(KIND == kStd ? sqrt(var) : var);
    }
  }
  // float32 with the table in LDS: rc[c] holds the refined reciprocal
  // 1/max(c, 1) that count_recips forms, and rc[max(c - 1, 0)] is its
  // 1/max(c - 1, 1), so the results are the same.
  __device__ static T out_tab(const S& s, int64_t mc, const double* rc) {
    const int64_t mce = mc > 2 ? mc : 2;
    const double m1 = s.s1.to_double();
    double var = (s.s11.to_double() - m1 * m1 * rc[s.c]) * rc[s.c > 0 ? s.c - 1 : 0];
    var = s.c >= mce ? var : canonical_nan<double>();
    return (T)(KIND == kStd ? sqrt_for_f32(var) : var);
  }
};
template <typename T>
struct Traits<T, kVar> : VarTraits<T, kVar> {};
template <typename T>
struct Traits<T, kStd> : VarTraits<T, kStd> {};
template <typename T>
struct Traits<T, kCov> {
  using S = MomStats<T, 3>;
  static constexpr int NARGS = 2;
  static constexpr bool SHIFT = true;
  __device__ static T out(const S& s, int64_t mc, const double* = nullptr) {
    const int64_t mce = mc > 2 ? mc : 2;
    if constexpr (sizeof(T) == 4) {
      const CountRecips q = count_recips(s.c);
      const double cov = (s.s12.to_double() - s.s1.to_double() * s.s2.to_double() * q.r) * q.r1;
      return (T)(s.c >= mce ? cov : canonical_nan<double>());
    } else {
      if (s.c < mce) return canonical_nan<T>();
      const double c = (double)s.c;
      return (T)((s.s12.to_double() - s.s1.to_double() * s.s2.to_double() / c) / (c - 1.0));
    }
  }
  // float32 with the reciprocal table (as VarTraits::out_tab)
  __device__ static T out_tab(const S& s, int64_t mc, const double* rc) {
    const int64_t mce = mc > 2 ? mc : 2;
    const double cov = (s.s12.to_double() - s.s1.to_double() * s.s2.to_double() * rc[s.c]) *
                       rc[s.c > 0 ? s.c - 1 : 0];
    return (T)(s.c >= mce ? cov : canonical_nan<double>());
  }
};
template <typename T>
struct Traits<T, kCorr> {
  using S = MomStats<T, 2>;
  static constexpr int NARGS = 2;
  static constexpr bool SHIFT = true;
  // float32: the correlation from r = 1/max(c, 1)
  __device__ static T from_r(const S& s, int64_t mce, double r) {
    const double ma = s.s1.to_double() * r, mb = s.s2.to_double() * r;
    const double va = s.s11.to_double() * r - ma * ma;
    const double vb = s.s22.to_double() * r - mb * mb;
    const double cv = s.s12.to_double() * r - ma * mb;
    const double vv = va * vb;
    const bool in_range = vv > 0x1p-1000 && vv < 0x1p1000;
    double res = cv * rsqrt_refined(in_range ? vv : 1.0);
    if (!in_range && vv > 0) res = cv / sqrt(vv);  // extreme scales only
    return (T)((s.c >= mce && vv > 0) ? res : canonical_nan<double>());
  }
  // float32 with the reciprocal table (the value recip forms)
  __device__ static T out_tab(const S& s, int64_t mc, const double* rc) {
    return from_r(s, mc > 1 ? mc : 1, rc[s.c]);
  }
  __device__ static T out(const S& s, int64_t mc, const double* = nullptr) {
    const int64_t mce = mc > 1 ? mc : 1;
    if constexpr (sizeof(T) == 4) {
      return from_r(s, mce, recip(s.c > 0 ? s.c : 1));
    } else {
      if (s.c < mce) return canonical_nan<T>();
      const double r = 1.0 / (double)s.c;
      const double ma = s.s1.to_double() * r, mb = s.s2.to_double() * r;
      const double va = s.s11.to_double() * r - ma * ma;
      const double vb = s.s22.to_double() * r - mb * mb;
      const double cv = s.s12.to_double() * r - ma * mb;
      const double vv = va * vb;
      return vv > 0 ? (T)(cv / sqrt(vv)) : canonical_nan<T>();
    }
  }
};

// ---- block-level helpers ------------------------------------------------------
constexpr int kMoveWaves = kMoveThreads / kWave;

template <typename S>
struct BlockShared {
  S wave_tot[kMoveWaves];
};

// Inclusive wave scan of S (sum-like merge) with DPP lane moves: row shifts
// by 1, 2, 4 and 8 scan each 16-lane row, then broadcasts of lane 15 (into
// rows 1 and 3) and lane 31 (into rows 2 and 3) carry row totals forward.
// A lane the move does not reach merges the zero statistics.
template <int CTRL, int RM, typename S>
__device__ __forceinline__ S dpp_scan_step(const S& v) {
  S o = v.template dpp<CTRL, RM>();
  o.merge(v);
  return o;
}
template <typename S>
__device__ __forceinline__ S wave_incl_scan(S v) {
  v = dpp_scan_step<0x111, 0xf>(v);
  v = dpp_scan_step<0x112, 0xf>(v);
  v = dpp_scan_step<0x114, 0xf>(v);
  v = dpp_scan_step<0x118, 0xf>(v);
  v = dpp_scan_step<0x142, 0xa>(v);
  v = dpp_scan_step<0x143, 0xc>(v);
  return v;
}

// Exclusive block scan of v fused with the block total of w, behind one
// barrier: wave scans, then the wave totals merged in wave order through
// LDS. The caller must pass a barrier before `sh` or `sh_w` is written again.
template <typename S>
__device__ __forceinline__ S block_excl_scan_and_total(S v, S w, S& w_total, BlockShared<S>& sh,
                                                       BlockShared<S>& sh_w) {
  S incl = wave_incl_scan(v);
  const S w_incl = wave_incl_scan(w);
  if (lane_id() == kWave - 1) {
    sh.wave_tot[wave_id()] = incl;
    sh_w.wave_tot[wave_id()] = w_incl;
  }
  __syncthreads();
  S prefix = S::zero();
  w_total = S::zero();
#pragma unroll
  for (int i = 0; i < kMoveWaves; ++i) {
    if (i < wave_id()) prefix.merge(sh.wave_tot[i]);
    w_total.merge(sh_w.wave_tot[i]);
  }
  S excl = incl;
  excl.unmerge(v);
  prefix.merge(excl);
  return prefix;
}

// 16 bytes of global memory at element alignment: gfx9+ executes dword-aligned
// multi-dword global accesses, so rows at any element offset (series packed
// back to back, sliced views) still move 16 bytes per load/store instruction.
template <typename T>
struct Vec16 {
  T v[16 / sizeof(T)];
};

// ---- the kernel -------------------------------------------------------------------
// Shared-memory scratch of one rolling tile besides the staged segment.
template <typename S>
struct RollingShared {
  BlockShared<S> sh;
  BlockShared<S> sh_w;  // halo totals of the fused scan
};

// Default outputs per thread: 16 float32 / 8 float64 (64 bytes of input).
template <typename T>
constexpr int default_items() {
  return 64 / (int)sizeof(T);
}

// Staged-segment LDS bytes of one rolling tile (both inputs), 16-aligned; a
// 1/c table for mean follows when `lut` is set.
template <typename T>
__host__ __device__ inline int64_t rolling_stage_bytes(int items, int64_t window, bool tail_lds, int nargs) {
  const int64_t seg = (int64_t)kMoveThreads * items + (tail_lds ? window + 16 : 0);
  return ((pad16(seg) + 16) * (int64_t)sizeof(T) * nargs + 15) / 16 * 16;
}

// The first non-NaN value among staged segment elements e0 .. e0 + 255 (0 if
// none): one LDS read and a ballot per 64 elements; wave-uniform result.
template <typename T>
__device__ __forceinline__ T first_valid_staged(const T* seg, int e0) {
#pragma unroll 1
  for (int q = 0; q < kMoveThreads; q += kWave) {
    const unsigned long long m = __ballot(!is_nan(seg[pad16(e0 + q + lane_id())]));
    if (m != 0ull) return seg[pad16(e0 + q + __ffsll((long long)m) - 1)];
  }
  return T(0);
}

// One output tile [tile*TILE, tile*TILE + TILE) of one row (ra/rb/ro point at
// the row), TILE = 256 threads x KP items. `smem_raw` holds the staged segment;
// `lut` (mean only, may be null) holds 1/c for c = 0..window in LDS; `staged`
// leading segment elements are already in LDS (span kernels carry the halo).
// `full` tells whether the tile lies wholly inside the row (1 / 0; -1: test
// here); callers that know it from 32-bit tile counts keep the test on the
// scalar unit. Ends with a barrier so the caller may reuse the shared memory
// for its next tile.
template <typename T, int KIND, bool TAIL_LDS, int VEC, int KP = default_items<T>(), bool LUT = false,
          bool LDS_OUT = false>
__device__ __forceinline__ void rolling_tile(const T* __restrict__ ra, const T* __restrict__ rb,
                                             T* __restrict__ ro, int64_t n, int64_t window,
                                             int64_t min_count, int64_t tile,
                                             unsigned char* smem_raw,
                                             RollingShared<typename Traits<T, KIND>::S>& rs,
                                             const double* lut = nullptr, int staged = 0,
                                             int full = -1) {
  using Tr = Traits<T, KIND>;
  using S = typename Tr::S;
  constexpr int K = KP;
  constexpr int TILE = kMoveThreads * K;
  BlockShared<S>& sh = rs.sh;
  const int64_t t0 = tile * TILE;
  const int tid = threadIdx.x;

  // Segment staged in LDS: [s0a, t0 + TILE) with s0a <= t0 - window (TAIL_LDS;
  // positions before the row start are staged as NaN, so every tail index is
  // simply head index - window), or [t0, t0 + TILE). s0a is a multiple of 16,
  // so each thread's head chunk starts a padded 16-slot group.
  const int64_t seg_begin = TAIL_LDS ? t0 - window : t0;
  const int64_t s0a = seg_begin & ~(int64_t)15;  // floor, also for negatives
  const int64_t seg_len = t0 + TILE - s0a;
  const int64_t lds_elems = pad16(seg_len) + 16;
  T* la = reinterpret_cast<T*>(smem_raw);
  T* lb = la + lds_elems;

  const bool tile_full = full >= 0 ? full != 0 : t0 + TILE <= n;

  // 1. stage the segment (wave-striped vector loads, NaN past the row end).
  //    The first `staged` elements (a multiple of 16) may already be in LDS:
  //    the halo carried over from the previous tile of a span. A segment
  //    wholly inside the row (block-uniform test) is staged without checks.
  if (VEC > 1 && tile_full && s0a >= 0) {
    const T* rsa = ra + s0a;
    const T* rsb = (Tr::NARGS == 2) ? rb + s0a : nullptr;
    for (int e = staged + tid * VEC; e < (int)seg_len; e += kMoveThreads * VEC) {
      const Vec16<T> va = *reinterpret_cast<const Vec16<T>*>(rsa + e);
#pragma unroll
      for (int v = 0; v < VEC; ++v) la[pad16(e + v)] = va.v[v];
      if (Tr::NARGS == 2) {
        const Vec16<T> vb = *reinterpret_cast<const Vec16<T>*>(rsb + e);
#pragma unroll
        for (int v = 0; v < VEC; ++v) lb[pad16(e + v)] = vb.v[v];
      }
    }
  } else {
    for (int64_t e = (int64_t)staged + (int64_t)tid * VEC; e < seg_len; e += (int64_t)kMoveThreads * VEC) {
      const int64_t p = s0a + e;
      if (VEC > 1 && p >= 0 && p + VEC <= n) {
        Vec16<T> va = *reinterpret_cast<const Vec16<T>*>(ra + p);
#pragma unroll
        for (int v = 0; v < VEC; ++v) la[pad16(e + v)] = va.v[v];
        if (Tr::NARGS == 2) {
          Vec16<T> vb = *reinterpret_cast<const Vec16<T>*>(rb + p);
#pragma unroll
          for (int v = 0; v < VEC; ++v) lb[pad16(e + v)] = vb.v[v];
        }
      } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const int64_t q = p + v;
          const bool in = q >= 0 && q < n;
          la[pad16(e + v)] = in ? ra[q] : canonical_nan<T>();
          if (Tr::NARGS == 2) lb[pad16(e + v)] = in ? rb[q] : canonical_nan<T>();
        }
      }
    }
  }
  __syncthreads();

  // Per-tile shifts for the moment kinds: the first valid value among the
  // tile's first 256 inputs (any constant works; one near the data keeps the
  // float64 moments small). Every wave finds them itself from the staged
  // segment (64 inputs per ballot), so no further barrier is needed.
  T ka = T(0), kb = T(0);
  if (Tr::SHIFT) {
    // first 256 inputs of the tile itself (always staged; NaN past the row end)
    const int e0 = (int)(t0 - s0a);
    ka = first_valid_staged(la, e0);
    if (Tr::NARGS == 2) kb = first_valid_staged(lb, e0);
  }

  auto valid = [&](T a, T b) -> bool {
    return (Tr::NARGS == 2) ? !(is_nan(a) || is_nan(b)) : !is_nan(a);
  };

  // 2. halo statistics W0 = window ending at t0-1.
  constexpr bool FAST = !Tr::SHIFT && std::is_same<typename S::Acc, D>::value;
  S w0 = S::zero();
  if constexpr (FAST && TAIL_LDS) {
    // float32 sum/mean: four halo inputs per iteration at fixed LDS offsets
    // (pad16(e + T) = pad16(e) + T + T / 16 for T a multiple of 16), NaN -> 0
    // selected in float32.
    constexpr int PS = kMoveThreads + kMoveThreads / 16;  // padded stride of kMoveThreads elements
    const int e_end = (int)(t0 - s0a);
    int e = (int)(t0 - window - s0a) + tid;
    double hs = 0.0;
    int hc = 0;
    for (; e + 3 * kMoveThreads < e_end; e += 4 * kMoveThreads) {
      const T* q = la + pad16(e);
      const T v0 = q[0], v1 = q[PS], v2 = q[2 * PS], v3 = q[3 * PS];
      hs += (double)nan_to_zero(v0);
      hs += (double)nan_to_zero(v1);
      hs += (double)nan_to_zero(v2);
      hs += (double)nan_to_zero(v3);
      hc += (int)!is_nan(v0) + (int)!is_nan(v1) + (int)!is_nan(v2) + (int)!is_nan(v3);
    }
    for (; e < e_end; e += kMoveThreads) {
      const T v0 = la[pad16(e)];
      hs += (double)nan_to_zero(v0);
      hc += (int)!is_nan(v0);
    }
    w0.c = hc;
    w0.s.v = hs;
  } else if (TAIL_LDS) {
    const int e_end = (int)(t0 - s0a);
    for (int e = (int)(t0 - window - s0a) + tid; e < e_end; e += kMoveThreads) {
      const T a = la[pad16(e)];
      const T b = (Tr::NARGS == 2) ? lb[pad16(e)] : T(0);
      if (valid(a, b)) w0.add_obs(a, b, ka, kb);
    }
  } else {
    const int64_t h_begin = t0 - window > 0 ? t0 - window : 0;
    for (int64_t p = h_begin + tid; p < t0; p += kMoveThreads) {
      const T a = ra[p];
      const T b = (Tr::NARGS == 2) ? rb[p] : T(0);
      if (valid(a, b)) w0.add_obs(a, b, ka, kb);
    }
  }
  // The halo reduction is folded into the delta scan below (one barrier).

  // 3. head/tail chunks and the per-thread delta. Chunks are re-read from LDS
  //    (or global memory for long windows) rather than held in registers.
  const int64_t c0 = t0 + (int64_t)tid * K;  // first output position of this thread
  // LDS element index of the thread's first head / tail input (32-bit).
  const int eh = TAIL_LDS ? (int)(c0 - s0a) : tid * K;
  const int et = eh - (int)(TAIL_LDS ? window : 0);
  const int ehp = pad16(eh);  // eh is a multiple of 16: pad16(eh + j) = ehp + j + j / 16
  auto head = [&](int j, T& a, T& b) {
    a = la[ehp + j + (j >> 4)];
    b = (Tr::NARGS == 2) ? lb[ehp + j + (j >> 4)] : T(0);
  };
  // With K a multiple of 16, tail inputs sit at a block-uniform misalignment
  // mis = (-window) & 15 behind the thread's 16-aligned head chunk:
  // pad16(et + j) = tbase + j + ((mis + j) >> 4), a per-thread base plus a
  // uniform offset per j.
  const int mis = (int)((-window) & 15);
  const int tbase = pad16(eh - (int)((window + 15) & ~(int64_t)15)) + mis;
  auto tail = [&](int j, T& a, T& b) {
    if (TAIL_LDS) {
      const int o = (K % 16 == 0) ? tbase + j + ((mis + j) >> 4) : pad16(et + j);
      a = la[o];
      b = (Tr::NARGS == 2) ? lb[o] : T(0);
    } else {
      const int64_t pt = c0 + j - window;
      const bool in = pt >= 0 && pt < n;
      a = in ? ra[pt] : canonical_nan<T>();
      b = (Tr::NARGS == 2 && in) ? rb[pt] : canonical_nan<T>();
    }
  };
  T res[K];
  if constexpr (FAST) {
    // float32 sum/mean with float64 accumulation, branch-free: exact
    // differences and count deltas are formed first, then one running float64
    // sum and int32 count produce the outputs (1/c from the LDS table).
    double dj[K];
    int dc[K];
    double dsum = 0.0;
    int dcnt = 0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      T h, t, u;
      head(j, h, u);
      tail(j, t, u);
      const bool hv = !is_nan(h), tv = !is_nan(t);
      // NaN -> 0 selected in float32 before widening (one select each)
      const T h0 = nan_to_zero(h), t0v = nan_to_zero(t);
      dj[j] = (double)h0 - (double)t0v;
      dc[j] = (int)hv - (int)tv;
      dsum += dj[j];
      dcnt += dc[j];
    }
    S delta = S::zero();
    delta.c = dcnt;
    delta.s.v = dsum;
    S w0_total;
    S st = block_excl_scan_and_total(delta, w0, w0_total, sh, rs.sh_w);
    st.merge(w0_total);  // window ending at c0 - 1 (the tile's closing barrier guards sh)
    const int64_t mce = KIND == kSum ? min_count : (min_count > 1 ? min_count : 1);
    const int mc32 = mce > 0x7fffffffLL ? 0x7fffffff : (int)mce;
    int cc[K];
    int c = st.c;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      c += dc[j];
      cc[j] = c;
    }
    // Outputs in groups of 8: the group's 1/c values are loaded first (table
    // loads in flight together, with half the registers of all K at once).
    double acc = st.s.v;
    constexpr int G = K < 8 ? K : 8;
#pragma unroll
    for (int g = 0; g < K; g += G) {
      double rc[G];
      if constexpr (KIND != kSum) {
#pragma unroll
        for (int i = 0; i < G; ++i) {
          if constexpr (LUT) rc[i] = lut[cc[g + i]];  // c in [0, window]
          else rc[i] = recip(cc[g + i] > 0 ? cc[g + i] : 1);
        }
      }
#pragma unroll
      for (int i = 0; i < G; ++i) {
        const int j = g + i;
        acc += dj[j];
        if constexpr (KIND != kSum && LUT) {
          // the table holds NaN below the effective min_count
          res[j] = (T)(acc * rc[i]);
        } else {
          const double v = (KIND == kSum) ? acc : acc * rc[i];
          // Adding +0 (enough observations) or NaN (too few) instead of
          // selecting keeps the compiler from branching around each output.
          res[j] = (T)v + (cc[j] >= mc32 ? T(0) : canonical_nan<T>());
        }
      }
    }
  } else if constexpr (!Tr::SHIFT) {
    // Sum/mean: the exact per-position differences head - tail stay in
    // registers between the scan and the output pass.
    using Acc = typename S::Acc;
    Acc dj[K];
    unsigned long long hbits = 0ull, tbits = 0ull;
    S delta = S::zero();
#pragma unroll
    for (int j = 0; j < K; ++j) {
      T h, t, u;
      head(j, h, u);
      tail(j, t, u);
      const bool hv = !is_nan(h), tv = !is_nan(t);
      dj[j] = exact_diff<Acc, T>(hv ? h : T(0), tv ? t : T(0));
      hbits |= (unsigned long long)hv << j;
      tbits |= (unsigned long long)tv << j;
      delta.s.add(dj[j]);
      delta.c += (int)hv - (int)tv;
    }
    S w0_total;
    S st = block_excl_scan_and_total(delta, w0, w0_total, sh, rs.sh_w);
    st.merge(w0_total);  // window ending at c0 - 1
#pragma unroll
    for (int j = 0; j < K; ++j) {
      st.c += (int)((hbits >> j) & 1ull) - (int)((tbits >> j) & 1ull);
      st.s.add(dj[j]);
      res[j] = Tr::out(st, min_count, lut);
    }
  } else {
    S delta = S::zero();
#pragma unroll
    for (int j = 0; j < K; ++j) {
      T a, b;
      head(j, a, b);
      if (valid(a, b)) delta.add_obs(a, b, ka, kb);
      tail(j, a, b);
      if (valid(a, b)) delta.sub_obs(a, b, ka, kb);
    }
    S w0_total;
    S st = block_excl_scan_and_total(delta, w0, w0_total, sh, rs.sh_w);
    st.merge(w0_total);  // window ending at c0 - 1

    // 4. sequential add/remove over the thread's chunk.
#pragma unroll
    for (int j = 0; j < K; ++j) {
      T a, b;
      head(j, a, b);
      if (valid(a, b)) st.add_obs(a, b, ka, kb);
      tail(j, a, b);
      if (valid(a, b)) st.sub_obs(a, b, ka, kb);
      if constexpr (LUT) res[j] = Tr::out_tab(st, min_count, lut);  // c in [0, window]
      else res[j] = Tr::out(st, min_count, nullptr);
    }
  }
  // LDS_OUT (sum, mean): a full tile's outputs go through LDS so that each
  // store instruction writes consecutive 16-byte pieces across the wave (a
  // thread's own K outputs are K consecutive positions, so storing them
  // directly spreads one instruction over 64 separate pieces of
  // K * sizeof(T) bytes). They take the staged segment's first pad16(TILE)
  // slots: the caller's carried halo comes from beyond them, and this tile's
  // last reads of its inputs came before the scan's barrier. The moment
  // kinds, bound by their arithmetic, store directly (staging measured no
  // faster for var / std and slower for cov, whose registers it raises:
  // 1.93 -> 2.50 ms on float32 (256, 2M), window 1000). The stores stream
  // past the caches: a (256, 2M) output is far larger than they are.
  constexpr bool XPOSE = LDS_OUT && VEC > 1 && K % 16 == 0 && !Tr::SHIFT;
  if (XPOSE && tile_full) {
#pragma unroll
    for (int j = 0; j < K; ++j) la[tid * (K + K / 16) + j + (j >> 4)] = res[j];  // pad16(tid K + j)
    __syncthreads();
    char* rot = reinterpret_cast<char*>(ro + t0);
#pragma unroll
    for (int m = 0; m < K / VEC; ++m) {
      const int e = (m * kMoveThreads + tid) * VEC;  // VEC consecutive slots of one 16-slot group
      Vec16<T> vo;
#pragma unroll
      for (int v = 0; v < VEC; ++v) vo.v[v] = la[pad16(e) + v];
      store_stream16(reinterpret_cast<T*>(rot + (uint32_t)e * (uint32_t)sizeof(T)), vo.v);
    }
  } else if (VEC > 1 && tile_full) {
    // full tile (block-uniform test): no per-thread checks, 32-bit offsets
    // from the tile's base
    char* rot = reinterpret_cast<char*>(ro + t0);
    const uint32_t b0 = (uint32_t)tid * (uint32_t)(K * sizeof(T));
#pragma unroll
    for (int j = 0; j < K; j += VEC) {
      Vec16<T> vo;
#pragma unroll
      for (int v = 0; v < VEC; ++v) vo.v[v] = res[j + v];
      *reinterpret_cast<Vec16<T>*>(rot + (b0 + (uint32_t)(j * sizeof(T)))) = vo;
    }
  } else if (c0 + K <= n && VEC > 1) {
#pragma unroll
    for (int j = 0; j < K; j += VEC) {
      Vec16<T> vo;
#pragma unroll
      for (int v = 0; v < VEC; ++v) vo.v[v] = res[j + v];
      *reinterpret_cast<Vec16<T>*>(ro + c0 + j) = vo;
    }
  } else {
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (c0 + j < n) ro[c0 + j] = res[j];
  }
  __syncthreads();
}

// Largest window for which mean (and the float32 moment kinds) keep a
// reciprocal table in LDS.
constexpr int64_t kMeanLutWindow = 4096;

// lut[c] = 1/c (hardware reciprocal + one Newton step) for c >= mce, NaN for
// c < mce (too few observations: the product sum * lut[c] is then NaN, the
// masked output), c = 0..window. mce = max(min_count, 1).
__device__ __forceinline__ void fill_recip_lut(double* lut, int64_t window, int64_t mce) {
  for (int64_t c = threadIdx.x; c <= window; c += kMoveThreads)
    lut[c] = c >= mce ? recip((int)c) : canonical_nan<double>();
}

// Moment table: rc[c] = 1/max(c, 1) (the refined reciprocal recip forms),
// c = 0..window; 1/max(c - 1, 1) is rc[max(c - 1, 0)].
__device__ __forceinline__ void fill_moment_lut(double* rc, int64_t window) {
  for (int64_t c = threadIdx.x; c <= window; c += kMoveThreads) rc[c] = recip(c > 0 ? (int)c : 1);
}

// Items per thread for float32 sum/mean: 16 (default) or 32, from
// NBG_MOVE_ITEMS; a tuning knob (more items amortise the per-tile scan and
// halo over more outputs, fewer keep more waves resident). Plain and ragged
// launches read the same setting, so their tiles stay identical.
inline int f32_sum_items() {
  return knob("NBG_MOVE_ITEMS", 16) == 32 ? 32 : 16;
}

// Whether a float32 mean of this window uses the table (plain and ragged
// launches must agree, so both ask here).
__host__ __device__ inline bool mean_uses_lut(int64_t window) { return window <= kMeanLutWindow; }

// Tiles per span: a workgroup evaluates up to this many consecutive tiles of a
// row, carrying each tile's halo (the last ceil16(window) staged inputs) to the
// front of LDS for the next tile instead of staging it again from global
// memory, and filling the 1/c table once.
constexpr int64_t kMaxSpan = 16;

// The staged-halo length of a tile, ceil16(window); carrying it needs the
// source [TILE, TILE + H) and destination [0, H) to be disjoint.
__host__ __device__ inline int64_t carried_halo(int64_t window) { return (window + 15) & ~(int64_t)15; }

// The float32 single-input prefetch variants are held to 128 VGPRs (4 waves
// per SIMD); every other variant keeps the compiler's register choice.
template <typename T, int KIND, bool PF, int KP>
constexpr int move_min_waves() {
  return (PF && sizeof(T) == 4 && KP == 16 && Traits<T, KIND>::NARGS == 1) ? 4 : 1;
}

template <typename T, int KIND, bool TAIL_LDS, int VEC, int KP = default_items<T>(), bool LUT = false,
          bool PF = false>
__global__ __launch_bounds__(kMoveThreads) __attribute__((amdgpu_waves_per_eu(move_min_waves<T, KIND, PF, KP>()))) void move_kernel(
    const T* __restrict__ xa, const T* __restrict__ xb, T* __restrict__ out, int64_t rows,
    int64_t n, int64_t ld_a, int64_t ld_b, int64_t ld_out, int64_t window, int64_t min_count,
    int64_t tiles_per_row, int64_t span) {
  using Tr = Traits<T, KIND>;
  constexpr int TILE = kMoveThreads * KP;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  __shared__ RollingShared<typename Tr::S> rs;
  // Consecutive logical ids walk consecutive spans of a row on one XCD, so a
  // span's first halo (the end of the previous span) is read from that XCD's L2.
  // Tile counts fit 32 bits (the host bounds rows x tiles by 2**31), so the
  // span loop and its full-tile tests run on the scalar unit.
  const uint32_t spans_per_row = (uint32_t)((tiles_per_row + span - 1) / span);
  const uint32_t v = xcd_order(blockIdx.x, gridDim.x);
  const int64_t row = (int64_t)(v / spans_per_row);
  const int t_begin = (int)(v % spans_per_row) * (int)span;
  const int t_end = t_begin + (int)span < (int)tiles_per_row ? t_begin + (int)span : (int)tiles_per_row;
  const int full_tiles = (int)(n / TILE);  // tiles lying wholly inside the row
  double* lut = nullptr;
  if constexpr (LUT) {
    lut = reinterpret_cast<double*>(smem_raw + rolling_stage_bytes<T>(KP, window, TAIL_LDS, Tr::NARGS));
    // ordered by the tile's first barrier
    if constexpr (KIND == kMean) fill_recip_lut(lut, window, min_count > 1 ? min_count : 1);
    else fill_moment_lut(lut, window);
  }
  const T* ra = xa + row * ld_a;
  const T* rb = (Tr::NARGS == 2) ? xb + row * ld_b : nullptr;
  T* ro = out + row * ld_out;
  const int64_t H = carried_halo(window);
  const bool carry = TAIL_LDS && H <= TILE;
  // Padded LDS element count of one input's staged segment (the second input
  // follows it), as rolling_tile lays it out.
  const int64_t lds_elems = pad16((int64_t)TILE + H) + 16;
  const int seg_len = (int)(TILE + H);
  T* la = reinterpret_cast<T*>(smem_raw);
  // PF: software pipeline over the span: the next tile's new inputs (positions
  // [t0 + TILE, t0 + 2 TILE), segment offsets [H, H + TILE)) are loaded into
  // registers before the current tile is evaluated, and written to LDS after
  // it, behind the carried halo. Thread tid holds the elements it would stage
  // itself: offsets H + tid*VEC + m*256*VEC + v.
  constexpr int NV = KP / VEC;  // vectors of VEC elements per thread and input
  T pa[KP], pb[(Tr::NARGS == 2) ? KP : 1];
  auto prefetch = [&](int next_tile) {
    const int64_t t0n = (int64_t)next_tile * TILE;
    if (VEC > 1 && next_tile < full_tiles) {
      // full tile (block-uniform test): no checks, 32-bit offsets
      const char* rat = reinterpret_cast<const char*>(ra + t0n);
      const char* rbt = (Tr::NARGS == 2) ? reinterpret_cast<const char*>(rb + t0n) : nullptr;
#pragma unroll
      for (int m = 0; m < NV; ++m) {
        const uint32_t o = (threadIdx.x + (uint32_t)m * kMoveThreads) * 16u;  // bytes
        const Vec16<T> va = load_stream(reinterpret_cast<const Vec16<T>*>(rat + o));
#pragma unroll
        for (int v = 0; v < VEC; ++v) pa[m * VEC + v] = va.v[v];
        if constexpr (Tr::NARGS == 2) {
          const Vec16<T> vb = load_stream(reinterpret_cast<const Vec16<T>*>(rbt + o));
#pragma unroll
          for (int v = 0; v < VEC; ++v) pb[m * VEC + v] = vb.v[v];
        }
      }
      return;
    }
#pragma unroll
    for (int m = 0; m < NV; ++m) {
      const int64_t p = t0n + ((int64_t)threadIdx.x + (int64_t)m * kMoveThreads) * VEC;
      if (VEC > 1 && p + VEC <= n) {
        const Vec16<T> va = *reinterpret_cast<const Vec16<T>*>(ra + p);
#pragma unroll
        for (int v = 0; v < VEC; ++v) pa[m * VEC + v] = va.v[v];
        if constexpr (Tr::NARGS == 2) {
          const Vec16<T> vb = *reinterpret_cast<const Vec16<T>*>(rb + p);
#pragma unroll
          for (int v = 0; v < VEC; ++v) pb[m * VEC + v] = vb.v[v];
        }
      } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
          const bool in = p + v < n;
          pa[m * VEC + v] = in ? ra[p + v] : canonical_nan<T>();
          if constexpr (Tr::NARGS == 2) pb[m * VEC + v] = in ? rb[p + v] : canonical_nan<T>();
        }
      }
    }
  };
  int staged = 0;
  for (int tile = t_begin; tile < t_end; ++tile) {
    const bool more = carry && tile + 1 < t_end;
    if (PF && more) prefetch(tile + 1);
    rolling_tile<T, KIND, TAIL_LDS, VEC, KP, LUT, true>(ra, rb, ro, n, window, min_count, tile, smem_raw,
                                                        rs, lut, staged, tile < full_tiles ? 1 : 0);
    if (more) {
      // The next tile's segment starts TILE elements later: its first H
      // elements are this segment's last H (source and destination disjoint,
      // H <= TILE), the rest comes from the prefetched registers. The barrier
      // orders both before the next tile reads the segment.
      for (int e = threadIdx.x; e < (int)H; e += kMoveThreads) {
        la[pad16(e)] = la[pad16(e + TILE)];
        if (Tr::NARGS == 2) la[lds_elems + pad16(e)] = la[lds_elems + pad16(e + TILE)];
      }
      __syncthreads();
      if constexpr (PF) {
#pragma unroll
        for (int m = 0; m < NV; ++m) {
          const int e = (int)H + ((int)threadIdx.x + m * kMoveThreads) * VEC;
#pragma unroll
          for (int v = 0; v < VEC; ++v) {
            la[pad16(e + v)] = pa[m * VEC + v];
            if constexpr (Tr::NARGS == 2) la[lds_elems + pad16(e + v)] = pb[m * VEC + v];
          }
        }
        staged = seg_len;  // fully staged; rolling_tile's first barrier publishes these writes
      } else {
        staged = (int)H;  // rolling_tile stages the rest
      }
    }
  }
}

// Largest window whose halo is staged in LDS: the padded segment of
// TILE + window elements per input must fit a 60 KB budget.
template <typename T, int NARGS>
constexpr int64_t lds_window_limit() {
  constexpr int K = default_items<T>();
  constexpr int64_t budget = 60 * 1024;
  return budget * 16 / (17 * (int64_t)sizeof(T) * NARGS) - kMoveThreads * K - 64;
}

}  // namespace
}  // namespace nbg
