// Grouped reductions (group_nan{sum,mean,count,argmax,argmin,first,last,var,std,
// min,max,prod,sum_of_squares,any,all}) for gfx950.
//
// Reference semantics (numbagg/grouped.py:7-270): values (rows, n) share one
// label vector (n,); for every row, each valid value (not NaN) whose label is
// >= 0 is reduced into out[row, label]. Labels >= num_labels are skipped (the
// reference leaves them undefined). Empty labels give 0 (sum, count,
// sum_of_squares, any), 1 (prod, all), NaN (mean, var, std, arg*, first, last,
// min, max; integer outputs get the minimum integer, which is what numba's
// NaN-to-integer store produces).
//
// Design:
//   * a workgroup reduces one (row, chunk) of values. With the LDS strategy its
//     per-label accumulators live in LDS (privatised: conflicts only between
//     the lanes of one workgroup) and are flushed once per chunk with global
//     atomics; with the global strategy (label counts too large for LDS) every
//     element updates the global accumulators directly;
//   * a finalize kernel turns accumulators into the output dtype.
// Accumulators: float sums in float64 (rounded once to the output dtype),
// exact integer counts, 64-bit packed (orderable key, ~index) maxima for
// float32/int32 argmax/argmin/min/max so that the first index wins ties in
// one pass (min/max then copy the value at that index), key-then-index passes
// for 64-bit values, index min/max for first/last, float64 (or wrapping
// integer) products by compare-and-swap, flags for any/all, and a corrected
// two-pass variance (per-label means, then sums of deviations).
#include <mutex>
#include <unordered_map>

#include "common.hpp"

namespace nbg {
namespace {

constexpr int kGroupThreads = 256;
constexpr int kGroupUnroll = 4;  // default vectors per thread and super-step

enum GKind {
  kGSum = 0,
  kGMean = 1,
  kGCount = 2,
  kGArgMax = 3,
  kGArgMin = 4,
  kGFirst = 5,
  kGLast = 6,
  kGVar = 7,
  kGStd = 8,
  kGMin = 9,
  kGMax = 10,
  kGProd = 11,
  kGSumSq = 12,
  kGAny = 13,
  kGAll = 14,
};

// Internal accumulation passes.
enum GPass {
  kPSum = 0,      // s (f64 or i64) [+ count]
  kPCount = 1,    // count
  kPPacked = 2,   // packed (key32, ~idx32) max
  kPKey = 3,      // 64-bit key max
  kPKeyIdx = 4,   // min index among elements whose key equals the max
  kPFirst = 5,    // min index
  kPLast = 6,     // max index + 1
  kPDev = 7,      // sum of squared deviations from the label mean
  kPSumSq = 8,    // sum of squares (f64 or wrapping u64)
  kPProd = 9,     // product (f64 or wrapping u64), compare-and-swap updates
  kPAny = 10,     // flag: a valid non-zero value was seen
  kPAll = 11,     // flag: a valid zero was seen
  kPMom = 12,     // shifted moments: count, sum and sum of squares of x - K (K: a value of the label)
};

template <typename T>
__device__ __forceinline__ bool gvalid(T v) {
  return !is_nan(v);
}

// Order-preserving maps to unsigned integers (argmax uses them directly, argmin
// their complement). -0.0 is folded onto +0.0: the reference compares values
// with `>`/`<`, under which the two zeros tie.
__device__ __forceinline__ uint32_t okey32(float v) {
  v = v + 0.0f;
  uint32_t b = __float_as_uint(v);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ uint32_t okey32(int32_t v) { return (uint32_t)v ^ 0x80000000u; }
__device__ __forceinline__ uint64_t okey64(double v) {
  v = v + 0.0;
  uint64_t b = (uint64_t)__double_as_longlong(v);
  return (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
}
// int64 values are compared as float64, like the reference (its running
// extreme is a float64 array).
__device__ __forceinline__ uint64_t okey64(int64_t v) { return okey64((double)v); }
__device__ __forceinline__ uint64_t okey64(float v) { return okey64((double)v); }
__device__ __forceinline__ uint64_t okey64(int32_t v) { return okey64((double)v); }

template <typename TV>
struct SumAcc {
  using type = double;
};
// Integer sums wrap modulo 2^64 and are narrowed at the end, which reproduces
// the reference's wrapping integer adds.
template <>
struct SumAcc<int32_t> {
  using type = unsigned long long;
};
template <>
struct SumAcc<int64_t> {
  using type = unsigned long long;
};

struct GArgs {
  const void* values;
  const void* labels;
  void* out;
  int64_t rows, n, ld_v, num_labels, ddof;
  int64_t chunk, chunks_per_row;  // set per launch (launch_pass)
  // global accumulators (rows x num_labels each)
  void* acc0;
  void* acc1;
  void* acc2;           // variance finalize: counts
  void* acc3;           // kPMom: per (row, label) shift (float64 bits, NaN until set)
  const double* means;  // kPDev: per (row, label) mean
  int argmin;           // kPPacked / kPKey / kPKeyIdx: complement keys
  int mom;              // variance finalize: from shifted moments (kPMom) instead of two passes
};

// Canonical NaN bit patterns: the "no shift yet" marker of kPMom slots.
template <typename TV>
constexpr unsigned long long nan_bits() {
  return sizeof(TV) == 4 ? 0x7FC00000ull : 0x7FF8000000000000ull;
}
template <typename TV>
__device__ __forceinline__ TV tv_from_bits(unsigned long long b) {
  if constexpr (sizeof(TV) == 4) return __uint_as_float((unsigned int)b);
  else return __longlong_as_double((long long)b);
}
template <typename TV>
__device__ __forceinline__ unsigned long long tv_bits(TV v) {
  if constexpr (sizeof(TV) == 4) return (unsigned long long)__float_as_uint(v);
  else return (unsigned long long)__double_as_longlong(v);
}
// Claim the shift of an LDS slot: the first value to arrive becomes the
// slot's shift; returns the shift in place afterwards.
template <typename TV>
__device__ __forceinline__ TV lds_claim_shift(TV* slot, TV v) {
  if constexpr (sizeof(TV) == 4) {
    const unsigned int old = atomicCAS(reinterpret_cast<unsigned int*>(slot), (unsigned int)nan_bits<TV>(),
                                       __float_as_uint(v));
    return old == (unsigned int)nan_bits<TV>() ? v : __uint_as_float(old);
  } else {
    const unsigned long long old = atomicCAS(reinterpret_cast<unsigned long long*>(slot), nan_bits<TV>(),
                                             (unsigned long long)__double_as_longlong(v));
    return old == nan_bits<TV>() ? v : __longlong_as_double((long long)old);
  }
}

__device__ __forceinline__ void g_atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }
__device__ __forceinline__ void g_atomic_add(unsigned long long* p, unsigned long long v) {
  atomicAdd(p, v);
}

// Multiply an accumulator (LDS or global) by v with a compare-and-swap loop:
// no hardware atomic multiplies. Products are formed in float64 (float data)
// or as wrapping 64-bit integers (integer data, narrowed at the end as the
// reference's wrapping integer multiplies are).
__device__ __forceinline__ double mul_acc(double a, double v) { return a * v; }
__device__ __forceinline__ unsigned long long mul_acc(unsigned long long a, unsigned long long v) { return a * v; }
template <typename SA>
__device__ __forceinline__ void atomic_mul(SA* p, SA v) {
  unsigned long long* q = reinterpret_cast<unsigned long long*>(p);
  unsigned long long old = *reinterpret_cast<volatile unsigned long long*>(q);
  while (true) {
    SA cur;
    __builtin_memcpy(&cur, &old, 8);
    const SA nv = mul_acc(cur, v);
    unsigned long long nb;
    __builtin_memcpy(&nb, &nv, 8);
    const unsigned long long seen = atomicCAS(q, old, nb);
    if (seen == old) break;
    old = seen;
  }
}
// The identity of the product accumulator, as a 64-bit word.
template <typename SA>
constexpr unsigned long long one_bits() {
  return std::is_same<SA, double>::value ? 0x3FF0000000000000ull : 1ull;
}

// One accumulation pass over R rows x one chunk of positions. The R rows
// share each label load. LDS (privatised strategy): per-(row, label)
// accumulators, flushed once with global atomics.
// HOT (float sum passes with LDS accumulators): in chunks whose labels are
// skewed towards labels 0 and 1, those two labels of every row accumulate in
// registers, and each wave adds them to the LDS slots once at the end. Skewed
// label distributions put most values on the first labels (Zipf 1.1 over 1000
// labels: 18% on label 0, 8% on label 1), whose LDS atomics would otherwise
// serialise within a wave.
template <typename TV, typename TL, int PASS, bool LDS, int R, bool COUNT, int U = kGroupUnroll,
          bool PIPE = false, bool HOT = false>
__global__ __launch_bounds__(kGroupThreads) void group_pass_kernel(GArgs A) {
  using SA = typename SumAcc<TV>::type;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int64_t L1 = A.num_labels;        // labels per row
  const int64_t L = L1 * R;               // LDS slots (row-major: r * L1 + label)
  // Chunk-major order on XCD-consecutive logical ids: the row groups of one
  // chunk run on the same XCD back to back and read its labels from that
  // XCD's L2 after the first of them.
  const uint64_t groups = (uint64_t)((A.rows + R - 1) / R);
  const uint32_t v = xcd_order(blockIdx.x, gridDim.x);
  const int64_t row0 = (int64_t)(v % groups) * R;
  const int64_t ck = (int64_t)(v / groups);
  const int64_t c0 = ck * A.chunk;
  const int64_t c1 = (c0 + A.chunk < A.n) ? c0 + A.chunk : A.n;
  const int nr = (int)((A.rows - row0) < R ? (A.rows - row0) : R);  // rows present
  const TV* vbase = static_cast<const TV*>(A.values) + row0 * A.ld_v;
  const TL* lab = static_cast<const TL*>(A.labels);
  const int tid = threadIdx.x;

  // Global accumulators of these rows (slot r * L1 + label, rows contiguous).
  SA* g_s = static_cast<SA*>(A.acc0) + row0 * L1;
  unsigned long long* g_u0 = static_cast<unsigned long long*>(A.acc0) + row0 * L1;
  unsigned long long* g_u1 = static_cast<unsigned long long*>(A.acc1) + row0 * L1;
  double* g_sq = static_cast<double*>(A.acc1) + row0 * L1;
  const double* g_mean = A.means ? A.means + row0 * L1 : nullptr;

  // LDS accumulators.
  SA* l_s = reinterpret_cast<SA*>(smem_raw);
  unsigned int* l_c = reinterpret_cast<unsigned int*>(smem_raw + (PASS == kPSum ? L * sizeof(SA) : 0));
  unsigned long long* l_u = reinterpret_cast<unsigned long long*>(smem_raw);
  unsigned int* l_u32 = reinterpret_cast<unsigned int*>(smem_raw);
  double* l_sq = reinterpret_cast<double*>(smem_raw);
  double* l_mean = l_sq + L;
  unsigned long long* l_key = l_u + L;  // kPKeyIdx: staged per-label maxima
  // kPMom: sums of y = x - K and y^2 (float64), shifts K (values dtype), counts
  double* l_m1 = reinterpret_cast<double*>(smem_raw);
  double* l_m2 = l_m1 + L;
  TV* l_mk = reinterpret_cast<TV*>(l_m2 + L);
  unsigned int* l_mc = reinterpret_cast<unsigned int*>(l_mk + L);
  const int64_t Lp = (int64_t)nr * L1;  // slots actually used
  if (LDS) {
    for (int64_t l = tid; l < Lp; l += kGroupThreads) {
      if constexpr (PASS == kPSum || PASS == kPSumSq) {
        l_s[l] = SA(0);
        if (PASS == kPSum && COUNT) l_c[l] = 0u;
      } else if constexpr (PASS == kPProd) {
        l_s[l] = SA(1);
      } else if constexpr (PASS == kPCount || PASS == kPAny || PASS == kPAll) {
        l_u32[l] = 0u;
      } else if constexpr (PASS == kPPacked || PASS == kPKey || PASS == kPLast) {
        l_u[l] = 0ull;
      } else if constexpr (PASS == kPKeyIdx) {
        l_u[l] = ~0ull;
        l_key[l] = g_u0[l];
      } else if constexpr (PASS == kPFirst) {
        l_u[l] = ~0ull;
      } else if constexpr (PASS == kPDev) {
        l_sq[l] = 0.0;
        l_mean[l] = g_mean[l];
      } else if constexpr (PASS == kPMom) {
        l_m1[l] = 0.0;
        l_m2[l] = 0.0;
        l_mc[l] = 0u;
        l_mk[l] = tv_from_bits<TV>(nan_bits<TV>());
      }
    }
    __syncthreads();
  }
  static_assert(!HOT || (LDS && (PASS == kPSum || PASS == kPDev || PASS == kPMom || PASS == kPSumSq ||
                                 PASS == kPProd) &&
                         std::is_same<SA, double>::value),
                "hot labels: float sum passes with LDS accumulators only");
  static_assert(PASS != kPMom || (LDS && std::is_floating_point<TV>::value),
                "shifted moments: float values with LDS accumulators");
  constexpr int NH = HOT ? 2 : 1;
  double h_s[R][NH];          // kPSum: sums; kPDev: sums of squared deviations; kPMom: sums of y
  unsigned int h_c[R][NH];    // kPSum with COUNT, kPMom: counts
  double h_mean[R][NH];       // kPDev: the hot labels' means
  double h_s2[R][NH];         // kPMom: sums of y^2
  TV h_k[R][NH];              // kPMom: this thread's shift per hot label (NaN: none yet)
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int h = 0; h < NH; ++h) {
      h_s[r][h] = PASS == kPProd ? 1.0 : 0.0;
      h_c[r][h] = 0u;
      h_mean[r][h] = 0.0;
      h_s2[r][h] = 0.0;
      h_k[r][h] = tv_from_bits<TV>(nan_bits<TV>());
      if constexpr (HOT && PASS == kPDev) {
        if (r < nr && h < L1) h_mean[r][h] = l_mean[r * L1 + h];
      }
    }

  // Update slot r * L1 + label with value v at position i.
  // One unsigned compare rejects both negative and too-large labels (32-bit
  // for int32 labels: every valid label is below 2**31).
  const uint32_t lim32 = L1 < 0x80000000LL ? (uint32_t)L1 : 0x80000000u;
  // hot_c: std::true_type on the register path for labels 0 and 1.
  auto update = [&](auto hot_c, TV v, TL lbl, int64_t i, int r) {
    constexpr bool H = decltype(hot_c)::value;
    if constexpr (sizeof(TL) == 4) {
      if ((uint32_t)lbl >= lim32) return;
    } else {
      if ((uint64_t)lbl >= (uint64_t)L1) return;
    }
    if (!gvalid(v)) return;
    if constexpr (H) {
      if ((uint32_t)lbl < 2u) {  // r is a compile-time index at every call site
        const bool first = lbl == 0;
        if constexpr (PASS == kPSum) {
          h_s[r][0] += first ? (double)v : 0.0;
          h_s[r][1] += first ? 0.0 : (double)v;
          if (COUNT) {
            h_c[r][0] += first ? 1u : 0u;
            h_c[r][1] += first ? 0u : 1u;
          }
        } else if constexpr (PASS == kPMom) {
          // the thread's first value of the label is its shift
          const TV kc = first ? h_k[r][0] : h_k[r][1];
          const TV k = is_nan(kc) ? v : kc;
          h_k[r][0] = first ? k : h_k[r][0];
          h_k[r][1] = first ? h_k[r][1] : k;
          const double y = (double)v - (double)k;
          h_s[r][0] += first ? y : 0.0;
          h_s[r][1] += first ? 0.0 : y;
          h_s2[r][0] += first ? y * y : 0.0;
          h_s2[r][1] += first ? 0.0 : y * y;
          h_c[r][0] += first ? 1u : 0u;
          h_c[r][1] += first ? 0u : 1u;
        } else if constexpr (PASS == kPSumSq) {
          const double sq = (double)v * (double)v;
          h_s[r][0] += first ? sq : 0.0;
          h_s[r][1] += first ? 0.0 : sq;
        } else if constexpr (PASS == kPProd) {
          h_s[r][0] *= first ? (double)v : 1.0;
          h_s[r][1] *= first ? 1.0 : (double)v;
        } else {
          const double d = (double)v - (first ? h_mean[r][0] : h_mean[r][1]);
          h_s[r][0] += first ? d * d : 0.0;
          h_s[r][1] += first ? 0.0 : d * d;
        }
        return;
      }
    }
    // LDS slots fit 32-bit indexing (the LDS budget bounds R * L1)
    const int64_t l = LDS ? (int64_t)(r * (int)L1 + (int)lbl) : (int64_t)r * L1 + (int64_t)lbl;
    if constexpr (PASS == kPSum) {
      if (LDS) {
        atomicAdd(&l_s[l], (SA)v);  // ds_add_f64 / ds_add_u64
        if (COUNT) atomicAdd(&l_c[l], 1u);
      } else {
        g_atomic_add(&g_s[l], (SA)v);
        if (COUNT) atomicAdd(&g_u1[l], 1ull);
      }
    } else if constexpr (PASS == kPSumSq) {
      const SA sq = (SA)v * (SA)v;  // exact float64 squares of float32 data; wrapping for integers
      if (LDS) atomicAdd(&l_s[l], sq);
      else g_atomic_add(&g_s[l], sq);
    } else if constexpr (PASS == kPProd) {
      if (LDS) atomic_mul(&l_s[l], (SA)v);
      else atomic_mul(&g_s[l], (SA)v);
    } else if constexpr (PASS == kPAny || PASS == kPAll) {
      // any: a valid non-zero value; all: a valid zero (-0.0 counts as zero)
      if ((PASS == kPAny) == (v != TV(0))) {
        if (LDS) atomicOr(&l_u32[l], 1u);
        else atomicOr(&g_u0[l], 1ull);
      }
    } else if constexpr (PASS == kPCount) {
      if (LDS) atomicAdd(&l_u32[l], 1u);
      else atomicAdd(&g_u0[l], 1ull);
    } else if constexpr (PASS == kPPacked) {
      uint32_t k = okey32(v);
      if (A.argmin) k = ~k;
      const unsigned long long w = ((unsigned long long)k << 32) | (unsigned long long)(~(uint32_t)i);
      if (LDS) atomicMax(&l_u[l], w);
      else atomicMax(&g_u0[l], w);
    } else if constexpr (PASS == kPKey) {
      uint64_t k = okey64(v);
      if (A.argmin) k = ~k;
      if (LDS) atomicMax(&l_u[l], (unsigned long long)k);
      else atomicMax(&g_u0[l], (unsigned long long)k);
    } else if constexpr (PASS == kPKeyIdx) {
      uint64_t k = okey64(v);
      if (A.argmin) k = ~k;
      if (k == (LDS ? l_key[l] : g_u0[l])) {
        if (LDS) atomicMin(&l_u[l], (unsigned long long)i);
        else atomicMin(&g_u1[l], (unsigned long long)i);
      }
    } else if constexpr (PASS == kPFirst) {
      if (LDS) atomicMin(&l_u[l], (unsigned long long)i);
      else atomicMin(&g_u0[l], (unsigned long long)i);
    } else if constexpr (PASS == kPLast) {
      if (LDS) atomicMax(&l_u[l], (unsigned long long)(i + 1));
      else atomicMax(&g_u0[l], (unsigned long long)(i + 1));
    } else if constexpr (PASS == kPDev) {
      const double d = (double)v - (LDS ? l_mean[l] : g_mean[l]);
      if (LDS) atomicAdd(&l_sq[l], d * d);
      else unsafeAtomicAdd(&g_sq[l], d * d);
    } else if constexpr (PASS == kPMom) {
      // y = x - K with K the slot's shift (the first value of the label to
      // reach this workgroup): exact in float64 for float32 data, and small
      // where the label's values are, so y^2 sums do not cancel against the
      // squared mean at the end
      TV k = l_mk[l];
      if (is_nan(k)) k = lds_claim_shift(&l_mk[l], v);
      const double y = (double)v - (double)k;
      atomicAdd(&l_m1[l], y);
      atomicAdd(&l_m2[l], y * y);
      atomicAdd(&l_mc[l], 1u);
    }
  };

  // Stream the chunk: every thread keeps U vectors of values and labels in
  // flight; full super-steps first, then a coalesced remainder.
  constexpr int VV = 16 / (int)sizeof(TV);
  constexpr int VL = 16 / (int)sizeof(TL);
  constexpr int VE = VV < VL ? VV : VL;  // elements per vector step
  constexpr int64_t STEP = (int64_t)kGroupThreads * VE;
  constexpr int64_t SUPER = STEP * U;
  const bool aligned = (((uintptr_t)(vbase + c0)) % 16 == 0) && (((uintptr_t)(lab + c0)) % 16 == 0) &&
                       (nr == 1 || (A.ld_v % VE) == 0);
  const int64_t nfull = aligned ? (c1 - c0) / SUPER : 0;
  struct alignas(VE * sizeof(TV)) V { TV v[VE]; };
  struct alignas(VE * sizeof(TL)) W { TL v[VE]; };
  // HOT: the register path runs when labels 0 and 1 are frequent in the
  // chunk, judged from its first 256 x VE labels (one block-wide count);
  // otherwise the plain loop runs, with no per-element test.
  bool hot_on = false;
  if constexpr (HOT) {
    int nh = 0;
#pragma unroll
    for (int e = 0; e < VE; ++e) {
      const int64_t j = c0 + (int64_t)tid * VE + e;
      if (j < c1) nh += (lab[j] >= 0 && lab[j] < 2) ? 1 : 0;
    }
    hot_on = __syncthreads_count(nh > 0) > kGroupThreads / 8;
  }
  auto stream = [&](auto hot_c) {
    if constexpr (!PIPE) {
      for (int64_t k = 0; k < nfull; ++k) {
        const int64_t base = c0 + k * SUPER + (int64_t)tid * VE;
        // All loads of the super-step (labels once, then every row's values)
        // are issued before any update, so one wait covers them.
        W ll[U];
        V vv[R][U];
#pragma unroll
        for (int u = 0; u < U; ++u) ll[u] = *reinterpret_cast<const W*>(lab + base + u * STEP);
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r < nr) {
            const TV* vr = vbase + (int64_t)r * A.ld_v;
#pragma unroll
            for (int u = 0; u < U; ++u) vv[r][u] = load_stream(reinterpret_cast<const V*>(vr + base + u * STEP));
          }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r >= nr) break;
#pragma unroll
          for (int u = 0; u < U; ++u)
#pragma unroll
            for (int e = 0; e < VE; ++e) update(hot_c, vv[r][u].v[e], ll[u].v[e], base + u * STEP + e, r);
        }
      }
    } else if (nfull > 0) {
      // Pipelined: the values of the next row (or the next super-step's labels
      // and first row) are loaded while the current row is folded in.
      W ll[U], lln[U];
      V vc[U], vn[U];
      {
        const int64_t b0 = c0 + (int64_t)tid * VE;
#pragma unroll
        for (int u = 0; u < U; ++u) {
          ll[u] = *reinterpret_cast<const W*>(lab + b0 + u * STEP);
          vc[u] = load_stream(reinterpret_cast<const V*>(vbase + b0 + u * STEP));
        }
      }
      for (int64_t k = 0; k < nfull; ++k) {
        const int64_t base = c0 + k * SUPER + (int64_t)tid * VE;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          if (r < nr) {
            if (r + 1 < nr) {
              const TV* vr = vbase + (int64_t)(r + 1) * A.ld_v;
#pragma unroll
              for (int u = 0; u < U; ++u) vn[u] = load_stream(reinterpret_cast<const V*>(vr + base + u * STEP));
            } else if (k + 1 < nfull) {
#pragma unroll
              for (int u = 0; u < U; ++u) {
                lln[u] = *reinterpret_cast<const W*>(lab + base + SUPER + u * STEP);
                vn[u] = load_stream(reinterpret_cast<const V*>(vbase + base + SUPER + u * STEP));
              }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
              for (int e = 0; e < VE; ++e) update(hot_c, vc[u].v[e], ll[u].v[e], base + u * STEP + e, r);
#pragma unroll
            for (int u = 0; u < U; ++u) vc[u] = vn[u];
          }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) ll[u] = lln[u];
      }
    }
    for (int64_t j = c0 + nfull * SUPER + tid; j < c1; j += kGroupThreads) {
      const TL lj = lab[j];
#pragma unroll
      for (int r = 0; r < R; ++r)
        if (r < nr) update(hot_c, vbase[(int64_t)r * A.ld_v + j], lj, j, r);
    }
  };
  if constexpr (HOT) {
    if (hot_on) stream(std::true_type{});
    else stream(std::false_type{});
  } else {
    stream(std::false_type{});
  }

  if (LDS) {
    if constexpr (HOT && PASS == kPMom) {
      if (hot_on) {  // the hot labels' moments, moved to the slot's shift, then wave sums
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int h = 0; h < NH; ++h) {
            const unsigned long long act = __ballot(h_c[r][h] != 0u);
            if (act == 0ull || r >= nr || h >= L1) continue;  // wave-uniform
            const int l = r * (int)L1 + h;
            const int src = __ffsll((long long)act) - 1;
            TV kslot = l_mk[l];
            if (lane_id() == src && is_nan(kslot)) kslot = lds_claim_shift(&l_mk[l], h_k[r][h]);
            kslot = __shfl(kslot, src, kWave);
            double c = (double)h_c[r][h];
            const double dl = h_c[r][h] ? (double)h_k[r][h] - (double)kslot : 0.0;
            double x1 = h_s[r][h] + c * dl;
            double x2 = h_s2[r][h] + 2.0 * dl * h_s[r][h] + c * dl * dl;
            unsigned int cc = h_c[r][h];
#pragma unroll
            for (int o = kWave / 2; o > 0; o >>= 1) {
              x1 += __shfl_xor(x1, o);
              x2 += __shfl_xor(x2, o);
              cc += __shfl_xor(cc, o);
            }
            (void)c;
            if (lane_id() == 0) {
              atomicAdd(&l_m1[l], x1);
              atomicAdd(&l_m2[l], x2);
              atomicAdd(&l_mc[l], cc);
            }
          }
      }
    } else if constexpr (HOT) {
      if (hot_on) {  // wave sums of the hot labels into their LDS slots (block-uniform)
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int h = 0; h < NH; ++h) {
          double x = h_s[r][h];
          unsigned int c = h_c[r][h];
#pragma unroll
          for (int o = kWave / 2; o > 0; o >>= 1) {
            if constexpr (PASS == kPProd) x *= __shfl_xor(x, o);
            else x += __shfl_xor(x, o);
            if (COUNT) c += __shfl_xor(c, o);
          }
          if (lane_id() == 0 && r < nr && h < L1) {
            const int l = r * (int)L1 + h;
            if constexpr (PASS == kPSum) {
              atomicAdd(&l_s[l], x);
              if (COUNT) atomicAdd(&l_c[l], c);
            } else if constexpr (PASS == kPSumSq) {
              atomicAdd(&l_s[l], x);
            } else if constexpr (PASS == kPProd) {
              if (x != 1.0) atomic_mul(&l_s[l], x);
            } else {
              atomicAdd(&l_sq[l], x);
            }
          }
        }
      }
    }
    __syncthreads();
    for (int64_t l = tid; l < Lp; l += kGroupThreads) {
      if constexpr (PASS == kPSum) {
        if (COUNT) {
          const unsigned int c = l_c[l];
          if (c) {
            g_atomic_add(&g_s[l], l_s[l]);
            atomicAdd(&g_u1[l], (unsigned long long)c);
          }
        } else if (l_s[l] != SA(0)) {
          g_atomic_add(&g_s[l], l_s[l]);
        }
      } else if constexpr (PASS == kPSumSq) {
        if (l_s[l] != SA(0)) g_atomic_add(&g_s[l], l_s[l]);
      } else if constexpr (PASS == kPProd) {
        if (l_s[l] != SA(1)) atomic_mul(&g_s[l], l_s[l]);
      } else if constexpr (PASS == kPAny || PASS == kPAll) {
        if (l_u32[l]) atomicOr(&g_u0[l], 1ull);
      } else if constexpr (PASS == kPCount) {
        if (l_u32[l]) atomicAdd(&g_u0[l], (unsigned long long)l_u32[l]);
      } else if constexpr (PASS == kPPacked || PASS == kPKey || PASS == kPLast) {
        if (l_u[l]) atomicMax(&g_u0[l], l_u[l]);
      } else if constexpr (PASS == kPKeyIdx) {
        if (l_u[l] != ~0ull) atomicMin(&g_u1[l], l_u[l]);
      } else if constexpr (PASS == kPFirst) {
        if (l_u[l] != ~0ull) atomicMin(&g_u0[l], l_u[l]);
      } else if constexpr (PASS == kPDev) {
        if (l_sq[l] != 0.0) unsafeAtomicAdd(&g_sq[l], l_sq[l]);
      } else if constexpr (PASS == kPMom) {
        // moved to the row's global shift Kg (the first workgroup to flush
        // the label sets it): sums of x - Kg = s1 + c d, of (x - Kg)^2 =
        // s2 + 2 d s1 + c d^2, with d = K - Kg
        const unsigned int c = l_mc[l];
        if (c) {
          const double k = (double)l_mk[l];
          unsigned long long* gk = static_cast<unsigned long long*>(A.acc3) + row0 * L1 + l;
          const unsigned long long kb = (unsigned long long)__double_as_longlong(k);
          const unsigned long long old = atomicCAS(gk, 0x7FF8000000000000ull, kb);
          const double kg = old == 0x7FF8000000000000ull ? k : __longlong_as_double((long long)old);
          const double d = k - kg;
          const double s1 = l_m1[l], s2 = l_m2[l], dc = (double)c;
          unsafeAtomicAdd(&static_cast<double*>(A.acc0)[row0 * L1 + l], s1 + dc * d);
          unsafeAtomicAdd(&static_cast<double*>(A.acc1)[row0 * L1 + l], s2 + 2.0 * d * s1 + dc * d * d);
          atomicAdd(&static_cast<unsigned long long*>(A.acc2)[row0 * L1 + l], (unsigned long long)c);
        }
      }
    }
  }
}

// Initialise accumulators: acc0/acc1 words set to the pass identities.
__global__ void group_init_kernel(unsigned long long* acc0, unsigned long long init0,
                                  unsigned long long* acc1, unsigned long long init1,
                                  int64_t count) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (int64_t)gridDim.x * blockDim.x) {
    acc0[i] = init0;
    if (acc1) acc1[i] = init1;
  }
}

template <typename TV>
__device__ __forceinline__ TV missing_value() {
  if (sizeof(TV) == 4 && (TV)0.5 != 0) return (TV)canonical_nan<float>();
  if ((TV)0.5 != 0) return (TV)canonical_nan<double>();
  return (TV)(sizeof(TV) == 4 ? (TV)0x80000000u : (TV)0x8000000000000000ull);
}

// Turn accumulators into outputs. `means` mode writes per-label means for the
// second variance pass instead of outputs.
template <typename TV, int KIND>
__global__ void group_finalize_kernel(GArgs A, double* means_out) {
  using SA = typename SumAcc<TV>::type;
  const int64_t total = A.rows * A.num_labels;
  TV* out = static_cast<TV*>(A.out);
  const TV* vals = static_cast<const TV*>(A.values);
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < total;
       q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = q / A.num_labels;
    if (means_out) {
      const double s = static_cast<const double*>(A.acc0)[q];
      const unsigned long long c = static_cast<const unsigned long long*>(A.acc1)[q];
      means_out[q] = c ? s / (double)c : 0.0;
      continue;
    }
    TV r;
    if constexpr (KIND == kGSum) {
      r = (TV)static_cast<const SA*>(A.acc0)[q];  // integers: modular narrowing
    } else if constexpr (KIND == kGMean) {
      const double s = static_cast<const double*>(A.acc0)[q];
      const unsigned long long c = static_cast<const unsigned long long*>(A.acc1)[q];
      r = c ? (TV)(s / (double)c) : missing_value<TV>();
    } else if constexpr (KIND == kGCount) {
      r = (TV)static_cast<const unsigned long long*>(A.acc0)[q];
    } else if constexpr (KIND == kGArgMax || KIND == kGArgMin || KIND == kGMin || KIND == kGMax) {
      // index of the extreme: packed (key32, ~idx32) words, or the key-index pass
      bool has;
      unsigned long long idx;
      if (A.acc1 == nullptr) {
        const unsigned long long w = static_cast<const unsigned long long*>(A.acc0)[q];
        has = w != 0ull;
        idx = (uint32_t)(~(uint32_t)(w & 0xffffffffull));
      } else {
        idx = static_cast<const unsigned long long*>(A.acc1)[q];
        has = idx != ~0ull;
      }
      if constexpr (KIND == kGArgMax || KIND == kGArgMin) {
        r = has ? (TV)(double)idx : missing_value<TV>();
      } else {
        // the value at that index, through float64 as the reference's float64
        // running extreme stores it (64-bit integers round there)
        r = has ? (TV)(double)vals[row * A.ld_v + (int64_t)idx] : missing_value<TV>();
      }
    } else if constexpr (KIND == kGSumSq || KIND == kGProd) {
      r = (TV)static_cast<const SA*>(A.acc0)[q];
    } else if constexpr (KIND == kGAny || KIND == kGAll) {
      const bool seen = static_cast<const unsigned long long*>(A.acc0)[q] != 0ull;
      r = (KIND == kGAny) == seen ? TV(1) : TV(0);
    } else if constexpr (KIND == kGFirst) {
      const unsigned long long i = static_cast<const unsigned long long*>(A.acc0)[q];
      r = (i != ~0ull) ? vals[row * A.ld_v + (int64_t)i] : missing_value<TV>();
    } else if constexpr (KIND == kGLast) {
      const unsigned long long i = static_cast<const unsigned long long*>(A.acc0)[q];
      r = i ? vals[row * A.ld_v + (int64_t)(i - 1)] : missing_value<TV>();
    } else {  // var / std from (count, sum of squared deviations)
      const unsigned long long c = static_cast<const unsigned long long*>(A.acc2)[q];
      double m2 = static_cast<const double*>(A.acc1)[q];
      if (A.mom && c) {
        // shifted moments: M2 = sum (x-Kg)^2 - (sum (x-Kg))^2 / n, kept >= 0
        const double s1 = static_cast<const double*>(A.acc0)[q];
        m2 = m2 - s1 * s1 / (double)c;
        m2 = m2 > 0.0 ? m2 : 0.0;
      }
      const double den = (double)((long long)c - A.ddof);
      if (c == 0 || !(den > 0)) {  // an empty label is NaN whatever ddof (0/0 in the reference)
        r = missing_value<TV>();
      } else {
        const double var = m2 / den;
        r = (TV)(KIND == kGStd ? sqrt(var) : var);
      }
    }
    out[q] = r;
  }
}

// LDS budget for privatised accumulators (bytes per workgroup).
constexpr int64_t kGroupLdsBudget = 48 * 1024;

// The most LDS one workgroup may use on the current device (the forced LDS
// strategy may go up to it; the automatic choice keeps to kGroupLdsBudget).
int64_t group_lds_max() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
    return kGroupLdsBudget;
  return v;
}

// LDS bytes per (row, label) slot of a pass.
int64_t pass_slot_bytes(int pass, bool count, int64_t sa_bytes) {
  switch (pass) {
    case kPSum: return sa_bytes + (count ? 4 : 0);
    case kPDev: return 16;
    case kPKeyIdx: return 16;
    case kPCount:
    case kPAny:
    case kPAll: return 4;
    default: return 8;
  }
}

int env_int(const char* name, int dflt) { return (int)knob(name, dflt); }

// Workgroups per CU of the LDS strategy's grid. Every workgroup ends with one
// global atomic per used label slot, and all of them end together, so the
// grid is kept at the fewest workgroups that still keep enough loads in
// flight: 2 for the sum passes (2 rows x 4 vectors per thread) and the
// others (4 rows x 4 vectors per thread), 4 for the deviation and moment
// passes. Measured on MI355X for (64, 4M) float32 x 1000 labels, launches
// back to back: group_nansum with uniform labels 0.180 ms at 2 workgroups per
// CU and 4 vectors against 0.200 ms at 4 and 2 vectors (Zipf(1.1) labels:
// 0.201 against 0.192 ms), count/argmax/first/last 5-8% faster at 2 than at
// the occupancy limit. Knob: NBG_GROUP_BLOCKS_PER_CU (0 = the occupancy
// limit); the autotuner may pick another value per shape. Results do not
// depend on it.
constexpr int group_blocks_per_cu(int pass) { return (pass == kPDev || pass == kPMom) ? 4 : 2; }

// Chunking: the row groups' positions are split so that the whole grid is
// one wave of resident workgroups (the device's capacity for this kernel and
// LDS size, capped per CU for the LDS strategy): no partly filled second
// wave, and each workgroup flushes its accumulators once. Chunks are
// multiples of 4096 positions (whole vector super-steps) and at least 4096.
template <typename TV, typename TL, int PASS, bool LDS, int R, bool COUNT, int U = kGroupUnroll,
          bool PIPE = false, bool HOT = false>
int launch_pass(GArgs a, size_t lds, hipStream_t s) {
  if (a.n == 0) return NBG_OK;  // nothing to accumulate: init and finalize give the empty-label outputs
  const int64_t groups = (a.rows + R - 1) / R;
  int cap = device_capacity_blocks((const void*)group_pass_kernel<TV, TL, PASS, LDS, R, COUNT, U, PIPE, HOT>,
                                   kGroupThreads, LDS ? lds : 0);
  if (cap <= 0) cap = 2048;
  if (LDS) {
    const int per_cu = env_int("NBG_GROUP_BLOCKS_PER_CU", group_blocks_per_cu(PASS));
    int dev = 0, cus = 0;
    if (per_cu > 0 && hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && cus > 0 &&
        cap > per_cu * cus)
      cap = per_cu * cus;
  }
  const int64_t per_group = cap / groups > 1 ? cap / groups : 1;
  int64_t chunk = (a.n + per_group - 1) / per_group;
  chunk = (chunk + 4095) / 4096 * 4096;
  a.chunk = chunk;
  a.chunks_per_row = (a.n + a.chunk - 1) / a.chunk;
  const int64_t blocks = groups * a.chunks_per_row;
  if (blocks > 0x7fffffffLL) return set_error(NBG_ERR_INVALID, "group: grid too large");
  hipLaunchKernelGGL((group_pass_kernel<TV, TL, PASS, LDS, R, COUNT, U, PIPE, HOT>), dim3((unsigned)blocks),
                     dim3(kGroupThreads), LDS ? lds : 0, s, a);
  return check_hip(hipGetLastError(), "group_pass_kernel launch");
}

// Rows per workgroup for the LDS strategy: 4 when four rows' accumulators fit
// the LDS budget (the rows then share each label load), else 2, else 1.
constexpr int kGroupRowsPerBlock = 4;

// Shape of the sum pass (sum, mean, sum of squares, and the first pass of
// var/std on global accumulators): at most 2 rows per workgroup and 4
// vectors per thread and super-step, with 2 workgroups per CU (see
// group_blocks_per_cu). Knobs: NBG_GROUP_ROWS (1, 2, 4), NBG_GROUP_UNROLL (2,
// 4), NBG_GROUP_PIPE=1 (load the next row while folding the current one).
// Results do not depend on them.
constexpr int kSumRows = 2;
constexpr int kSumUnroll = 4;
// Hot labels 0 and 1 in registers for the float sum passes (see
// group_pass_kernel). Knob: NBG_GROUP_HOT=0/1.
constexpr int kGroupHotDefault = 1;

template <typename TV, typename TL, int PASS, bool COUNT, int R>
int launch_rows(GArgs a, size_t lds, hipStream_t s) {
  constexpr bool HOT_OK = std::is_floating_point<TV>::value &&
                         (PASS == kPSum || PASS == kPDev || PASS == kPMom || PASS == kPSumSq || PASS == kPProd);
  const bool hot = HOT_OK && env_int("NBG_GROUP_HOT", kGroupHotDefault) == 1;
  if constexpr (PASS == kPSum) {
    const int u = env_int("NBG_GROUP_UNROLL", kSumUnroll);
    const bool pipe = env_int("NBG_GROUP_PIPE", 0) == 1;
    if (pipe) {
      if (u == 2) return launch_pass<TV, TL, PASS, true, R, COUNT, 2, true>(a, lds, s);
      return launch_pass<TV, TL, PASS, true, R, COUNT, 4, true>(a, lds, s);
    }
    if (u == 2) {
      if constexpr (HOT_OK)
        if (hot) return launch_pass<TV, TL, PASS, true, R, COUNT, 2, false, true>(a, lds, s);
      return launch_pass<TV, TL, PASS, true, R, COUNT, 2, false>(a, lds, s);
    }
  }
  if constexpr (HOT_OK)
    if (hot) return launch_pass<TV, TL, PASS, true, R, COUNT, kGroupUnroll, false, true>(a, lds, s);
  return launch_pass<TV, TL, PASS, true, R, COUNT>(a, lds, s);
}

template <typename TV, typename TL, int PASS, bool COUNT = true>
int run_pass(GArgs a, bool lds, hipStream_t s) {
  const int64_t L = a.num_labels;
  // kPMom: two float64 sums, a shift in the values dtype and a count
  const int64_t slot = PASS == kPMom ? 20 + (int64_t)sizeof(TV)
                                     : pass_slot_bytes(PASS, COUNT, sizeof(typename SumAcc<TV>::type));
  if constexpr (PASS != kPMom) {  // shifted moments run in LDS only
    if (!lds) return launch_pass<TV, TL, PASS, false, 1, COUNT>(a, 0, s);
  }
  const int rmax = PASS == kPSum ? env_int("NBG_GROUP_ROWS", kSumRows) : kGroupRowsPerBlock;
  auto bytes = [&](int64_t r) { return L * r * slot; };
  if (rmax >= 4 && a.rows >= 4 && bytes(4) <= kGroupLdsBudget)
    return launch_rows<TV, TL, PASS, COUNT, 4>(a, (size_t)bytes(4), s);
  if (rmax >= 2 && a.rows >= 2 && bytes(2) <= kGroupLdsBudget)
    return launch_rows<TV, TL, PASS, COUNT, 2>(a, (size_t)bytes(2), s);
  return launch_rows<TV, TL, PASS, COUNT, 1>(a, (size_t)bytes(1), s);
}

int launch_init(void* acc0, unsigned long long init0, void* acc1, unsigned long long init1,
                int64_t count, hipStream_t s) {
  int64_t blocks = (count + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(group_init_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     (unsigned long long*)acc0, init0, (unsigned long long*)acc1, init1, count);
  return check_hip(hipGetLastError(), "group_init_kernel launch");
}

template <typename TV, int KIND>
int launch_finalize(GArgs a, double* means_out, hipStream_t s) {
  const int64_t total = a.rows * a.num_labels;
  int64_t blocks = (total + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((group_finalize_kernel<TV, KIND>), dim3((unsigned)blocks), dim3(256), 0, s, a,
                     means_out);
  return check_hip(hipGetLastError(), "group_finalize_kernel launch");
}

constexpr bool float_only_kind(int kind) { return kind == kGMean || kind == kGVar || kind == kGStd; }

int64_t lds_bytes_per_label(int kind, int vdtype) {
  switch (kind) {
    case kGSum: return 12;
    case kGMean: return 12;
    case kGCount: return 4;
    case kGVar:
    case kGStd: return 28;  // shifted-moment slots (the two-pass fallback needs 16)
    case kGArgMax:
    case kGArgMin:
    case kGMin:
    case kGMax: return 16;
    case kGAny:
    case kGAll: return 4;
    default: return 8;
  }
  (void)vdtype;
}

// Workspace layout: acc0 | acc1 | acc2 | acc3 (rows x num_labels x 8 bytes each).
size_t group_ws_bytes(int64_t rows, int64_t L) { return (size_t)(4 * rows * L) * 8; }

template <typename TV, typename TL, int KIND>
int group_impl(GArgs a, int strategy, void* ws, size_t ws_bytes, hipStream_t s) {
  const int64_t L = a.num_labels;
  const size_t need = group_ws_bytes(a.rows, L);
  if (ws == nullptr || ws_bytes < need)
    return set_error(NBG_ERR_INVALID, "group: workspace too small (%zu < %zu)", ws_bytes, need);
  const int64_t per = (int64_t)a.rows * L;
  unsigned long long* acc0 = static_cast<unsigned long long*>(ws);
  unsigned long long* acc1 = acc0 + per;
  unsigned long long* acc2 = acc1 + per;
  unsigned long long* acc3 = acc2 + per;
  a.acc2 = nullptr;
  const int64_t lds_need = L * lds_bytes_per_label(KIND, 0);
  bool lds = lds_need <= kGroupLdsBudget;
  if (strategy == NBG_GROUP_LDS && !lds) {
    // forced: one row per workgroup, up to the device's LDS per workgroup
    if (lds_need > group_lds_max())
      return set_error(NBG_ERR_INVALID, "group: %lld labels do not fit the LDS strategy", (long long)L);
    lds = true;
  }
  if (strategy == NBG_GROUP_GLOBAL) lds = false;

  a.means = nullptr;
  a.argmin = (KIND == kGArgMin || KIND == kGMin);
  int rc = 0;
  const bool packed = a.n < 0xffffffffLL;  // indices fit the low 32 bits
  switch (KIND) {
    case kGSum:
      if ((rc = launch_init(acc0, 0ull, nullptr, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = nullptr;
      if ((rc = run_pass<TV, TL, kPSum, false>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGMean:
      if ((rc = launch_init(acc0, 0ull, acc1, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = acc1;
      if ((rc = run_pass<TV, TL, kPSum>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGCount:
      if ((rc = launch_init(acc0, 0ull, nullptr, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = nullptr;
      if ((rc = run_pass<TV, TL, kPCount>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGSumSq:
      if ((rc = launch_init(acc0, 0ull, nullptr, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = nullptr;
      if ((rc = run_pass<TV, TL, kPSumSq>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGProd:
      if ((rc = launch_init(acc0, one_bits<typename SumAcc<TV>::type>(), nullptr, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = nullptr;
      if ((rc = run_pass<TV, TL, kPProd>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGAny:
    case kGAll:
      if ((rc = launch_init(acc0, 0ull, nullptr, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = nullptr;
      if ((rc = run_pass<TV, TL, KIND == kGAny ? kPAny : kPAll>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGArgMax:
    case kGArgMin:
    case kGMin:
    case kGMax:
      if constexpr (sizeof(TV) == 4) {
        if (packed) {
          if ((rc = launch_init(acc0, 0ull, nullptr, 0ull, per, s))) return rc;
          a.acc0 = acc0;
          a.acc1 = nullptr;
          if ((rc = run_pass<TV, TL, kPPacked>(a, lds, s))) return rc;
          return launch_finalize<TV, KIND>(a, nullptr, s);
        }
      }
      if ((rc = launch_init(acc0, 0ull, acc1, ~0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = acc1;
      if ((rc = run_pass<TV, TL, kPKey>(a, lds, s))) return rc;
      if ((rc = run_pass<TV, TL, kPKeyIdx>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGFirst:
      if ((rc = launch_init(acc0, ~0ull, nullptr, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = nullptr;
      if ((rc = run_pass<TV, TL, kPFirst>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    case kGLast:
      if ((rc = launch_init(acc0, 0ull, nullptr, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = nullptr;
      if ((rc = run_pass<TV, TL, kPLast>(a, lds, s))) return rc;
      return launch_finalize<TV, KIND>(a, nullptr, s);
    default: {  // var / std
      if constexpr (std::is_floating_point<TV>::value) {
        if (lds) {
          // One pass: per-(row, label) shifted moments in LDS, moved to a
          // global shift per (row, label) when each workgroup flushes them.
          if ((rc = launch_init(acc0, 0ull, acc1, 0ull, per, s))) return rc;
          if ((rc = launch_init(acc2, 0ull, acc3, 0x7FF8000000000000ull, per, s))) return rc;
          a.acc0 = acc0;
          a.acc1 = acc1;
          a.acc2 = acc2;
          a.acc3 = acc3;
          a.mom = 1;
          if ((rc = run_pass<TV, TL, kPMom>(a, true, s))) return rc;
          return launch_finalize<TV, KIND>(a, nullptr, s);
        }
      }
      // Global accumulators: per-label means, then squared deviations.
      // The float64 means come from float64 sums, so the deviations' own sum
      // (the usual two-pass correction term) is negligible: with n values
      // per label its share of sum d^2 is about (n 2**-53 |mean| / std)^2.
      if ((rc = launch_init(acc0, 0ull, acc1, 0ull, per, s))) return rc;
      a.acc0 = acc0;
      a.acc1 = acc1;
      if ((rc = run_pass<TV, TL, kPSum>(a, lds, s))) return rc;
      double* means = reinterpret_cast<double*>(acc2);
      if ((rc = launch_finalize<TV, KIND>(a, means, s))) return rc;
      GArgs b = a;
      b.means = means;
      b.acc1 = acc3;  // sum of squared deviations
      b.acc2 = acc1;  // counts from the first pass
      if ((rc = launch_init(acc3, 0ull, nullptr, 0ull, per, s))) return rc;
      if ((rc = run_pass<TV, TL, kPDev>(b, lds, s))) return rc;
      return launch_finalize<TV, KIND>(b, nullptr, s);
    }
  }
}

template <int KIND>
int group_entry(const char* name, int vdtype, const void* values, int ldtype, const void* labels,
                void* out, int64_t rows, int64_t n, int64_t ld_v, int64_t num_labels, int64_t ddof,
                int strategy, void* ws, size_t ws_bytes, void* stream) {
  if (int rc = injected_error(name)) return rc;
  if (rows < 0 || n < 0 || num_labels < 0) return set_error(NBG_ERR_INVALID, "%s: negative shape", name);
  if (strategy < NBG_GROUP_AUTO || strategy > NBG_GROUP_GLOBAL)
    return set_error(NBG_ERR_INVALID, "%s: unknown strategy %d", name, strategy);
  if (rows == 0 || num_labels == 0) return NBG_OK;
  if (values == nullptr && n > 0) return set_error(NBG_ERR_INVALID, "%s: null values", name);
  if ((labels == nullptr && n > 0) || out == nullptr) return set_error(NBG_ERR_INVALID, "%s: null pointer", name);
  if (rows > 1 && ld_v < n) return set_error(NBG_ERR_INVALID, "%s: row stride < n", name);
  const bool float_only = (KIND == kGMean || KIND == kGVar || KIND == kGStd);
  if (float_only && vdtype != NBG_F32 && vdtype != NBG_F64)
    return set_error(NBG_ERR_INVALID, "%s: values must be float32 or float64", name);
  GArgs a{};
  a.values = values;
  a.labels = labels;
  a.out = out;
  a.rows = rows;
  a.n = n;
  a.ld_v = ld_v;
  a.num_labels = num_labels;
  a.ddof = ddof;
  hipStream_t s = static_cast<hipStream_t>(stream);
#define NBG_G(TV, TL) return group_impl<TV, TL, KIND>(a, strategy, ws, ws_bytes, s)
  if (ldtype == NBG_I32) {
    switch (vdtype) {
      case NBG_F32: NBG_G(float, int32_t);
      case NBG_F64: NBG_G(double, int32_t);
      case NBG_I32: if constexpr (!float_only_kind(KIND)) NBG_G(int32_t, int32_t); break;
      case NBG_I64: if constexpr (!float_only_kind(KIND)) NBG_G(int64_t, int32_t); break;
    }
  } else if (ldtype == NBG_I64) {
    switch (vdtype) {
      case NBG_F32: NBG_G(float, int64_t);
      case NBG_F64: NBG_G(double, int64_t);
      case NBG_I32: if constexpr (!float_only_kind(KIND)) NBG_G(int32_t, int64_t); break;
      case NBG_I64: if constexpr (!float_only_kind(KIND)) NBG_G(int64_t, int64_t); break;
    }
  } else {
    return set_error(NBG_ERR_INVALID, "%s: labels must be int32 or int64 (code %d)", name, ldtype);
  }
#undef NBG_G
  return set_error(NBG_ERR_INVALID, "%s: unsupported values dtype code %d", name, vdtype);
}

}  // namespace

// ---- allocation-free variant of group_nansum --------------------------------
// Scratch of nbg_group_nansum_noalloc: one static array per device (each GPU
// the library runs on holds it until the process ends). Host code never uses
// the array's name as an address: the per-device address comes from
// hipGetSymbolAddress with that device current.
static __device__ __align__(256) unsigned char scratch[NBG_GROUP_SCRATCH_BYTES];

// Device address of the current device's copy of `scratch`, resolved on the
// first call on each device; nullptr (with the error recorded) on failure.
static void* group_scratch_ptr() {
  static std::mutex mu;
  static std::unordered_map<int, void*> ptrs;
  int dev = 0;
  if (check_hip(hipGetDevice(&dev), "hipGetDevice") != NBG_OK) return nullptr;
  std::lock_guard<std::mutex> g(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  void* ptr = nullptr;
  if (check_hip(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)") != NBG_OK)
    return nullptr;
  ptrs[dev] = ptr;
  return ptr;
}

}  // namespace nbg

extern "C" {

int nbg_group_nansum_noalloc(int vdtype, const void* values, int ldtype, const void* labels, void* out,
                             int64_t rows, int64_t n, int64_t ld_v, int64_t num_labels, int strategy,
                             void* stream) {
  if (rows > 0 && num_labels > 0) {
    // rows * num_labels bounded first, so the size query cannot overflow
    const size_t need = (rows > (int64_t)NBG_GROUP_SCRATCH_BYTES || num_labels > (int64_t)NBG_GROUP_SCRATCH_BYTES)
                            ? (size_t)-1
                            : nbg::group_ws_bytes(rows, num_labels);
    if (rows * num_labels > (int64_t)NBG_GROUP_SCRATCH_BYTES || need > (size_t)NBG_GROUP_SCRATCH_BYTES)
      return nbg::set_error(NBG_ERR_INVALID,
                            "nbg_group_nansum_noalloc: %lld rows x %lld labels need more than the %d bytes "
                            "of static scratch",
                            (long long)rows, (long long)num_labels, NBG_GROUP_SCRATCH_BYTES);
  }
  void* ptr = nbg::group_scratch_ptr();
  if (ptr == nullptr) return NBG_ERR_HIP;
  return nbg::group_entry<nbg::kGSum>("nbg_group_nansum_noalloc", vdtype, values, ldtype, labels, out, rows, n,
                                      ld_v, num_labels, 0, strategy, ptr, (size_t)NBG_GROUP_SCRATCH_BYTES, stream);
}

size_t nbg_group_workspace_bytes(int vdtype, int64_t rows, int64_t n, int64_t num_labels,
                                 int strategy) {
  (void)vdtype;
  (void)n;
  (void)strategy;
  if (rows <= 0 || num_labels <= 0) return 0;
  return nbg::group_ws_bytes(rows, num_labels);
}

#define NBG_GROUP(NAME, KIND)                                                                  \
  int NAME(int vdtype, const void* values, int ldtype, const void* labels, void* out,          \
           int64_t rows, int64_t n, int64_t ld_v, int64_t num_labels, int strategy, void* ws,  \
           size_t ws_bytes, void* stream) {                                                    \
    return nbg::group_entry<KIND>(#NAME, vdtype, values, ldtype, labels, out, rows, n, ld_v,   \
                                  num_labels, 0, strategy, ws, ws_bytes, stream);              \
  }
#define NBG_GROUP_DDOF(NAME, KIND)                                                             \
  int NAME(int vdtype, const void* values, int ldtype, const void* labels, void* out,          \
           int64_t rows, int64_t n, int64_t ld_v, int64_t num_labels, int64_t ddof,            \
           int strategy, void* ws, size_t ws_bytes, void* stream) {                            \
    return nbg::group_entry<KIND>(#NAME, vdtype, values, ldtype, labels, out, rows, n, ld_v,   \
                                  num_labels, ddof, strategy, ws, ws_bytes, stream);           \
  }

NBG_GROUP(nbg_group_nansum, nbg::kGSum)
NBG_GROUP(nbg_group_nanmean, nbg::kGMean)
NBG_GROUP(nbg_group_nancount, nbg::kGCount)
NBG_GROUP(nbg_group_nanargmax, nbg::kGArgMax)
NBG_GROUP(nbg_group_nanargmin, nbg::kGArgMin)
NBG_GROUP(nbg_group_nanfirst, nbg::kGFirst)
NBG_GROUP(nbg_group_nanlast, nbg::kGLast)
NBG_GROUP_DDOF(nbg_group_nanvar, nbg::kGVar)
NBG_GROUP_DDOF(nbg_group_nanstd, nbg::kGStd)
NBG_GROUP(nbg_group_nanmin, nbg::kGMin)
NBG_GROUP(nbg_group_nanmax, nbg::kGMax)
NBG_GROUP(nbg_group_nanprod, nbg::kGProd)
NBG_GROUP(nbg_group_nansum_of_squares, nbg::kGSumSq)
NBG_GROUP(nbg_group_nanany, nbg::kGAny)
NBG_GROUP(nbg_group_nanall, nbg::kGAll)

}  // extern "C"
