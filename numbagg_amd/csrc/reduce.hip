// Plain reductions over one axis for gfx950: nansum, nanmean, nancount,
// allnan, anynan, nanvar, nanstd (ndaggregate, numbagg/funcs.py:23-159,
// numbagg/decorators.py:188-272) and nanmax, nanmin, nanargmax, nanargmin
// (ndreduce, numbagg/funcs.py:160-228, numbagg/decorators.py:904-1028).
//
// The host moves the reduced axes last and flattens them, so every kernel sees
// a row-major (rows, n) view with a unit-stride core axis and reduces each
// row to one output.
//
// Design: one wavefront reduces one (row, chunk) item. Rows are split into
// chunks so that a launch holds enough wavefronts to fill the device; each
// lane streams 16-byte vectors of its chunk (strided by the wave), keeps a
// private state, and the wave merges its 64 lane states in a fixed butterfly
// order and writes one 32-byte partial. A finalize kernel merges each row's
// chunk partials in chunk order and writes the output. No atomics: results
// are the same bits on every run.
//
// States: float32 sums in float64 (exact differences of 24-bit significands
// stay far inside float64), float64 sums in double-double, integer sums
// wrapping mod 2**64 (narrowed like the reference's fixed-width adds); the
// variance keeps (count, mean, M2) per lane, updated batch by batch with the
// batch's own two-pass moments and Chan's merge (the reference makes two
// passes over the data, numbagg/funcs.py:108-127); extremes as orderable
// 64-bit keys with the index of the first (arg*: the reference replaces on a
// strict `>` / `<`) or last (min/max: it replaces on `>=` / `<=`) occurrence.
#include <type_traits>

#include "common.hpp"

namespace nbg {
namespace {

constexpr int kRedThreads = 256;
constexpr int kRedWaves = kRedThreads / kWave;

enum RKind {
  kRSum = 0,
  kRMean = 1,
  kRCount = 2,
  kRAllNan = 3,
  kRAnyNan = 4,
  kRVar = 5,
  kRStd = 6,
  kRMax = 7,
  kRMin = 8,
  kRArgMax = 9,
  kRArgMin = 10,
};

constexpr bool is_var_kind(int k) { return k == kRVar || k == kRStd; }
constexpr bool is_ext_kind(int k) { return k == kRMax || k == kRMin || k == kRArgMax || k == kRArgMin; }
constexpr bool first_wins(int k) { return k == kRArgMax || k == kRArgMin; }

// Partial state of one (row, chunk): 32 bytes.
//   sum/mean: w0 sum (float64 bits, or wrapping u64 for integers), w1 the
//             double-double low part (float64 data), w2 count, w3 the sum of
//             the infinite values (float64 data: +-inf or NaN, else 0);
//   count/allnan/anynan: w2 count of valid values, w3 count of NaNs;
//   var/std: w0 mean, w1 M2 (float64 bits), w2 count;
//   extremes: w0 key (0 = no valid value), w1 index.
struct RPart {
  unsigned long long w0, w1, w2, w3;
};

__device__ __forceinline__ double f64_of(unsigned long long b) { return __longlong_as_double((long long)b); }
__device__ __forceinline__ unsigned long long bits_of(double v) {
  return (unsigned long long)__double_as_longlong(v);
}

// Orderable key of a value compared as float64 (the reference's running
// extremes are float64, also for integer input); -0.0 and +0.0 tie. Maximum
// kinds use the key, minimum kinds its complement; every non-NaN value maps
// to a non-zero word, so 0 stands for "no valid value".
template <typename T>
__device__ __forceinline__ unsigned long long ext_key(T v, bool minimum) {
  const unsigned long long b = bits_of((double)v + 0.0);
  const unsigned long long k = (b & 0x8000000000000000ull) ? ~b : (b | 0x8000000000000000ull);
  return minimum ? ~k : k;
}

// (count, mean, M2) and Chan's merge of another such triple.
struct Chan {
  double n, mean, m2;
  __device__ __forceinline__ void merge(double nb, double mb, double m2b) {
    if (nb == 0.0) return;
    const double tot = n + nb;
    const double d = mb - mean;
    const double f = nb / tot;
    mean += d * f;
    m2 += m2b + d * d * n * f;
    n = tot;
  }
};

// float32 extremes keep 32-bit keys, the best value and a 32-bit index
// within the chunk per lane (chunks are below 2**31 positions); the 64-bit
// key and the absolute index are formed once, when the lane's part is taken.
template <typename T, int KIND>
constexpr bool small_ext() {
  return is_ext_kind(KIND) && sizeof(T) == 4 && std::is_floating_point<T>::value;
}

// Orderable 32-bit key of a float32 value (-0 as +0): larger values give
// larger keys (complemented for minima); every non-NaN value's key is nonzero.
__device__ __forceinline__ uint32_t ext_key32(float v, bool minimum) {
  const uint32_t b = __float_as_uint(v + 0.0f);
  const uint32_t k = (b & 0x80000000u) ? ~b : (b | 0x80000000u);
  return minimum ? ~k : k;
}

template <typename T, int KIND>
struct Lane {
  double s = 0.0;                                 // float32 data: float64 sum
  Comp<double> cs{0.0, 0.0};                      // float64 data: double-double sum of finite values
  double sinf = 0.0;                              // float64 data: sum of the infinite values
  unsigned long long us = 0ull;                   // integers: wrapping sum
  unsigned long long cnt = 0ull, nans = 0ull;
  Chan ch = Chan{0.0, 0.0, 0.0};
  unsigned long long key = 0ull, idx = 0ull;
  uint32_t key32 = 0u, idx32 = 0u;  // small_ext: key (0: none yet), index within the chunk
  T best = T(0);                    // small_ext: the value of key32

  // small_ext: one element at position rel within the chunk, by selects (no
  // branch per element).
  __device__ __forceinline__ void fold32(T v, uint32_t rel) {
    const uint32_t k = ext_key32((float)v, KIND == kRMin || KIND == kRArgMin);
    const bool take = !is_nan(v) && (first_wins(KIND) ? k > key32 : k >= key32);
    key32 = take ? k : key32;
    idx32 = take ? rel : idx32;
    best = take ? v : best;
  }

  // One element at position i (non-variance kinds).
  __device__ __forceinline__ void fold(T v, int64_t i) {
    const bool valid = !is_nan(v);
    if constexpr (KIND == kRSum || KIND == kRMean) {
      if constexpr (!std::is_floating_point<T>::value) {
        us += (unsigned long long)(long long)v;
      } else if constexpr (sizeof(T) == 8) {
        // infinities stay out of the double-double pair, whose error terms
        // would turn them into NaN
        if (valid) {
          if (isfinite((double)v)) cs.add((double)v);
          else sinf += (double)v;
        }
      } else {
        s += valid ? (double)v : 0.0;
      }
      cnt += valid ? 1ull : 0ull;
    } else if constexpr (KIND == kRCount || KIND == kRAllNan || KIND == kRAnyNan) {
      cnt += valid ? 1ull : 0ull;
      nans += valid ? 0ull : 1ull;
    } else if constexpr (is_ext_kind(KIND)) {
      if (valid) {
        const unsigned long long k = ext_key(v, KIND == kRMin || KIND == kRArgMin);
        // a lane visits increasing positions: strictly larger keys replace
        // (first occurrence), or larger-or-equal keys (last occurrence)
        const bool take = first_wins(KIND) ? k > key : k >= key;
        if (take) {
          key = k;
          idx = (unsigned long long)i;
        }
      }
    }
  }

  // A batch of B elements (variance kinds): its own two-pass moments, merged.
  template <int B>
  __device__ __forceinline__ void fold_batch(const T (&v)[B]) {
    double sb = 0.0;
    int nb = 0;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const bool valid = !is_nan(v[j]);
      sb += valid ? (double)v[j] : 0.0;
      nb += valid ? 1 : 0;
    }
    if (nb == 0) return;
    const double mb = sb / (double)nb;
    double m2b = 0.0;
#pragma unroll
    for (int j = 0; j < B; ++j) {
      const double d = (double)v[j] - mb;
      m2b += !is_nan(v[j]) ? d * d : 0.0;
    }
    ch.merge((double)nb, mb, m2b);
  }

  // c0: the chunk's first position (small_ext indices are relative to it).
  __device__ __forceinline__ RPart to_part(int64_t c0) const {
    RPart p{0ull, 0ull, 0ull, 0ull};
    if constexpr (small_ext<T, KIND>()) {
      if (key32 != 0u) {
        p.w0 = ext_key(best, KIND == kRMin || KIND == kRArgMin);
        p.w1 = (unsigned long long)(c0 + (int64_t)idx32);
      }
      return p;
    }
    if constexpr (KIND == kRSum || KIND == kRMean) {
      if constexpr (!std::is_floating_point<T>::value) {
        p.w0 = us;
      } else if constexpr (sizeof(T) == 8) {
        p.w0 = bits_of(cs.hi);
        p.w1 = bits_of(cs.lo);
        p.w3 = bits_of(sinf);
      } else {
        p.w0 = bits_of(s);
      }
      p.w2 = cnt;
    } else if constexpr (KIND == kRCount || KIND == kRAllNan || KIND == kRAnyNan) {
      p.w2 = cnt;
      p.w3 = nans;
    } else if constexpr (is_var_kind(KIND)) {
      p.w0 = bits_of(ch.mean);
      p.w1 = bits_of(ch.m2);
      p.w2 = (unsigned long long)ch.n;
    } else {
      p.w0 = key;
      p.w1 = idx;
    }
    return p;
  }
};

// Merge partial b (a later or any other part of the row) into a.
template <typename T, int KIND>
__device__ __forceinline__ void merge_part(RPart& a, const RPart& b) {
  if constexpr (KIND == kRSum || KIND == kRMean) {
    if constexpr (!std::is_floating_point<T>::value) {
      a.w0 += b.w0;
    } else if constexpr (sizeof(T) == 8) {
      Comp<double> x{f64_of(a.w0), f64_of(a.w1)};
      x.add(Comp<double>{f64_of(b.w0), f64_of(b.w1)});
      a.w0 = bits_of(x.hi);
      a.w1 = bits_of(x.lo);
      a.w3 = bits_of(f64_of(a.w3) + f64_of(b.w3));
    } else {
      a.w0 = bits_of(f64_of(a.w0) + f64_of(b.w0));
    }
    a.w2 += b.w2;
  } else if constexpr (KIND == kRCount || KIND == kRAllNan || KIND == kRAnyNan) {
    a.w2 += b.w2;
    a.w3 += b.w3;
  } else if constexpr (is_var_kind(KIND)) {
    Chan c{(double)a.w2, f64_of(a.w0), f64_of(a.w1)};
    c.merge((double)b.w2, f64_of(b.w0), f64_of(b.w1));
    a.w0 = bits_of(c.mean);
    a.w1 = bits_of(c.m2);
    a.w2 = a.w2 + b.w2;
  } else {
    // larger key wins; equal keys: the smaller (first) or larger (last) index
    const bool take = b.w0 > a.w0 || (b.w0 == a.w0 && b.w0 != 0ull &&
                                      (first_wins(KIND) ? b.w1 < a.w1 : b.w1 > a.w1));
    if (take) {
      a.w0 = b.w0;
      a.w1 = b.w1;
    }
  }
}

__device__ __forceinline__ RPart shfl_xor_part(const RPart& p, int m) {
  RPart q;
  q.w0 = (unsigned long long)__shfl_xor((long long)p.w0, m, kWave);
  q.w1 = (unsigned long long)__shfl_xor((long long)p.w1, m, kWave);
  q.w2 = (unsigned long long)__shfl_xor((long long)p.w2, m, kWave);
  q.w3 = (unsigned long long)__shfl_xor((long long)p.w3, m, kWave);
  return q;
}

// One (row, chunk) item per wavefront: item = block * 4 + wave.
template <typename T, int KIND, int VEC, int U>
__global__ __launch_bounds__(kRedThreads) void reduce_kernel(const T* __restrict__ x, int64_t n, int64_t ld,
                                                             int64_t chunk, int64_t chunks_per_row,
                                                             int64_t items, RPart* __restrict__ parts) {
  const int64_t item = (int64_t)blockIdx.x * kRedWaves + wave_id();
  if (item >= items) return;  // whole wavefronts leave; no barriers below
  const int lane = lane_id();
  const int64_t row = item / chunks_per_row;
  const int64_t ck = item - row * chunks_per_row;
  const int64_t c0 = ck * chunk;
  const int64_t c1 = c0 + chunk < n ? c0 + chunk : n;
  const T* xr = x + row * ld;
  Lane<T, KIND> st;
  constexpr int64_t STEP = (int64_t)kWave * VEC;
  constexpr int64_t SUPER = STEP * U;
  struct alignas(VEC * sizeof(T)) V {
    T v[VEC];
  };
  int64_t e = c0;
  if (VEC > 1) {
    // full super-steps: U vectors per lane, all loads issued first
    for (; e + SUPER <= c1; e += SUPER) {
      V vv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) vv[u] = load_stream(reinterpret_cast<const V*>(xr + e + u * STEP + lane * VEC));
      if constexpr (is_var_kind(KIND)) {
        T b[U * VEC];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < VEC; ++j) b[u * VEC + j] = vv[u].v[j];
        st.template fold_batch<U * VEC>(b);
      } else if constexpr (small_ext<T, KIND>()) {
        const uint32_t rel = (uint32_t)(e - c0) + (uint32_t)(lane * VEC);
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < VEC; ++j) st.fold32(vv[u].v[j], rel + (uint32_t)(u * STEP + j));
      } else {
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
          for (int j = 0; j < VEC; ++j) st.fold(vv[u].v[j], e + u * STEP + lane * VEC + j);
      }
    }
  }
  // remainder (or the whole chunk for unaligned rows): one element per lane
  for (int64_t i = e + lane; i < c1; i += kWave) {
    if constexpr (is_var_kind(KIND)) {
      const T b[1] = {xr[i]};
      st.template fold_batch<1>(b);
    } else if constexpr (small_ext<T, KIND>()) {
      st.fold32(xr[i], (uint32_t)(i - c0));
    } else {
      st.fold(xr[i], i);
    }
  }
  RPart p = st.to_part(c0);
#pragma unroll
  for (int m = 1; m < kWave; m <<= 1) {
    const RPart q = shfl_xor_part(p, m);
    // merge in position order: the lower lane's state first
    if (lane & m) {
      RPart lo = q;
      merge_part<T, KIND>(lo, p);
      p = lo;
    } else {
      merge_part<T, KIND>(p, q);
    }
  }
  if (lane == 0) parts[item] = p;
}

// The output of one row from its merged partial.
template <typename T, typename TO, int KIND>
__device__ __forceinline__ TO part_value(const RPart& p, int64_t row, int64_t ddof, const T* __restrict__ x,
                                         int64_t ld) {
  TO r;
  if constexpr (KIND == kRSum) {
    if constexpr (!std::is_floating_point<T>::value) {
      r = (TO)(T)(long long)p.w0;  // modular narrowing
    } else {
      const double inf = f64_of(p.w3);
      r = (TO)(inf != 0.0 || inf != inf ? inf : f64_of(p.w0) + f64_of(p.w1));
    }
  } else if constexpr (KIND == kRMean) {
    const double inf = f64_of(p.w3);
    const double s = inf != 0.0 || inf != inf ? inf : f64_of(p.w0) + f64_of(p.w1);
    r = p.w2 ? (TO)(s / (double)p.w2) : (TO)canonical_nan<double>();
  } else if constexpr (KIND == kRCount) {
    r = (TO)p.w2;
  } else if constexpr (KIND == kRAllNan) {
    r = (TO)(p.w2 == 0 ? 1 : 0);
  } else if constexpr (KIND == kRAnyNan) {
    r = (TO)(p.w3 != 0 ? 1 : 0);
  } else if constexpr (is_var_kind(KIND)) {
    const long long c = (long long)p.w2;
    if (c > ddof) {
      const double var = f64_of(p.w1) / (double)(c - ddof);
      r = (TO)(KIND == kRStd ? sqrt(var) : var);
    } else {
      r = (TO)canonical_nan<double>();
    }
  } else if constexpr (KIND == kRArgMax || KIND == kRArgMin) {
    r = p.w0 ? (TO)p.w1 : (TO)-1;  // -1: no valid value (the host raises)
  } else {
    // the value at the extreme's index, through float64 as the reference's
    // float64 running extreme returns it (64-bit integers round there)
    if (p.w0 == 0ull) r = (TO)canonical_nan<double>();
    else r = (TO)(double)x[row * ld + (int64_t)p.w1];
  }
  return r;
}

// Output of one row from its chunk partials (merged in chunk order), a thread
// per row.
template <typename T, typename TO, int KIND>
__global__ void reduce_finalize_kernel(const RPart* __restrict__ parts, int64_t rows, int64_t chunks_per_row,
                                       int64_t ddof, const T* __restrict__ x, int64_t ld, TO* __restrict__ out) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= rows) return;
  RPart p = parts[row * chunks_per_row];
  for (int64_t c = 1; c < chunks_per_row; ++c) merge_part<T, KIND>(p, parts[row * chunks_per_row + c]);
  out[row] = part_value<T, TO, KIND>(p, row, ddof, x, ld);
}

// The same for rows of many chunks, a workgroup per row: each thread merges a
// contiguous run of chunks, then a pairwise tree over the threads in order
// (left before right, so every merge keeps chunk order; deterministic).
constexpr int kRedFinishThreads = 256;

template <typename T, typename TO, int KIND>
__global__ __launch_bounds__(kRedFinishThreads) void reduce_finalize_wide_kernel(
    const RPart* __restrict__ parts, int64_t chunks_per_row, int64_t ddof, const T* __restrict__ x, int64_t ld,
    TO* __restrict__ out) {
  __shared__ RPart sp[kRedFinishThreads];
  __shared__ int hv[kRedFinishThreads];
  const int64_t row = blockIdx.x;
  const int t = threadIdx.x;
  const int64_t per = (chunks_per_row + kRedFinishThreads - 1) / kRedFinishThreads;
  const int64_t c0 = t * per;
  const int64_t c1 = c0 + per < chunks_per_row ? c0 + per : chunks_per_row;
  const RPart* rp = parts + row * chunks_per_row;
  RPart p{};
  int have = 0;
  for (int64_t c = c0; c < c1; ++c) {
    if (have) merge_part<T, KIND>(p, rp[c]);
    else p = rp[c];
    have = 1;
  }
  sp[t] = p;
  hv[t] = have;
  for (int stride = 1; stride < kRedFinishThreads; stride <<= 1) {
    __syncthreads();
    if ((t & (2 * stride - 1)) == 0 && hv[t + stride]) {
      if (hv[t]) {
        RPart a = sp[t];
        merge_part<T, KIND>(a, sp[t + stride]);
        sp[t] = a;
      } else {
        sp[t] = sp[t + stride];
        hv[t] = 1;
      }
    }
  }
  if (t == 0) out[row] = part_value<T, TO, KIND>(sp[0], row, ddof, x, ld);
}

// Chunks per row: about 8 wavefronts per SIMD over the device, chunks of
// whole super-steps and at least 16 KiB of data.
template <typename T>
int64_t reduce_chunk(int64_t rows, int64_t n, int64_t super_elems) {
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) !=
                                              hipSuccess || cus <= 0)
    cus = 256;
  const int64_t target = (int64_t)cus * 4 * 8;
  int64_t per_row = target / (rows > 0 ? rows : 1);
  if (per_row < 1) per_row = 1;
  int64_t chunk = (n + per_row - 1) / per_row;
  const int64_t min_chunk = 16384 / (int64_t)sizeof(T);
  if (chunk < min_chunk) chunk = min_chunk;
  chunk = (chunk + super_elems - 1) / super_elems * super_elems;
  // below 2**31 positions (float32 extremes index within the chunk in 32 bits)
  const int64_t max_chunk = ((int64_t)1 << 31) / super_elems * super_elems;
  return chunk < max_chunk ? chunk : max_chunk;
}

template <typename T>
constexpr int red_unroll() {
  return 4;
}

size_t reduce_ws_bytes(int64_t rows, int64_t n, int elem) {
  if (rows <= 0) return 0;
  const int64_t super = (int64_t)kWave * (16 / elem) * 4;
  // the same chunking as the launch, computed without the device query
  // (at most rows * ceil(n / min_chunk) items; bounded by rows * 8192)
  int64_t min_chunk = 16384 / elem;
  min_chunk = (min_chunk + super - 1) / super * super;
  int64_t per_row = n > 0 ? (n + min_chunk - 1) / min_chunk : 1;
  if (per_row < 1) per_row = 1;
  return (size_t)(rows * per_row) * sizeof(RPart);
}

template <typename T, typename TO, int KIND>
int reduce_impl(const T* x, TO* out, int64_t rows, int64_t n, int64_t ld, int64_t ddof, void* ws,
                size_t ws_bytes, hipStream_t s) {
  constexpr int VEC = 16 / (int)sizeof(T);
  constexpr int U = red_unroll<T>();
  const int64_t super = (int64_t)kWave * VEC * U;
  const int64_t chunk = n > 0 ? reduce_chunk<T>(rows, n, super) : super;
  const int64_t cpr = n > 0 ? (n + chunk - 1) / chunk : 1;
  const int64_t items = rows * cpr;
  if ((size_t)items * sizeof(RPart) > ws_bytes || ws == nullptr)
    return set_error(NBG_ERR_INVALID, "reduce: workspace too small (%zu < %zu)", ws_bytes,
                     (size_t)items * sizeof(RPart));
  const int64_t blocks = (items + kRedWaves - 1) / kRedWaves;
  if (blocks > 0x7fffffffLL) return set_error(NBG_ERR_INVALID, "reduce: too many rows");
  RPart* parts = static_cast<RPart*>(ws);
  const bool vec_ok = pick_vec<T>(x, ld, rows) == VEC;
  if (vec_ok)
    hipLaunchKernelGGL((reduce_kernel<T, KIND, VEC, U>), dim3((unsigned)blocks), dim3(kRedThreads), 0, s, x, n, ld,
                       chunk, cpr, items, parts);
  else
    hipLaunchKernelGGL((reduce_kernel<T, KIND, 1, 1>), dim3((unsigned)blocks), dim3(kRedThreads), 0, s, x, n, ld,
                       chunk, cpr, items, parts);
  if (int rc = check_hip(hipGetLastError(), "reduce_kernel launch")) return rc;
  if (cpr >= 64 && rows <= 0x7fffffffLL) {
    hipLaunchKernelGGL((reduce_finalize_wide_kernel<T, TO, KIND>), dim3((unsigned)rows), dim3(kRedFinishThreads), 0,
                       s, parts, cpr, ddof, x, ld, out);
  } else {
    const int64_t fblocks = (rows + 255) / 256;
    hipLaunchKernelGGL((reduce_finalize_kernel<T, TO, KIND>), dim3((unsigned)fblocks), dim3(256), 0, s, parts, rows,
                       cpr, ddof, x, ld, out);
  }
  return check_hip(hipGetLastError(), "reduce_finalize_kernel launch");
}

// Output dtype per kind: sum -> input; mean/var/std -> input (floats only);
// count, argmax, argmin -> int64; allnan/anynan -> uint8 (bool); max/min ->
// input for floats, int64 for integers.
template <int KIND>
int reduce_entry(const char* name, int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld,
                 int64_t ddof, void* ws, size_t ws_bytes, void* stream) {
  if (int rc = injected_error(name)) return rc;
  if (rows < 0 || n < 0) return set_error(NBG_ERR_INVALID, "%s: negative shape", name);
  if (rows == 0) return NBG_OK;
  if ((x == nullptr && n > 0) || out == nullptr) return set_error(NBG_ERR_INVALID, "%s: null pointer", name);
  if (rows > 1 && ld < n) return set_error(NBG_ERR_INVALID, "%s: row stride < n", name);
  if (n == 0 && (KIND == kRArgMax || KIND == kRArgMin || KIND == kRMax || KIND == kRMin))
    return set_error(NBG_ERR_INVALID, "%s: zero-size reduction", name);
  constexpr bool FLOAT_ONLY = KIND == kRMean || is_var_kind(KIND);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool flt = dtype == NBG_F32 || dtype == NBG_F64;
  if (FLOAT_ONLY && !flt) return set_error(NBG_ERR_INVALID, "%s: float32 or float64 input required", name);
#define NBG_R(T, TO) return reduce_impl<T, TO, KIND>((const T*)x, (TO*)out, rows, n, ld, ddof, ws, ws_bytes, s)
  if constexpr (KIND == kRCount || KIND == kRArgMax || KIND == kRArgMin) {
    switch (dtype) {
      case NBG_F32: NBG_R(float, int64_t);
      case NBG_F64: NBG_R(double, int64_t);
      case NBG_I32: NBG_R(int32_t, int64_t);
      case NBG_I64: NBG_R(int64_t, int64_t);
    }
  } else if constexpr (KIND == kRAllNan || KIND == kRAnyNan) {
    switch (dtype) {
      case NBG_F32: NBG_R(float, uint8_t);
      case NBG_F64: NBG_R(double, uint8_t);
      case NBG_I32: NBG_R(int32_t, uint8_t);
      case NBG_I64: NBG_R(int64_t, uint8_t);
    }
  } else if constexpr (FLOAT_ONLY) {
    switch (dtype) {
      case NBG_F32: NBG_R(float, float);
      case NBG_F64: NBG_R(double, double);
    }
  } else if constexpr (KIND == kRSum) {
    switch (dtype) {
      case NBG_F32: NBG_R(float, float);
      case NBG_F64: NBG_R(double, double);
      case NBG_I32: NBG_R(int32_t, int32_t);
      case NBG_I64: NBG_R(int64_t, int64_t);
    }
  } else {  // max / min
    switch (dtype) {
      case NBG_F32: NBG_R(float, float);
      case NBG_F64: NBG_R(double, double);
      case NBG_I32: NBG_R(int32_t, int64_t);
      case NBG_I64: NBG_R(int64_t, int64_t);
    }
  }
#undef NBG_R
  return set_error(NBG_ERR_INVALID, "%s: unsupported dtype code %d", name, dtype);
}

}  // namespace
}  // namespace nbg

extern "C" {

size_t nbg_reduce_workspace_bytes(int64_t rows, int64_t n, int dtype) {
  const int elem = (dtype == NBG_F64 || dtype == NBG_I64) ? 8 : 4;
  return nbg::reduce_ws_bytes(rows, n, elem);
}

#define NBG_REDUCE(NAME, KIND)                                                                          \
  int NAME(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld, void* ws,          \
           size_t ws_bytes, void* stream) {                                                             \
    return nbg::reduce_entry<KIND>(#NAME, dtype, x, out, rows, n, ld, 0, ws, ws_bytes, stream);       \
  }
#define NBG_REDUCE_DDOF(NAME, KIND)                                                                     \
  int NAME(int dtype, const void* x, void* out, int64_t rows, int64_t n, int64_t ld, int64_t ddof,      \
           void* ws, size_t ws_bytes, void* stream) {                                                   \
    return nbg::reduce_entry<KIND>(#NAME, dtype, x, out, rows, n, ld, ddof, ws, ws_bytes, stream);    \
  }

NBG_REDUCE(nbg_nansum, nbg::kRSum)
NBG_REDUCE(nbg_nanmean, nbg::kRMean)
NBG_REDUCE(nbg_nancount, nbg::kRCount)
NBG_REDUCE(nbg_allnan, nbg::kRAllNan)
NBG_REDUCE(nbg_anynan, nbg::kRAnyNan)
NBG_REDUCE_DDOF(nbg_nanvar, nbg::kRVar)
NBG_REDUCE_DDOF(nbg_nanstd, nbg::kRStd)
NBG_REDUCE(nbg_nanmax, nbg::kRMax)
NBG_REDUCE(nbg_nanmin, nbg::kRMin)
NBG_REDUCE(nbg_nanargmax, nbg::kRArgMax)
NBG_REDUCE(nbg_nanargmin, nbg::kRArgMin)

}  // extern "C"
