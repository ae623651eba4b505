// Device label factorisation for gfx950: codes and uniques of a label array,
// the caller-side step before the grouped reductions (the reference requires
// labels already factorised, numbagg/decorators.py:570-573, and its benchmarks
// factorise with pd.factorize(labels, sort=True), numbagg/test/conftest.py:175-189).
//
// Semantics (pandas.factorize): codes[i] = position of values[i] among the
// uniques, -1 for NaN; uniques sorted ascending (sort) or in order of first
// appearance (no sort); -0.0 and +0.0 are one value.
//
// Design (hash, then sort the uniques):
//   1. insert: every value's 64-bit key goes into an open-addressing table
//      (linear probing, compare-and-swap on an empty marker); each slot also
//      records the first position of its key (atomic minimum); each
//      workgroup gathers its chunk's distinct keys in LDS first;
//   2. compact: occupied slots append (sort key, slot) pairs to a list;
//   3. sort: a bitonic network sorts the list by sort key (the order-preserving
//      image of the value, or the first position), padded to a power of two;
//   4. rank: each sorted entry writes its rank into its slot and its value into
//      the uniques;
//   5. codes: every value finds its slot again and reads the rank (for up to
//      2048 uniques in a key -> rank table that each workgroup builds in LDS).
// The empty marker is the all-ones word: no float key can equal it (it is a
// NaN pattern, and NaN is never inserted); the integer -1 (sign-extended to
// the same bits) is tracked by a flag outside the table.
#include <type_traits>

#include "common.hpp"

namespace nbg {
namespace {

constexpr unsigned long long kEmpty = ~0ull;
constexpr int kFzThreads = 256;

__device__ __forceinline__ unsigned long long mix64(unsigned long long k) {
  // finaliser of a 64-bit multiplicative hash (xor-shift / multiply rounds)
  k ^= k >> 33;
  k *= 0xff51afd7ed558ccdull;
  k ^= k >> 33;
  k *= 0xc4ceb9fe1a85ec53ull;
  k ^= k >> 33;
  return k;
}

// The table key of a value (-0.0 folded onto +0.0) and whether it is NaN.
template <typename T>
__device__ __forceinline__ unsigned long long fz_key(T v, bool& nan) {
  if constexpr (std::is_floating_point<T>::value) {
    nan = v != v;
    const double d = (double)v + 0.0;
    return (unsigned long long)__double_as_longlong(d);
  } else {
    nan = false;
    return (unsigned long long)(long long)v;
  }
}

// Order-preserving image of a key as an unsigned word (floats by their
// sign-magnitude order, integers by offsetting the sign).
template <typename T>
__device__ __forceinline__ unsigned long long fz_order(unsigned long long k) {
  if constexpr (std::is_floating_point<T>::value) {
    return (k & 0x8000000000000000ull) ? ~k : (k | 0x8000000000000000ull);
  } else {
    return k ^ 0x8000000000000000ull;
  }
}

// One open-addressing table: keys (kEmpty: free), first position of each
// key, and the key's code.
struct FzTable {
  unsigned long long mask;  // size - 1 (a power of two)
  unsigned long long* keys;
  unsigned long long* first;
  long long* rank;
};

// Two tables: a small one (at most kFzSmallSlots) that is tried first, and
// the big one of 2n slots, used only when the small one overflows (more than
// half full, or a probe longer than kFzSmallProbe) — so few uniques do not
// pay for initialising and scanning 2n slots.
constexpr int64_t kFzSmallSlots = 65536;
constexpr int kFzSmallProbe = 64;

struct FzArgs {
  int64_t n;
  FzTable big, small;
  unsigned int* count;                 // [0] uniques in the table, [1] minus-one flag,
                                       // [2] small-table slots taken, [3] small table overflowed
  unsigned long long* minus_one_first; // first position of an int64 -1
  unsigned long long* skey;            // sort keys (padded list)
  unsigned long long* sval;            // slot of each list entry (kEmpty: the int64 -1)
  int64_t padded;                      // list length, a power of two
};

// The table in use after the insert passes.
__device__ __forceinline__ FzTable fz_table(const FzArgs& A) { return A.count[3] ? A.big : A.small; }

// Key k (not the int64 -1) seen at position i: into table t (BIG: the 2n-slot
// table; else the small one, which flags its overflow and gives up).
// `over` (small table only): the workgroup's LDS copy of the overflow flag,
// set here when this call finds the overflow, and read at every probe so
// that a workgroup stops probing a full table as soon as one of its threads
// has seen it overflow.
template <bool BIG>
__device__ __forceinline__ void fz_table_insert(const FzArgs& A, const FzTable& t, unsigned long long k,
                                                unsigned long long i, unsigned int* over) {
  // The big table holds at most half its slots, so a free slot is found. A
  // slot changes only once (empty -> key) and a first position only
  // decreases, so plain reads that see an older value are safe: the common
  // case (the key is present, with an earlier first position) then takes
  // no atomic, which keeps labels with many repeats free of contention.
  unsigned long long s = mix64(k) & t.mask;
  for (int p = 0;; ++p) {
    if (!BIG) {
      if (__atomic_load_n(over, __ATOMIC_RELAXED)) return;
      if (p >= kFzSmallProbe) {
        atomicOr(&A.count[3], 1u);
        __atomic_store_n(over, 1u, __ATOMIC_RELAXED);
        return;
      }
    }
    unsigned long long cur = t.keys[s];
    if (cur == kEmpty) {
      cur = atomicCAS(&t.keys[s], kEmpty, k);
      if (!BIG && cur == kEmpty && atomicAdd(&A.count[2], 1u) + 1 > (unsigned int)((t.mask + 1) / 2)) {
        atomicOr(&A.count[3], 1u);
        __atomic_store_n(over, 1u, __ATOMIC_RELAXED);
      }
    }
    if (cur == kEmpty || cur == k) {
      if (i < t.first[s]) atomicMin(&t.first[s], i);
      return;
    }
    s = (s + 1) & t.mask;
  }
}

// Insert pass. Each workgroup first gathers its contiguous chunk's keys in an
// LDS table (key, first position), then puts each distinct key into the
// global table once: few uniques then cost one L2 access per key and
// workgroup instead of two per value. A key that finds no LDS slot within
// kFzLdsProbe probes (the table holds kFzLdsSlots) goes to the global table
// directly, so any number of uniques is correct.
constexpr int kFzInsThreads = 1024;
constexpr int kFzLdsSlots = 4096;  // 64 KiB of keys and first positions
constexpr int kFzLdsProbe = 16;

// BIG: the second pass, into the big table, run only when the small table
// overflowed.
template <typename T, bool BIG>
__global__ __launch_bounds__(kFzInsThreads) void fz_insert_kernel(const T* __restrict__ x, FzArgs A, int64_t chunk) {
  __shared__ unsigned long long lk[kFzLdsSlots], lf[kFzLdsSlots];
  __shared__ unsigned int lused, over;
  if (BIG && !A.count[3]) return;
  const FzTable t = BIG ? A.big : A.small;
  for (int j = threadIdx.x; j < kFzLdsSlots; j += kFzInsThreads) {
    lk[j] = kEmpty;
    lf[j] = kEmpty;
  }
  if (threadIdx.x == 0) {
    lused = 0u;
    over = 0u;
  }
  __syncthreads();
  const int64_t lo = (int64_t)blockIdx.x * chunk;
  const int64_t hi = lo + chunk < A.n ? lo + chunk : A.n;
  // four values per thread in flight: their loads are issued before the
  // first one is looked up
  constexpr int U = 4;
  for (int64_t i0 = lo + threadIdx.x; i0 < hi; i0 += U * kFzInsThreads) {
    // once the small table has overflowed its contents are dropped: stop
    if (!BIG && __atomic_load_n(&over, __ATOMIC_RELAXED)) break;
    // a table three-quarters full (many distinct keys in this chunk) only
    // looks keys up near their home slot and claims no more slots
    const bool claim = __atomic_load_n(&lused, __ATOMIC_RELAXED) < (unsigned int)(kFzLdsSlots * 3 / 4);
    const int lim = claim ? kFzLdsProbe : 4;
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + (int64_t)u * kFzInsThreads;
      v[u] = i < hi ? x[i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int64_t i = i0 + (int64_t)u * kFzInsThreads;
      if (i >= hi) continue;
      bool nan;
      const unsigned long long k = fz_key(v[u], nan);
      if (nan) continue;
      if (k == kEmpty) {  // int64 -1: outside the table
        atomicOr(&A.count[1], 1u);
        atomicMin(A.minus_one_first, (unsigned long long)i);
        continue;
      }
      unsigned int h = (unsigned int)mix64(k) & (kFzLdsSlots - 1);
      bool placed = false;
      for (int p = 0; p < lim; ++p) {
        unsigned long long cur = lk[h];
        if (cur == kEmpty) {
          if (!claim) break;
          cur = atomicCAS(&lk[h], kEmpty, k);
          if (cur == kEmpty) atomicAdd(&lused, 1u);
        }
        if (cur == kEmpty || cur == k) {
          if ((unsigned long long)i < lf[h]) atomicMin(&lf[h], (unsigned long long)i);
          placed = true;
          break;
        }
        h = (h + 1) & (kFzLdsSlots - 1);
      }
      if (!placed) fz_table_insert<BIG>(A, t, k, (unsigned long long)i, &over);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0 && !BIG && __atomic_load_n(&A.count[3], __ATOMIC_RELAXED)) over = 1u;
  __syncthreads();
  if (over) return;
  for (int j = threadIdx.x; j < kFzLdsSlots; j += kFzInsThreads)
    if (lk[j] != kEmpty) fz_table_insert<BIG>(A, t, lk[j], lf[j], &over);
}

// Occupied slots append (sort key, slot) to the list; the bitonic sort fixes
// the order. Each workgroup takes a contiguous range of slots, counts its
// occupied ones, claims their list positions with one atomic, then writes
// them in slot order (one atomic per workgroup: the same-address atomics of
// per-wave claims serialise in L2, 1.5 ms for 4M uniques).
template <typename T>
__global__ __launch_bounds__(kFzThreads) void fz_compact_kernel(FzArgs A, int by_value) {
  __shared__ unsigned int wcount[kFzThreads / 64], red[kFzThreads / 64];
  __shared__ unsigned int base;
  const FzTable t = fz_table(A);
  const int64_t size = (int64_t)t.mask + 1;
  int64_t per = (size + gridDim.x - 1) / gridDim.x;
  per = (per + kFzThreads - 1) / kFzThreads * kFzThreads;
  const int64_t r0 = (int64_t)blockIdx.x * per;
  const int64_t r1 = r0 + per < size ? r0 + per : size;
  const int lane = __lane_id(), wave = threadIdx.x / 64;
  unsigned int c = 0u;
  for (int64_t s = r0 + threadIdx.x; s < r1; s += kFzThreads) c += t.keys[s] != kEmpty ? 1u : 0u;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off);
  if (lane == 0) red[wave] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned int total = 0u;
    for (int w = 0; w < kFzThreads / 64; ++w) total += red[w];
    base = total ? atomicAdd(&A.count[0], total) : 0u;
  }
  __syncthreads();
  unsigned int off = base;
  // the workgroup steps together, so every wave has all its lanes here
  for (int64_t s0 = r0; s0 < r1; s0 += kFzThreads) {
    const int64_t s = s0 + threadIdx.x;
    const unsigned long long k = s < r1 ? t.keys[s] : kEmpty;
    const bool occ = k != kEmpty;
    const unsigned long long m = __ballot(occ);
    if (lane == 0) wcount[wave] = (unsigned int)__popcll(m);
    __syncthreads();
    unsigned int before = 0u, all = 0u;
    for (int w = 0; w < kFzThreads / 64; ++w) {
      before += w < wave ? wcount[w] : 0u;
      all += wcount[w];
    }
    if (occ) {
      const unsigned int p = off + before + (unsigned int)__popcll(m & ((1ull << lane) - 1ull));
      A.skey[p] = by_value ? fz_order<T>(k) : t.first[s];
      A.sval[p] = (unsigned long long)s;
    }
    off += all;
    __syncthreads();  // wcount is rewritten by the next step
  }
}

// The number of uniques (table entries plus an int64 -1).
__device__ __forceinline__ int64_t fz_total(const FzArgs& A) {
  return (int64_t)A.count[0] + (A.count[1] ? 1 : 0);
}

__global__ void fz_total_kernel(FzArgs A, long long* total) { *total = fz_total(A); }

// Entry K - 1 for an int64 -1 (value kEmpty, outside the table), and the
// padding [K, padded): the largest key with a value above every slot, so
// that (key, value) pairs order padding after every real entry.
template <typename T>
__global__ void fz_pad_kernel(FzArgs A, int64_t k_total, int by_value) {
  if (fz_total(A) != k_total) return;  // not the count of these values: write nothing
  const int64_t table_entries = (int64_t)A.count[0];
  for (int64_t p = table_entries + (int64_t)blockIdx.x * kFzThreads + threadIdx.x; p < A.padded;
       p += (int64_t)gridDim.x * kFzThreads) {
    if (p < k_total) {  // the int64 -1
      A.skey[p] = by_value ? fz_order<T>(kEmpty) : *A.minus_one_first;
      A.sval[p] = kEmpty;
    } else {
      A.skey[p] = kEmpty;
      A.sval[p] = kEmpty - 1;
    }
  }
}

// Bitonic stages in LDS. Each workgroup holds one chunk of kFzLdsSort
// entries (the whole list when it is shorter) and runs, for every stage size
// k from k_first to k_last, the steps j = min(k / 2, chunk / 2) .. 1, whose
// partners lie inside the chunk; the direction comes from the global index,
// so the chunks together form the same network as the one-step global
// kernel. (2, padded) on one workgroup sorts lists of up to kFzLdsSort
// entries in one launch; longer lists run (2, kFzLdsSort) once and then, for
// each k above it, global steps down to j = kFzLdsSort and one (k, k) launch
// for the rest.
constexpr int kFzLdsSort = 4096;
constexpr int kFzSortThreads = 1024;

__global__ __launch_bounds__(kFzSortThreads) void fz_lds_sort_kernel(unsigned long long* key, unsigned long long* val,
                                                                      int64_t padded, int64_t k_first, int64_t k_last) {
  __shared__ unsigned long long sk[kFzLdsSort], sv[kFzLdsSort];
  const int len = padded < kFzLdsSort ? (int)padded : kFzLdsSort;
  const int64_t base = (int64_t)blockIdx.x * len;
  for (int i = threadIdx.x; i < len; i += kFzSortThreads) {
    sk[i] = key[base + i];
    sv[i] = val[base + i];
  }
  __syncthreads();
  constexpr int M = kFzLdsSort / kFzSortThreads;  // entries per thread
  for (int64_t k = k_first; k <= k_last; k <<= 1) {
    const int jtop = (int)((k >> 1) < (len >> 1) ? (k >> 1) : (len >> 1));
    // strides of a wave or more: through LDS, a barrier per step
    for (int j = jtop; j >= 64; j >>= 1) {
      for (int i = threadIdx.x; i < len; i += kFzSortThreads) {
        const int p = i ^ j;
        if (p <= i) continue;
        const bool up = ((base + i) & k) == 0;
        const unsigned long long a = sk[i], b = sk[p];
        const bool greater = a > b || (a == b && sv[i] > sv[p]);
        if (greater == up) {
          sk[i] = b;
          sk[p] = a;
          const unsigned long long t = sv[i];
          sv[i] = sv[p];
          sv[p] = t;
        }
      }
      __syncthreads();
    }
    // strides below 64: partners in the same wave, exchanged by lane shuffles
    // (entry e sits in lane e mod 64), no barriers
    const int jr = jtop < 32 ? jtop : 32;
    if (jr >= 1) {
      unsigned long long rk[M], rv[M];
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int e = threadIdx.x + m * kFzSortThreads;
        if (e < len) {
          rk[m] = sk[e];
          rv[m] = sv[e];
        }
      }
      for (int j = jr; j > 0; j >>= 1) {
#pragma unroll
        for (int m = 0; m < M; ++m) {
          const int e = threadIdx.x + m * kFzSortThreads;
          if (e < len) {  // len is a power of two: the partner e ^ j is active too
            const unsigned long long pk = __shfl_xor(rk[m], j), pv = __shfl_xor(rv[m], j);
            const bool lower = (e & j) == 0;
            const bool up = ((base + e) & k) == 0;
            // (a, va): the pair's lower entry, (b, vb): its upper entry
            const unsigned long long a = lower ? rk[m] : pk, b = lower ? pk : rk[m];
            const unsigned long long va = lower ? rv[m] : pv, vb = lower ? pv : rv[m];
            const bool greater = a > b || (a == b && va > vb);
            if (greater == up) {
              rk[m] = pk;
              rv[m] = pv;
            }
          }
        }
      }
#pragma unroll
      for (int m = 0; m < M; ++m) {
        const int e = threadIdx.x + m * kFzSortThreads;
        if (e < len) {
          sk[e] = rk[m];
          sv[e] = rv[m];
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < len; i += kFzSortThreads) {
    key[base + i] = sk[i];
    val[base + i] = sv[i];
  }
}

// One compare-exchange stage of the bitonic network (ascending overall).
__global__ __launch_bounds__(kFzThreads) void fz_bitonic_kernel(unsigned long long* key, unsigned long long* val,
                                                               int64_t padded, int64_t j, int64_t k) {
  for (int64_t i = (int64_t)blockIdx.x * kFzThreads + threadIdx.x; i < padded; i += (int64_t)gridDim.x * kFzThreads) {
    const int64_t p = i ^ j;
    if (p <= i) continue;
    const bool up = (i & k) == 0;
    const unsigned long long a = key[i], b = key[p];
    const bool greater = a > b || (a == b && val[i] > val[p]);  // (key, value) order
    if (greater == up) {
      key[i] = b;
      key[p] = a;
      const unsigned long long t = val[i];
      val[i] = val[p];
      val[p] = t;
    }
  }
}

// Ranks into the table, and into the uniques the value at each key's first
// position (so -0.0 or +0.0, whichever came first, as pandas keeps it).
template <typename T>
__global__ __launch_bounds__(kFzThreads) void fz_rank_kernel(const T* __restrict__ x, FzArgs A, int64_t k_total,
                                                            T* __restrict__ uniques,
                                                            long long* __restrict__ minus_one_rank) {
  if (fz_total(A) != k_total) return;
  const FzTable t = fz_table(A);
  for (int64_t r = (int64_t)blockIdx.x * kFzThreads + threadIdx.x; r < k_total; r += (int64_t)gridDim.x * kFzThreads) {
    const unsigned long long s = A.sval[r];
    unsigned long long pos;
    if (s == kEmpty) {
      *minus_one_rank = r;
      pos = *A.minus_one_first;
    } else {
      t.rank[s] = r;
      pos = t.first[s];
    }
    uniques[r] = x[pos];
  }
}

template <typename T>
__global__ __launch_bounds__(kFzThreads) void fz_codes_kernel(const T* __restrict__ x, FzArgs A,
                                                             const long long* __restrict__ minus_one_rank,
                                                             long long* __restrict__ codes, int64_t k_total) {
  if (fz_total(A) != k_total) return;
  const FzTable t = fz_table(A);
  for (int64_t i = (int64_t)blockIdx.x * kFzThreads + threadIdx.x; i < A.n; i += (int64_t)gridDim.x * kFzThreads) {
    bool nan;
    const unsigned long long k = fz_key(x[i], nan);
    long long c = -1;
    if (!nan) {
      if (k == kEmpty) {
        c = *minus_one_rank;
      } else {
        // The table is at most half full, so the probe ends at the key or at a
        // free slot (a key not inserted: values changed since the count).
        for (unsigned long long s = mix64(k) & t.mask;; s = (s + 1) & t.mask) {
          const unsigned long long ks = t.keys[s];
          if (ks == k) {
            c = t.rank[s];
            break;
          }
          if (ks == kEmpty) break;
        }
      }
    }
    codes[i] = c;
  }
}

// Codes for at most kFzLdsCodes uniques: each workgroup builds an LDS hash
// table (key -> rank, twice as many slots as uniques) from the sorted list,
// then looks every value up there instead of in the global table.
constexpr int kFzLdsCodes = 2048;
constexpr int kFzCodeSlots = 2 * kFzLdsCodes;
constexpr int kFzCodeThreads = 1024;

template <typename T>
__global__ __launch_bounds__(kFzCodeThreads) void fz_codes_lds_kernel(const T* __restrict__ x, FzArgs A,
                                                                     const long long* __restrict__ minus_one_rank,
                                                                     long long* __restrict__ codes, int64_t k_total) {
  __shared__ unsigned long long tk[kFzCodeSlots];
  __shared__ int tr[kFzCodeSlots];
  if (fz_total(A) != k_total || k_total > kFzLdsCodes) return;
  const FzTable t = fz_table(A);
  for (int j = threadIdx.x; j < kFzCodeSlots; j += kFzCodeThreads) tk[j] = kEmpty;
  __syncthreads();
  for (int r = threadIdx.x; r < (int)k_total; r += kFzCodeThreads) {
    const unsigned long long s = A.sval[r];
    if (s == kEmpty) continue;  // the int64 -1
    const unsigned long long k = t.keys[s];
    // distinct keys, at most half the slots: a free slot is always found
    for (unsigned int h = (unsigned int)mix64(k) & (kFzCodeSlots - 1);; h = (h + 1) & (kFzCodeSlots - 1)) {
      if (atomicCAS(&tk[h], kEmpty, k) == kEmpty) {
        tr[h] = r;
        break;
      }
    }
  }
  __syncthreads();
  const long long m1 = *minus_one_rank;
  for (int64_t i = (int64_t)blockIdx.x * kFzCodeThreads + threadIdx.x; i < A.n;
       i += (int64_t)gridDim.x * kFzCodeThreads) {
    bool nan;
    const unsigned long long k = fz_key(x[i], nan);
    long long c = -1;
    if (!nan) {
      if (k == kEmpty) {
        c = m1;
      } else {
        // ends at the key or at a free slot (a key not inserted: values
        // changed since the count)
        for (unsigned int h = (unsigned int)mix64(k) & (kFzCodeSlots - 1);; h = (h + 1) & (kFzCodeSlots - 1)) {
          const unsigned long long ks = tk[h];
          if (ks == k) {
            c = tr[h];
            break;
          }
          if (ks == kEmpty) break;
        }
      }
    }
    codes[i] = c;
  }
}

// Clears the small table and the counters (BIG false), or the big table when
// the small one overflowed (BIG true).
template <bool BIG>
__global__ void fz_init_kernel(FzArgs A) {
  if (BIG && !A.count[3]) return;
  const FzTable t = BIG ? A.big : A.small;
  const int64_t size = (int64_t)t.mask + 1;
  for (int64_t s = (int64_t)blockIdx.x * kFzThreads + threadIdx.x; s < size; s += (int64_t)gridDim.x * kFzThreads) {
    t.keys[s] = kEmpty;
    t.first[s] = kEmpty;
  }
  if (!BIG && blockIdx.x == 0 && threadIdx.x == 0) {
    for (int c = 0; c < 4; ++c) A.count[c] = 0u;
    *A.minus_one_first = kEmpty;
  }
}

int64_t next_pow2(int64_t v) {
  int64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Table of 2 n slots (rounded up to a power of two, at least 1024): load
// factor at most one half.
int64_t fz_table_size(int64_t n) { return next_pow2(2 * (n > 512 ? n : 512)); }

int64_t fz_small_size(int64_t n) {
  const int64_t big = fz_table_size(n);
  return big < kFzSmallSlots ? big : kFzSmallSlots;
}

size_t fz_ws_bytes(int64_t n) {
  const int64_t size = fz_table_size(n) + fz_small_size(n);
  const int64_t padded = next_pow2(n + 1);
  return (size_t)size * 24 + (size_t)padded * 16 + 32;
}

// Workspace: keys | first | rank of the big table, then of the small one
// (table size each, 8 bytes) | sort keys | sort values (n + 1 rounded up to
// a power of two each) | four 32-bit counters | minus-one first position |
// minus-one rank.
FzTable fz_table_at(char*& p, int64_t size) {
  FzTable t;
  t.mask = (unsigned long long)(size - 1);
  t.keys = reinterpret_cast<unsigned long long*>(p);
  t.first = t.keys + size;
  t.rank = reinterpret_cast<long long*>(t.first + size);
  p = reinterpret_cast<char*>(t.rank + size);
  return t;
}

FzArgs fz_layout(int64_t n, void* ws) {
  FzArgs a{};
  const int64_t padded_max = next_pow2(n + 1);
  char* p = static_cast<char*>(ws);
  a.n = n;
  a.big = fz_table_at(p, fz_table_size(n));
  a.small = fz_table_at(p, fz_small_size(n));
  a.skey = reinterpret_cast<unsigned long long*>(p);
  a.sval = a.skey + padded_max;
  a.count = reinterpret_cast<unsigned int*>(a.sval + padded_max);
  a.minus_one_first = reinterpret_cast<unsigned long long*>(a.count + 4);
  a.padded = padded_max;
  return a;
}
long long* fz_minus_one_rank(int64_t n, void* ws) {
  return reinterpret_cast<long long*>(fz_layout(n, ws).minus_one_first + 1);
}

int fz_grid(int64_t work) {
  int64_t b = (work + kFzThreads - 1) / kFzThreads;
  if (b > 4096) b = 4096;
  return (int)(b < 1 ? 1 : b);
}

}  // namespace
}  // namespace nbg

extern "C" {

size_t nbg_factorize_workspace_bytes(int64_t n) { return n < 0 ? 0 : nbg::fz_ws_bytes(n); }

int nbg_factorize_count(int dtype, const void* values, int64_t n, int sort, void* ws, size_t ws_bytes,
                        int64_t* num_uniques, void* stream) {
  if (int rc = nbg::injected_error("nbg_factorize_count")) return rc;
  if (n < 0) return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_count: negative length");
  if ((values == nullptr && n > 0) || num_uniques == nullptr || ws == nullptr)
    return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_count: null pointer");
  if (ws_bytes < nbg::fz_ws_bytes(n))
    return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_count: workspace too small");
  nbg::FzArgs a = nbg::fz_layout(n, ws);
  hipStream_t s = (hipStream_t)stream;
  const int64_t size = (int64_t)a.big.mask + 1;
  const int64_t small = (int64_t)a.small.mask + 1;
  hipLaunchKernelGGL(nbg::fz_init_kernel<false>, dim3(nbg::fz_grid(small)), dim3(nbg::kFzThreads), 0, s, a);
  // Insert: at most 512 workgroups (two per CU in 64 KiB of LDS each), each
  // over a contiguous chunk, a multiple of the workgroup size.
  int64_t ins_blocks = (n + 16 * nbg::kFzInsThreads - 1) / (16 * nbg::kFzInsThreads);
  ins_blocks = ins_blocks < 1 ? 1 : (ins_blocks > 512 ? 512 : ins_blocks);
  int64_t chunk = (n + ins_blocks - 1) / ins_blocks;
  chunk = (chunk + nbg::kFzInsThreads - 1) / nbg::kFzInsThreads * nbg::kFzInsThreads;
  if (chunk < nbg::kFzInsThreads) chunk = nbg::kFzInsThreads;
  ins_blocks = (n + chunk - 1) / chunk;
  if (ins_blocks < 1) ins_blocks = 1;
  // small table, then (only if it overflowed: the kernels exit at once
  // otherwise) the big one
#define NBG_FZ1(T)                                                                                          \
  hipLaunchKernelGGL((nbg::fz_insert_kernel<T, false>), dim3((unsigned)ins_blocks), dim3(nbg::kFzInsThreads), \
                     0, s, (const T*)values, a, chunk);                                                     \
  hipLaunchKernelGGL(nbg::fz_init_kernel<true>, dim3(nbg::fz_grid(size)), dim3(nbg::kFzThreads), 0, s, a);    \
  hipLaunchKernelGGL((nbg::fz_insert_kernel<T, true>), dim3((unsigned)ins_blocks), dim3(nbg::kFzInsThreads),  \
                     0, s, (const T*)values, a, chunk);                                                     \
  hipLaunchKernelGGL((nbg::fz_compact_kernel<T>), dim3(nbg::fz_grid(size)), dim3(nbg::kFzThreads), 0, s, a,   \
                     sort ? 1 : 0)
  switch (dtype) {
    case NBG_F32: NBG_FZ1(float); break;
    case NBG_F64: NBG_FZ1(double); break;
    case NBG_I32: NBG_FZ1(int32_t); break;
    case NBG_I64: NBG_FZ1(int64_t); break;
    default: return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_count: unsupported dtype code %d", dtype);
  }
#undef NBG_FZ1
  hipLaunchKernelGGL(nbg::fz_total_kernel, dim3(1), dim3(1), 0, s, a, (long long*)num_uniques);
  return nbg::check_hip(hipGetLastError(), "factorize launch");
}

int nbg_factorize_finish(int dtype, const void* values, int64_t n, int64_t num_uniques, int sort, void* codes,
                         void* uniques, void* ws, size_t ws_bytes, void* stream) {
  if (int rc = nbg::injected_error("nbg_factorize_finish")) return rc;
  if (n < 0 || num_uniques < 0 || num_uniques > n)
    return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_finish: bad sizes");
  if ((codes == nullptr && n > 0) || (uniques == nullptr && num_uniques > 0) || ws == nullptr)
    return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_finish: null pointer");
  if (ws_bytes < nbg::fz_ws_bytes(n))
    return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_finish: workspace too small");
  if (n == 0) return NBG_OK;
  nbg::FzArgs a = nbg::fz_layout(n, ws);
  a.padded = nbg::next_pow2(num_uniques > 1 ? num_uniques : 1);
  hipStream_t s = (hipStream_t)stream;
  long long* minus_one_rank = nbg::fz_minus_one_rank(n, ws);
  // LDS codes: about 16 values per thread, at most 512 workgroups (each
  // builds its own table)
  int64_t codes_blocks = (n + 16 * nbg::kFzCodeThreads - 1) / (16 * nbg::kFzCodeThreads);
  codes_blocks = codes_blocks < 1 ? 1 : (codes_blocks > 512 ? 512 : codes_blocks);
#define NBG_FZ2(T)                                                                                            \
  do {                                                                                                        \
    hipLaunchKernelGGL((nbg::fz_pad_kernel<T>), dim3(nbg::fz_grid(a.padded)), dim3(nbg::kFzThreads), 0, s, a,   \
                       num_uniques, sort ? 1 : 0);                                                            \
    if (a.padded <= nbg::kFzLdsSort) {                                                                        \
      hipLaunchKernelGGL(nbg::fz_lds_sort_kernel, dim3(1), dim3(nbg::kFzSortThreads), 0, s, a.skey, a.sval,     \
                         a.padded, (int64_t)2, a.padded);                                                     \
    } else {                                                                                                  \
      const unsigned chunks = (unsigned)(a.padded / nbg::kFzLdsSort);                                         \
      hipLaunchKernelGGL(nbg::fz_lds_sort_kernel, dim3(chunks), dim3(nbg::kFzSortThreads), 0, s, a.skey,        \
                         a.sval, a.padded, (int64_t)2, (int64_t)nbg::kFzLdsSort);                             \
      for (int64_t k = 2 * (int64_t)nbg::kFzLdsSort; k <= a.padded; k <<= 1) {                                \
        for (int64_t j = k >> 1; j >= nbg::kFzLdsSort; j >>= 1)                                               \
          hipLaunchKernelGGL(nbg::fz_bitonic_kernel, dim3(nbg::fz_grid(a.padded)), dim3(nbg::kFzThreads), 0,   \
                             s, a.skey, a.sval, a.padded, j, k);                                              \
        hipLaunchKernelGGL(nbg::fz_lds_sort_kernel, dim3(chunks), dim3(nbg::kFzSortThreads), 0, s, a.skey,      \
                           a.sval, a.padded, k, k);                                                           \
      }                                                                                                       \
    }                                                                                                         \
    hipLaunchKernelGGL((nbg::fz_rank_kernel<T>), dim3(nbg::fz_grid(num_uniques)), dim3(nbg::kFzThreads), 0, s,  \
                       (const T*)values, a, num_uniques, (T*)uniques, minus_one_rank);                        \
    if (num_uniques <= nbg::kFzLdsCodes)                                                                      \
      hipLaunchKernelGGL((nbg::fz_codes_lds_kernel<T>), dim3(codes_blocks), dim3(nbg::kFzCodeThreads), 0, s,    \
                         (const T*)values, a, (const long long*)minus_one_rank, (long long*)codes, num_uniques);  \
    else                                                                                                      \
      hipLaunchKernelGGL((nbg::fz_codes_kernel<T>), dim3(nbg::fz_grid(n)), dim3(nbg::kFzThreads), 0, s,         \
                         (const T*)values, a, (const long long*)minus_one_rank, (long long*)codes, num_uniques);  \
  } while (0)
  switch (dtype) {
    case NBG_F32: NBG_FZ2(float); break;
    case NBG_F64: NBG_FZ2(double); break;
    case NBG_I32: NBG_FZ2(int32_t); break;
    case NBG_I64: NBG_FZ2(int64_t); break;
    default: return nbg::set_error(NBG_ERR_INVALID, "nbg_factorize_finish: unsupported dtype code %d", dtype);
  }
#undef NBG_FZ2
  return nbg::check_hip(hipGetLastError(), "factorize launch");
}

}  // extern "C"
