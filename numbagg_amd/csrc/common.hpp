// Shared device/host helpers for the numbagg_amd HIP kernels (gfx950 / CDNA4).
//
// Everything here is written for 64-wide wavefronts. Most kernels are
// bandwidth-bound streaming scans and reductions; the static correlation /
// covariance matrices (matrix.hip) are the one dense contraction and run on the
// float64 matrix cores.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

#include "../../include/numbagg_amd.h"

namespace nbg {

constexpr int kWave = 64;

// ---------------------------------------------------------------------------
// Error reporting (host side). Each C entry point returns an NBG_* status code
// and, on failure, records a message retrievable with nbg_last_error().
// ---------------------------------------------------------------------------
int set_error(int code, const char* fmt, ...);
int check_hip(hipError_t e, const char* what);
int injected_error(const char* entry);  // NBG_INJECT_ERROR=<entry> forces a failure
int device_capacity_blocks(const void* kernel, int block, size_t lds_bytes);

#define NBG_HIP_TRY(expr)                                         \
  do {                                                            \
    hipError_t _e = (expr);                                       \
    if (_e != hipSuccess) return ::nbg::check_hip(_e, #expr);     \
  } while (0)

// Tuning knobs (tile shapes, grid caps, path choices). The value of knob
// `name` is this host thread's override (nbg_tune_set: the autotuner's choice
// for the calls the thread is about to make), else the environment variable of
// the same name, else `dflt`. Results never depend on a knob beyond the
// float rounding the parity bars allow.
long long knob(const char* name, long long dflt);

// Device address of this device's counter of bounded spin-waits that expired
// (should always stay 0). Kernels that wait on other workgroups take it as an
// argument; the host reads it through nbg_spin_timeouts().
unsigned int* spin_counter_ptr();

// ---------------------------------------------------------------------------
// Numerics helpers.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ bool is_nan(T v) { return v != v; }

template <typename T> __device__ __forceinline__ T canonical_nan();
template <> __device__ __forceinline__ float canonical_nan<float>() { return __builtin_nanf(""); }
template <> __device__ __forceinline__ double canonical_nan<double>() { return __builtin_nan(""); }

// Error-free transformations (Knuth TwoSum / Dekker Fast2Sum). The library is
// compiled with -ffp-contract=off so the compiler cannot fuse these.
template <typename F>
__device__ __forceinline__ void two_sum(F a, F b, F& s, F& e) {
  s = a + b;
  F bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}
template <typename F>
__device__ __forceinline__ void fast_two_sum(F a, F b, F& s, F& e) {
  s = a + b;
  e = b - (s - a);
}

// Compensated ("double-F") accumulator: value = hi + lo with |lo| <= ulp(hi)/2.
// For F=float this is the FP32 TwoSum accumulator used for first moments of
// float32 data; for F=double it is double-double, used for float64 data.
template <typename F>
struct Comp {
  F hi, lo;
  __device__ __forceinline__ static Comp zero() { return Comp{F(0), F(0)}; }
  __device__ __forceinline__ void add(F x) {
    F s, e;
    two_sum(hi, x, s, e);
    e += lo;
    fast_two_sum(s, e, hi, lo);
  }
  __device__ __forceinline__ void add(const Comp& o) {
    F s, e, t, f;
    two_sum(hi, o.hi, s, e);
    two_sum(lo, o.lo, t, f);
    e += t;
    fast_two_sum(s, e, s, e);
    e += f;
    fast_two_sum(s, e, hi, lo);
  }
  __device__ __forceinline__ void sub(const Comp& o) {
    Comp n{-o.hi, -o.lo};
    add(n);
  }
  __device__ __forceinline__ double to_double() const { return (double)hi + (double)lo; }
};

// ---------------------------------------------------------------------------
// Wave-level primitives.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int lane_id() { return threadIdx.x & (kWave - 1); }
__device__ __forceinline__ int wave_id() { return threadIdx.x / kWave; }

template <typename T>
__device__ __forceinline__ T shfl_up(T v, int d) { return __shfl_up(v, d, kWave); }

// DPP lane moves: cross-lane reads inside the VALU, without the LDS permute
// path that __shfl takes. CTRL: row_shr:n = 0x110 + n (n = 1..15),
// wave_shr:1 (whole wave, by one lane) = 0x138,
// row_bcast:15 = 0x142, row_bcast:31 = 0x143; ROW_MASK selects the 16-lane
// rows written. Lanes that are not written (source outside the row, or row
// masked off) return `old`. The whole wave must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ int dpp_move(int old, int v) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, ROW_MASK, 0xf, false);
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_move(float old, float v) {
  return __int_as_float(dpp_move<CTRL, ROW_MASK>(__float_as_int(old), __float_as_int(v)));
}
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double old, double v) {
  const long long o = __double_as_longlong(old), x = __double_as_longlong(v);
  const int lo = dpp_move<CTRL, ROW_MASK>((int)o, (int)x);
  const int hi = dpp_move<CTRL, ROW_MASK>((int)(o >> 32), (int)(x >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned int)hi << 32) | (unsigned int)lo));
}

// Value of v in one lane (a wave-uniform index), read into a scalar register.
template <typename T>
__device__ __forceinline__ T read_lane(T v, int lane);
template <>
__device__ __forceinline__ int read_lane<int>(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}
template <>
__device__ __forceinline__ float read_lane<float>(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
template <>
__device__ __forceinline__ long long read_lane<long long>(long long v, int lane) {
  const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)v, lane);
  const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(v >> 32), lane);
  return (long long)(((unsigned long long)hi << 32) | lo);
}
template <>
__device__ __forceinline__ double read_lane<double>(double v, int lane) {
  const long long x = __double_as_longlong(v);
  const unsigned int lo = (unsigned int)__builtin_amdgcn_readlane((int)x, lane);
  const unsigned int hi = (unsigned int)__builtin_amdgcn_readlane((int)(x >> 32), lane);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

template <typename T>
__device__ __forceinline__ T shfl_down(T v, int d) { return __shfl_down(v, d, kWave); }
template <typename T>
__device__ __forceinline__ T shfl(T v, int src) { return __shfl(v, src, kWave); }

// ---------------------------------------------------------------------------
// Agent-scope status words for cross-workgroup (decoupled look-back) protocols.
// Relaxed agent-scope atomics bypass the non-coherent per-XCD caches, so a word
// published by a workgroup on one XCD is visible to a workgroup on another.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long ld_agent(const unsigned long long* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Multi-word publication (flag + payload) where every payload word is itself
// written and read with the relaxed agent-scope atomics below. Agent-scope
// release/acquire would write back / invalidate the whole per-XCD L2 on every
// publish and every poll; since no payload word lives in a non-coherent cache,
// ordering is all that is needed: the writer waits until its payload stores
// have completed (s_waitcnt with all counters at zero) before storing the
// flag, and the reader issues its payload loads only after the flag load has
// returned (the spin loop branches on its value). The signal fences keep the
// compiler from moving the atomics across these points.
__device__ __forceinline__ unsigned int ld_flag(const unsigned int* p) {
  const unsigned int v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return v;
}
__device__ __forceinline__ void st_flag(unsigned int* p, unsigned int v) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double ld_agent_f64(const double* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_agent_f64(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Bounded spin: every wait in this library ends. ~2^24 sleeps of 64 cycles is
// well over a second; a wait that long means the protocol was broken, and the
// kernel records it in g_spin_timeouts instead of hanging the device.
constexpr unsigned int kSpinLimit = 1u << 24;

__device__ __forceinline__ void spin_pause() { __builtin_amdgcn_s_sleep(1); }

// Workgroups of one launch are handed to the 8 XCDs round-robin (workgroup b
// runs on XCD b % 8), and each XCD has its own L2. xcd_order(b, n) renumbers
// the grid so that consecutive logical ids run on the same XCD one after the
// other: logical id v = (b % 8) * (n / 8) + b / 8 for the first 8 * (n / 8)
// workgroups (a bijection onto [0, 8 * (n / 8))), the rest keep their id.
// Kernels whose neighbouring work items share inputs (a rolling tile and the
// halo of the next one, the row groups of one label chunk) then find those
// inputs in their XCD's L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_order(uint32_t b, uint32_t n) {
  const uint32_t per = n / kXcds;
  return b < per * kXcds ? (b % kXcds) * per + b / kXcds : b;
}

// LDS index with one padding slot per 16 elements: threads that each own 16
// consecutive elements then hit distinct banks on every strided access.
__host__ __device__ __forceinline__ int64_t pad16(int64_t e) { return e + (e >> 4); }
__host__ __device__ __forceinline__ int pad16(int e) { return e + (e >> 4); }

// Integer vectors of SIZE bytes at alignment ALIGN (the alignment of the
// values they carry: element-aligned 16-byte accesses are legal on gfx9+).
template <int SIZE, int ALIGN>
struct NtVec;
template <>
struct NtVec<4, 4> {
  typedef int t;
};
template <>
struct NtVec<8, 4> {
  typedef int t __attribute__((ext_vector_type(2), aligned(4)));
};
template <>
struct NtVec<8, 8> {
  typedef int t __attribute__((ext_vector_type(2), aligned(8)));
};
template <>
struct NtVec<16, 4> {
  typedef int t __attribute__((ext_vector_type(4), aligned(4)));
};
template <>
struct NtVec<16, 8> {
  typedef int t __attribute__((ext_vector_type(4), aligned(8)));
};
template <>
struct NtVec<16, 16> {
  typedef int t __attribute__((ext_vector_type(4), aligned(16)));
};

// A load of a 4-, 8- or 16-byte value (vector of values) that is read once:
// non-temporal, so the stream does not displace reused data from the caches
// (measured on MI355X: a 16-byte-per-thread copy of 4 GB runs at 6.5 TB/s
// with non-temporal loads and stores, 6.26 TB/s with plain ones).
template <typename S>
__device__ __forceinline__ S load_stream(const S* p) {
  using R = typename NtVec<(int)sizeof(S), (int)alignof(S)>::t;
  const R x = __builtin_nontemporal_load(reinterpret_cast<const R*>(p));
  S r;
  __builtin_memcpy(&r, &x, sizeof(S));
  return r;
}

// A 16-byte store of 16 / sizeof(T) values (at the alignment of T),
// non-temporal: for outputs written once in whole-wave contiguous runs, which
// then go to HBM without lingering as dirty lines in the caches (their later
// write-back would take bandwidth from whatever kernel runs next).
template <typename T>
__device__ __forceinline__ void store_stream16(T* p, const T (&v)[16 / sizeof(T)]) {
  using R = typename NtVec<16, (int)alignof(T)>::t;
  R x;
  __builtin_memcpy(&x, v, 16);
  __builtin_nontemporal_store(x, reinterpret_cast<R*>(p));
}

// Zero-fill kernel used to reset look-back status words and ticket counters.
// (A kernel rather than a memset node so the call stays graph-capture safe.)
__global__ void zero_words_kernel(unsigned long long* p, int64_t nwords);
int launch_zero(void* p, size_t bytes, hipStream_t s);

// Widest vector (in elements, up to 16 bytes) that every row of a (rows, ld)
// view can be loaded with: the base must be aligned, and for more than one row
// so must the row stride.
template <typename T>
inline int pick_vec(const void* p, int64_t ld, int64_t rows) {
  constexpr int maxv = 16 / (int)sizeof(T);
  for (int v = maxv; v > 1; v >>= 1) {
    if (((uintptr_t)p % (v * sizeof(T))) == 0 && (rows <= 1 || (ld % v) == 0)) return v;
  }
  return 1;
}

// Widest vector for the rolling kernels, whose 16-byte global accesses need
// element alignment only (see Vec16): MAXV elements when every pointer is
// element-aligned, else 1.
template <typename T>
inline int pick_vec_elem(const void* a, const void* b, const void* out) {
  const bool ok = ((uintptr_t)a % sizeof(T)) == 0 && ((uintptr_t)b % sizeof(T)) == 0 &&
                  ((uintptr_t)out % sizeof(T)) == 0;
  return ok ? 16 / (int)sizeof(T) : 1;
}

}  // namespace nbg
