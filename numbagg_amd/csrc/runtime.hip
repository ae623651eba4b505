// Library-wide runtime pieces of the C ABI: error strings, version, the
// spin-timeout counter, fault injection for host-side error-path tests, and the
// graph-safe zero-fill kernel used to reset cross-workgroup status words.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include "common.hpp"

namespace nbg {

__device__ unsigned int g_spin_timeouts = 0;

unsigned int* spin_counter_ptr() {
  static std::mutex mu;
  static std::unordered_map<int, unsigned int*> ptrs;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return nullptr;
  std::lock_guard<std::mutex> g(mu);
  auto it = ptrs.find(dev);
  if (it != ptrs.end()) return it->second;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(g_spin_timeouts)) != hipSuccess) return nullptr;
  ptrs[dev] = (unsigned int*)p;
  return (unsigned int*)p;
}

static thread_local char t_err[512] = "";

// This thread's knob overrides (a handful of entries at most).
struct KnobEntry {
  char name[40];
  long long value;
};
constexpr int kMaxKnobs = 32;
static thread_local KnobEntry t_knobs[kMaxKnobs];
static thread_local int t_nknobs = 0;

long long knob(const char* name, long long dflt) {
  for (int i = 0; i < t_nknobs; ++i)
    if (strcmp(t_knobs[i].name, name) == 0) return t_knobs[i].value;
  const char* v = getenv(name);
  if (v != nullptr && v[0] != '\0') return atoll(v);
  return dflt;
}

int set_error(int code, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(t_err, sizeof(t_err), fmt, ap);
  va_end(ap);
  return code;
}

int check_hip(hipError_t e, const char* what) {
  if (e == hipSuccess) return NBG_OK;
  return set_error(NBG_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
}

int injected_error(const char* entry) {
  const char* v = getenv("NBG_INJECT_ERROR");
  if (v == nullptr || v[0] == '\0') return NBG_OK;
  if (strcmp(v, "*") == 0 || strcmp(v, entry) == 0)
    return set_error(NBG_ERR_INJECTED, "%s: injected failure (NBG_INJECT_ERROR=%s)", entry, v);
  return NBG_OK;
}

int device_capacity_blocks(const void* kernel, int block, size_t lds_bytes) {
  static std::mutex mu;
  static std::unordered_map<unsigned long long, int> cache;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  unsigned long long key = ((unsigned long long)(uintptr_t)kernel) ^
                           ((unsigned long long)dev << 56) ^ ((unsigned long long)block << 40) ^
                           (unsigned long long)lds_bytes;
  {
    std::lock_guard<std::mutex> g(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int per_cu = 0, cus = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, lds_bytes) != hipSuccess)
    return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  int cap = per_cu * cus;
  std::lock_guard<std::mutex> g(mu);
  cache[key] = cap;
  return cap;
}

__global__ void zero_words_kernel(unsigned long long* p, int64_t nwords) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nwords; i += stride) p[i] = 0ull;
}

int launch_zero(void* p, size_t bytes, hipStream_t s) {
  int64_t nwords = (int64_t)((bytes + 7) / 8);
  if (nwords == 0) return NBG_OK;
  int64_t blocks = (nwords + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                     (unsigned long long*)p, nwords);
  return check_hip(hipGetLastError(), "zero_words_kernel launch");
}

// Streaming copy used to calibrate bandwidth measurements: the practical HBM
// ceiling of a read+write stream, and the scale of the FETCH_SIZE/WRITE_SIZE
// counters for a kernel of known traffic. One 16-byte vector per thread,
// streaming (non-temporal) load and store, and a one-shot grid: short-lived
// independent workgroups are the fastest copy on MI355X (a grid-stride loop
// keeps fewer loads in flight per byte moved; plain loads and stores reach
// 6.26 TB/s against 6.51).
__global__ __launch_bounds__(256) void stream_copy_kernel(const uint4* __restrict__ src,
                                                          uint4* __restrict__ dst, int64_t n16) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n16) {
    const uint4 v = load_stream(src + i);
    const unsigned int w[4] = {v.x, v.y, v.z, v.w};
    store_stream16(reinterpret_cast<unsigned int*>(dst + i), w);
  }
}

}  // namespace nbg

extern "C" {

int nbg_stream_copy(const void* src, void* dst, size_t bytes, void* stream) {
  if (src == nullptr || dst == nullptr || bytes % 16 != 0 || ((uintptr_t)src % 16) || ((uintptr_t)dst % 16))
    return nbg::set_error(NBG_ERR_INVALID, "nbg_stream_copy: 16-byte aligned pointers and size required");
  const int64_t n16 = (int64_t)(bytes / 16);
  if (n16 == 0) return NBG_OK;
  // one-shot grids of at most 2**31 - 1 workgroups, one vector per thread
  constexpr int64_t kMaxVecs = (int64_t)0x7fffffff * 256;
  for (int64_t off = 0; off < n16; off += kMaxVecs) {
    const int64_t cnt = n16 - off < kMaxVecs ? n16 - off : kMaxVecs;
    const int64_t blocks = (cnt + 255) / 256;
    hipLaunchKernelGGL(nbg::stream_copy_kernel, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                       (const uint4*)src + off, (uint4*)dst + off, cnt);
    if (int rc = nbg::check_hip(hipGetLastError(), "stream_copy_kernel launch")) return rc;
  }
  return NBG_OK;
}

int nbg_tune_set(const char* name, int64_t value) {
  if (name == nullptr || strncmp(name, "NBG_", 4) != 0 || strlen(name) >= sizeof(nbg::KnobEntry::name))
    return nbg::set_error(NBG_ERR_INVALID, "nbg_tune_set: knob names are NBG_* (under 40 characters)");
  for (int i = 0; i < nbg::t_nknobs; ++i)
    if (strcmp(nbg::t_knobs[i].name, name) == 0) {
      nbg::t_knobs[i].value = value;
      return NBG_OK;
    }
  if (nbg::t_nknobs == nbg::kMaxKnobs) return nbg::set_error(NBG_ERR_INVALID, "nbg_tune_set: too many knobs");
  strcpy(nbg::t_knobs[nbg::t_nknobs].name, name);
  nbg::t_knobs[nbg::t_nknobs].value = value;
  ++nbg::t_nknobs;
  return NBG_OK;
}

int nbg_tune_clear(void) {
  nbg::t_nknobs = 0;
  return NBG_OK;
}

int64_t nbg_tune_get(const char* name, int64_t dflt) { return name ? nbg::knob(name, dflt) : dflt; }

const char* nbg_version(void) { return "numbagg_amd 0.1.0 (gfx950)"; }

const char* nbg_last_error(void) { return nbg::t_err; }

int nbg_spin_timeouts(unsigned int* count) {
  if (count == nullptr) return nbg::set_error(NBG_ERR_INVALID, "nbg_spin_timeouts: null pointer");
  NBG_HIP_TRY(hipMemcpyFromSymbol(count, HIP_SYMBOL(nbg::g_spin_timeouts), sizeof(unsigned int), 0,
                                  hipMemcpyDeviceToHost));
  return NBG_OK;
}

int nbg_reset_spin_timeouts(void) {
  unsigned int z = 0;
  NBG_HIP_TRY(hipMemcpyToSymbol(HIP_SYMBOL(nbg::g_spin_timeouts), &z, sizeof(z), 0,
                                hipMemcpyHostToDevice));
  return NBG_OK;
}

}  // extern "C"
