// Shared infrastructure of libdeco_hip.so: error state, the library stream, device buffers,
// launch helpers and wave64 reduction primitives for gfx950.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <string>
#include <vector>
#include <stdexcept>
#include <type_traits>
#include "deco_hip.h"

namespace deco {

// ------------------------------------------------------------------ errors -------------
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const char* fmt, ...);
void set_error(const char* msg);

#define DECO_HIP(expr)                                                                  \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      ::deco::fail(e_ == hipErrorOutOfMemory ? DECO_ERR_NOMEM : DECO_ERR_DEVICE,        \
                   "%s:%d %s: %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));   \
  } while (0)

#define DECO_CHECK_LAUNCH() DECO_HIP(hipGetLastError())

// Wraps a C-ABI body: converts exceptions into status codes.
template <class F>
int guard(F&& f) {
  try {
    f();
    return DECO_OK;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return DECO_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return DECO_ERR_ARG;
  }
}

// ------------------------------------------------------------------ context ------------
struct Context {
  int device = -1;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  size_t lds_per_block = 0;
  // pinned host staging for scalar readbacks (coherent: kernels may write it through
  // pinned_dev, its device address, and the host reads it after an event behind them or once
  // the matching tags below are set)
  double* pinned = nullptr;
  double* pinned_dev = nullptr;
  // per-value sequence tags beside it (coherent pinned): a kernel stores tags[v] = seq after
  // value v and a system-scope fence, so the host may read value v once its tag reads seq
  long long* tags = nullptr;
  long long* tags_dev = nullptr;
  long long tag_seq = 0;
  // library buffers come from the device's stream-ordered pool (allocation and release queued
  // on the library stream: no device-wide synchronisation, freed blocks reused)
  bool pooled = false;
};
Context& ctx();
void ensure_init();
inline hipStream_t stream() { ensure_init(); return ctx().stream; }
void sync();

// Phase timer tree (profile.hip): no-ops unless enabled with deco_profile_enable.
void prof_start(const char* name);
void prof_stop();
// a phase that ends at end() or, at the latest, when it goes out of scope (also on a throw)
struct ProfPhase {
  bool open = true;
  explicit ProfPhase(const char* name) { prof_start(name); }
  void end() {
    if (open) { open = false; prof_stop(); }
  }
  ~ProfPhase() {
    try { end(); } catch (...) {}
  }
  ProfPhase(const ProfPhase&) = delete;
  ProfPhase& operator=(const ProfPhase&) = delete;
};

// Algorithmic bytes of the device work queued while counting is on (deco_count_bytes): each
// launch adds what it must move at least once -- its matrix's stored values and indices, its row
// data, one pass over every vector it reads or writes -- and a copy or fill its bytes.
struct ByteCount {
  bool on = false;
  double bytes = 0.0;
  long long launches = 0;
};
ByteCount& byte_count();
inline void count_bytes(double b) {
  ByteCount& c = byte_count();
  if (c.on) {
    c.bytes += b;
    ++c.launches;
  }
}

// ------------------------------------------------------------------ device buffer ------
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    ensure_init();
    if (count) {
      if (ctx().pooled) DECO_HIP(hipMallocAsync(reinterpret_cast<void**>(&p), count * sizeof(T), ctx().stream));
      else DECO_HIP(hipMalloc(&p, count * sizeof(T)));
    }
    n = count;
  }
  void ensure(size_t count) { if (count > n) alloc(count); }
  void release() {
    if (p) {
      if (ctx().pooled && ctx().stream) (void)hipFreeAsync(p, ctx().stream);
      else (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  T* release_ownership() { T* q = p; p = nullptr; n = 0; return q; }
  void upload(const T* h, size_t count) {
    ensure(count);
    if (count) DECO_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, stream()));
  }
  void download(T* h, size_t count) const {
    if (count) DECO_HIP(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
    sync();
  }
  void zero() { if (n) DECO_HIP(hipMemsetAsync(p, 0, n * sizeof(T), stream())); }
  std::vector<T> to_host() const { std::vector<T> h(n); download(h.data(), n); return h; }
};

template <class T>
void copy_d2d(T* dst, const T* src, size_t count) {
  if (count) DECO_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToDevice, stream()));
  count_bytes(2.0 * sizeof(T) * count);
}
// clears bytes of device memory on the library stream (counted as written bytes)
inline void zero_async(void* p, size_t bytes) {
  if (bytes) DECO_HIP(hipMemsetAsync(p, 0, bytes, stream()));
  count_bytes((double)bytes);
}
template <class T>
void copy_h2d(T* dst, const T* src, size_t count) {
  if (count) DECO_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyHostToDevice, stream()));
}
template <class T>
void copy_d2h(T* dst, const T* src, size_t count) {
  if (count) DECO_HIP(hipMemcpyAsync(dst, src, count * sizeof(T), hipMemcpyDeviceToHost, stream()));
  sync();
}

inline unsigned grid_for(int64_t work, int block) {
  int64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  return (unsigned)g;
}

// ------------------------------------------------------------------ wave64 reductions ---
// Sum over the G lanes of a lane group (G a power of two <= 64, groups aligned in the wave).
template <int G, class T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, G);
  return v;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) { return group_sum<64>(v); }

// Workgroups are dispatched round-robin over the 8 XCDs (workgroup i runs on XCD i % 8), each
// with its own L2. xcd_block maps the hardware index to a logical one so that every XCD gets a
// contiguous range of logical blocks (a bijection on [0, G)): neighbouring row chunks, and the
// x entries they gather, then share one L2.
__device__ __forceinline__ int xcd_block(int i, int G) {
  const int q = G >> 3, r = G & 7, x = i & 7, k = i >> 3;
  return (x < r) ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}

// Appends `item` to list `bin` (lists[bin*stride + pos], pos from cnt[bin]); bin < 0 appends
// nothing. One atomic per wavefront and bin: lanes with the same bin take consecutive slots
// from a single atomicAdd by the lowest such lane. Every lane of the wave must call it.
template <int NBINS>
__device__ __forceinline__ void wave_bin_append(int bin, int item, int* cnt, int* lists, size_t stride) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int q = 0; q < NBINS; ++q) {
    const unsigned long long m = __ballot(bin == q);
    if (m == 0) continue;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(&cnt[q], __popcll(m));
    base = __shfl(base, leader);
    if (bin == q) lists[(size_t)q * stride + base + __popcll(m & ((1ull << lane) - 1ull))] = item;
  }
}

// A HIP event owned by its scope (destroyed on every exit path, errors included).
struct Event {
  hipEvent_t e = nullptr;
  explicit Event(unsigned flags = hipEventDefault) { DECO_HIP(hipEventCreateWithFlags(&e, flags)); }
  ~Event() {
    if (e) (void)hipEventDestroy(e);
  }
  Event(const Event&) = delete;
  Event& operator=(const Event&) = delete;
};

// A device-array argument of the C ABI: NULL is an error unless the array is empty (a kernel
// would fault on it).
inline void need(const void* p, int64_t count, const char* what) {
  if (count > 0 && !p) fail(DECO_ERR_ARG, "null %s", what);
}

// Block-wide sum of one value per thread, fixed order (deterministic). BLOCK multiple of 64.
template <int BLOCK, class T>
__device__ __forceinline__ T block_sum(T v, T* smem /* BLOCK/64 entries */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  T s = T(0);
  if (threadIdx.x < 64) {
    s = (lane < BLOCK / 64) ? smem[lane] : T(0);
    s = wave_sum(s);
  }
  __syncthreads();
  return s;  // valid in wave 0
}

// Fixed-grid size for reductions over n elements: depends only on n, so results are
// bit-reproducible run to run.
constexpr int kRedBlock = 256;
constexpr int kRedMaxGrid = 1024;
inline int red_grid(int64_t n) {
  int64_t g = (n + kRedBlock * 4 - 1) / (kRedBlock * 4);
  if (g < 1) g = 1;
  if (g > kRedMaxGrid) g = kRedMaxGrid;
  return (int)g;
}

// Sum `count` partials in a fixed order with the first wave of the block; every block that
// calls it gets the bitwise-same value.
__device__ __forceinline__ double reduce_partials(const double* __restrict__ part, int count) {
  __shared__ double sh_rp[1];
  if (threadIdx.x < 64) {
    double s = 0.0;
    for (int k = threadIdx.x; k < count; k += 64) s += part[k];
    s = wave_sum(s);
    if (threadIdx.x == 0) sh_rp[0] = s;
  }
  __syncthreads();
  double r = sh_rp[0];
  __syncthreads();
  return r;
}

}  // namespace deco
