// Runtime (errors, stream, memory) and the device CSR container with its kernels:
// create/copy/download, filters, pressure extraction, transpose, SpMV.
// Reference: core/CSR.hpp (container and methods), core/alg.hpp:51-64 (spmv).
#include "csr.hpp"
#include <algorithm>
#include <climits>
#include <memory>
#include <cstring>
#include <mutex>

namespace deco {

static thread_local std::string t_last_error;
static Context g_ctx;
static std::mutex g_init_mu;

void set_error(const char* msg) { t_last_error = msg ? msg : ""; }

void fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

Context& ctx() { return g_ctx; }

static void init_device(int device) {
  std::lock_guard<std::mutex> lk(g_init_mu);
  if (g_ctx.stream && (device < 0 || device == g_ctx.device)) return;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) fail(DECO_ERR_DEVICE, "no HIP device available (%s)", hipGetErrorString(e));
  if (device < 0) device = 0;
  if (device >= count) fail(DECO_ERR_ARG, "device %d out of range (%d devices)", device, count);
  DECO_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  DECO_HIP(hipGetDeviceProperties(&prop, device));
  if (g_ctx.stream) { (void)hipStreamDestroy(g_ctx.stream); g_ctx.stream = nullptr; }
  // default-priority, non-blocking stream
  DECO_HIP(hipStreamCreateWithFlags(&g_ctx.stream, hipStreamNonBlocking));
  // stream-ordered pool for the library's buffers, keeping freed memory for reuse
  hipMemPool_t pool = nullptr;
  if (!g_ctx.pooled && hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
    uint64_t keep = UINT64_MAX;
    g_ctx.pooled = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep) == hipSuccess;
  }
  (void)hipGetLastError();
  if (!g_ctx.pinned) {
    DECO_HIP(hipHostMalloc(&g_ctx.pinned, 4096 * sizeof(double), hipHostMallocCoherent));
    DECO_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&g_ctx.pinned_dev), g_ctx.pinned, 0));
  }
  if (!g_ctx.tags) {
    DECO_HIP(hipHostMalloc(reinterpret_cast<void**>(&g_ctx.tags), 4096 * sizeof(long long), hipHostMallocCoherent));
    std::memset(g_ctx.tags, 0, 4096 * sizeof(long long));
    DECO_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&g_ctx.tags_dev), g_ctx.tags, 0));
  }
  g_ctx.device = device;
  g_ctx.num_cu = prop.multiProcessorCount;
  g_ctx.lds_per_block = prop.sharedMemPerBlock;
}

void ensure_init() {
  if (!g_ctx.stream) init_device(-1);
}

void sync() { DECO_HIP(hipStreamSynchronize(stream())); }

// ------------------------------------------------------------------ scan ---------------
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

// inclusive scan of one int per thread over the block
__device__ __forceinline__ int block_inclusive_scan(int v, int* smem /* 4 */) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int t = __shfl_up(v, off, 64);
    if (lane >= off) v += t;
  }
  if (lane == 63) smem[wid] = v;
  __syncthreads();
  int add = 0;
  for (int w = 0; w < wid; ++w) add += smem[w];
  __syncthreads();
  return v + add;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(const int* __restrict__ in, int* __restrict__ out,
                                                          int n, int* __restrict__ tile_sums) {
  __shared__ int sm[4];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  int vals[kScanItems];
  int tsum = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    vals[k] = (i < n) ? in[i] : 0;
    tsum += vals[k];
  }
  int incl = block_inclusive_scan(tsum, sm);
  int run = incl - tsum;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    int64_t i = base + k;
    if (i < n) out[i] = run;
    run += vals[k];
  }
  if (threadIdx.x == kScanBlock - 1) tile_sums[blockIdx.x] = incl;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_add(int* __restrict__ out, int n,
                                                        const int* __restrict__ tile_off) {
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  const int add = tile_off[blockIdx.x];
  for (int k = threadIdx.x; k < kScanTile; k += kScanBlock) {
    int64_t i = base + k;
    if (i < n) out[i] += add;
  }
}

__global__ void k_set_total(const int* __restrict__ in, int* __restrict__ out, int n) {
  // out[n] = out[n-1] + in[n-1]  (in already overwritten when aliasing is handled by caller)
  if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = (n > 0) ? out[n - 1] + in[0] : 0;
}

// Exclusive scan of in[0..n) into out[0..n) (out may alias in), recursive over tile sums.
static void exclusive_scan(const int* in, int* out, int n) {
  if (n <= 0) return;
  const int tiles = (n + kScanTile - 1) / kScanTile;
  DBuf<int> sums(tiles + 1);
  hipLaunchKernelGGL(k_scan_tiles, dim3(tiles), dim3(kScanBlock), 0, stream(), in, out, n, sums.p);
  DECO_CHECK_LAUNCH();
  if (tiles > 1) {
    exclusive_scan(sums.p, sums.p, tiles);
    hipLaunchKernelGGL(k_scan_add, dim3(tiles), dim3(kScanBlock), 0, stream(), out, n, sums.p);
    DECO_CHECK_LAUNCH();
  }
}

__global__ void k_last_plus(const int* __restrict__ last_in, int* __restrict__ out, int n) {
  if (threadIdx.x == 0 && blockIdx.x == 0) out[n] = out[n - 1] + last_in[0];
}

void scan_to_ptr_async(const int* in, int* out, int n) {
  if (n <= 0) {
    DECO_HIP(hipMemsetAsync(out, 0, sizeof(int), stream()));
    return;
  }
  // keep the last count before a possibly aliasing scan overwrites it
  DBuf<int> last(1);
  copy_d2d(last.p, in + (n - 1), 1);
  exclusive_scan(in, out, n);
  hipLaunchKernelGGL(k_last_plus, dim3(1), dim3(1), 0, stream(), last.p, out, n);
  DECO_CHECK_LAUNCH();
}

int64_t scan_to_ptr(const int* in, int* out, int n) {
  scan_to_ptr_async(in, out, n);
  int tot = 0;
  copy_d2h(&tot, out + n, 1);
  return tot;
}

deco_csr* csr_new(int n, int m) {
  auto* A = new deco_csr();
  A->n = n;
  A->m = m;
  return A;
}

int build_row_ptr(deco_csr* dst, const int* counts, int n) {
  dst->r.alloc(n + 1);
  int64_t nnz = scan_to_ptr(counts, dst->r.p, n);
  if (nnz > INT_MAX) fail(DECO_ERR_UNSUPPORTED, "nnz %lld exceeds int32", (long long)nnz);
  dst->nnz = (int)nnz;
  dst->c.alloc(nnz);
  dst->v.alloc(nnz);
  dst->invalidate();
  return (int)nnz;
}

// ------------------------------------------------------------------ row sort -----------
__global__ void k_row_len(int n, const int* __restrict__ r, int* __restrict__ len) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) len[i] = r[i + 1] - r[i];
}

__global__ void k_max_reduce(const int* __restrict__ a, int n, int* __restrict__ out) {
  __shared__ int sm[4];
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, a[i]);
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    int mm = sm[0];
    for (int w = 1; w < (int)(blockDim.x / 64); ++w) mm = max(mm, sm[w]);
    atomicMax(out, mm);
  }
}

int max_row_length(const deco_csr* A) {
  if (A->n == 0) return 0;
  DBuf<int> len(A->n), mx(1);
  mx.zero();
  hipLaunchKernelGGL(k_row_len, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, A->r.p, len.p);
  DECO_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_max_reduce, dim3(std::min<unsigned>(1024, grid_for(A->n, 256))), dim3(256), 0, stream(),
                     len.p, A->n, mx.p);
  DECO_CHECK_LAUNCH();
  int h = 0;
  copy_d2h(&h, mx.p, 1);
  return h;
}

// insertion sort per row for short rows (one thread per row)
__global__ void k_sort_rows_small(int n, const int* __restrict__ r, int* __restrict__ c, double* __restrict__ v,
                                  int limit) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = r[i], e = r[i + 1];
  if (e - b > limit) return;
  for (int j = b + 1; j < e; ++j) {
    int key = c[j];
    double val = v ? v[j] : 0.0;
    int k = j - 1;
    while (k >= b && c[k] > key) {
      c[k + 1] = c[k];
      if (v) v[k + 1] = v[k];
      --k;
    }
    c[k + 1] = key;
    if (v) v[k + 1] = val;
  }
}

// bitonic sort in LDS for rows longer than the small limit (one block per row, <= 4096)
constexpr int kSortMax = 4096;
__global__ __launch_bounds__(256) void k_sort_rows_large(int n, const int* __restrict__ r, int* __restrict__ c,
                                                         double* __restrict__ v, int limit) {
  __shared__ int sk[kSortMax];
  __shared__ double sv[kSortMax];
  const int i = blockIdx.x;
  const int b = r[i], e = r[i + 1], len = e - b;
  if (len <= limit) return;
  int P = 1;
  while (P < len) P <<= 1;
  for (int k = threadIdx.x; k < P; k += 256) {
    sk[k] = (k < len) ? c[b + k] : INT_MAX;
    sv[k] = (k < len && v) ? v[b + k] : 0.0;
  }
  __syncthreads();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = threadIdx.x; k < P; k += 256) {
        int partner = k ^ stride;
        if (partner > k) {
          bool up = ((k & size) == 0);
          if ((sk[k] > sk[partner]) == up) {
            int tk = sk[k]; sk[k] = sk[partner]; sk[partner] = tk;
            double tv = sv[k]; sv[k] = sv[partner]; sv[partner] = tv;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int k = threadIdx.x; k < len; k += 256) {
    c[b + k] = sk[k];
    if (v) v[b + k] = sv[k];
  }
}

void sort_rows(int n, const int* r, int* c, double* v, int max_row) {
  if (n <= 0) return;
  const int small = 32;
  hipLaunchKernelGGL(k_sort_rows_small, dim3(grid_for(n, 128)), dim3(128), 0, stream(), n, r, c, v, small);
  DECO_CHECK_LAUNCH();
  if (max_row > small) {
    if (max_row > kSortMax) fail(DECO_ERR_UNSUPPORTED, "row of %d entries exceeds the sort limit %d", max_row, kSortMax);
    hipLaunchKernelGGL(k_sort_rows_large, dim3(n), dim3(256), 0, stream(), n, r, c, v, small);
    DECO_CHECK_LAUNCH();
  }
}

// ------------------------------------------------------------------ transpose ----------
__global__ void k_col_count(int nnz, const int* __restrict__ c, int* __restrict__ cnt) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nnz) atomicAdd(&cnt[c[j]], 1);
}

__global__ void k_transpose_scatter(int n, const int* __restrict__ r, const int* __restrict__ c,
                                    const double* __restrict__ v, int* __restrict__ cursor,
                                    int* __restrict__ ct, double* __restrict__ vt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = r[i]; j < r[i + 1]; ++j) {
    int pos = atomicAdd(&cursor[c[j]], 1);
    ct[pos] = i;
    vt[pos] = v[j];
  }
}

deco_csr* csr_transpose(const deco_csr* A) {
  deco_csr* T = csr_new(A->m, A->n);
  DBuf<int> cnt(A->m + 1);
  cnt.zero();
  if (A->nnz > 0) {
    hipLaunchKernelGGL(k_col_count, dim3(grid_for(A->nnz, 256)), dim3(256), 0, stream(), A->nnz, A->c.p, cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(T, cnt.p, A->m);
  if (T->nnz > 0) {
    DBuf<int> cursor(A->m + 1);
    copy_d2d(cursor.p, T->r.p, A->m + 1);
    hipLaunchKernelGGL(k_transpose_scatter, dim3(grid_for(A->n, 128)), dim3(128), 0, stream(), A->n, A->r.p,
                       A->c.p, A->v.p, cursor.p, T->c.p, T->v.p);
    DECO_CHECK_LAUNCH();
    sort_rows(T->n, T->r.p, T->c.p, T->v.p, max_row_length(T));
  }
  return T;
}

deco_csr* csr_clone(const deco_csr* A) {
  deco_csr* B = csr_new(A->n, A->m);
  B->nnz = A->nnz;
  B->r.alloc(A->n + 1);
  B->c.alloc(A->nnz);
  B->v.alloc(A->nnz);
  copy_d2d(B->r.p, A->r.p, A->n + 1);
  copy_d2d(B->c.p, A->c.p, A->nnz);
  copy_d2d(B->v.p, A->v.p, A->nnz);
  return B;
}

// ------------------------------------------------------------------ filters ------------
// Diagonal position search: core::binarysearch (core/alg.hpp:16-25) returns the last index
// whose column is <= target inside [lower, upper). Reproduced exactly (used for the diagonal
// lookup of filter_proper, CSR.hpp:486).
__device__ __forceinline__ int dev_binarysearch(const int* __restrict__ arr, int lower, int upper, int target) {
  int mn = lower, mx = upper, mid = (lower + upper) >> 1;
  while (mx - mn > 1) {
    if (arr[mid] > target) mx = mid; else mn = mid;
    mid = (mn + mx) >> 1;
  }
  return mid;
}

template <int MODE>  // 0: |a| > tol ; 1: |a| > |a_ii * tol|
__global__ void k_filter_count(int n, const int* __restrict__ r, const int* __restrict__ c,
                               const double* __restrict__ v, double tol, int* __restrict__ cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = r[i], e = r[i + 1];
  double D = tol;
  if (MODE == 1) D = fabs(v[dev_binarysearch(c, b, e, i)] * tol);
  int k = 0;
  for (int j = b; j < e; ++j)
    if (fabs(v[j]) > D || c[j] == i) ++k;
  cnt[i] = k;
}

template <int MODE>
__global__ void k_filter_fill(int n, const int* __restrict__ r, const int* __restrict__ c,
                              const double* __restrict__ v, double tol, const int* __restrict__ rn,
                              int* __restrict__ cn, double* __restrict__ vn) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = r[i], e = r[i + 1];
  double D = tol;
  if (MODE == 1) D = fabs(v[dev_binarysearch(c, b, e, i)] * tol);
  int o = rn[i];
  for (int j = b; j < e; ++j)
    if (fabs(v[j]) > D || c[j] == i) {
      cn[o] = c[j];
      vn[o] = v[j];
      ++o;
    }
}

template <int MODE>
static void filter_impl(deco_csr* A, double tol) {
  const int n = A->n;
  if (n == 0) return;
  DBuf<int> cnt(n + 1);
  hipLaunchKernelGGL((k_filter_count<MODE>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, A->r.p, A->c.p,
                     A->v.p, tol, cnt.p);
  DECO_CHECK_LAUNCH();
  deco_csr tmp;
  tmp.n = n;
  tmp.m = A->m;
  build_row_ptr(&tmp, cnt.p, n);
  hipLaunchKernelGGL((k_filter_fill<MODE>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, A->r.p, A->c.p,
                     A->v.p, tol, tmp.r.p, tmp.c.p, tmp.v.p);
  DECO_CHECK_LAUNCH();
  A->r = std::move(tmp.r);
  A->c = std::move(tmp.c);
  A->v = std::move(tmp.v);
  A->nnz = tmp.nnz;
  A->invalidate();
  if (A->At) { delete A->At; A->At = nullptr; }
}

// ------------------------------------------------------------------ local_matrix -------
__global__ void k_local_count(int nl, int bs, int p, const int* __restrict__ r, const int* __restrict__ c,
                              int* __restrict__ cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  const int row = bs * i + p;
  int k = 0;
  for (int j = r[row]; j < r[row + 1]; ++j)
    if (c[j] % bs == p) ++k;
  cnt[i] = k;
}

__global__ void k_local_fill(int nl, int bs, int p, const int* __restrict__ r, const int* __restrict__ c,
                             const double* __restrict__ v, const double* __restrict__ w,
                             const int* __restrict__ rl, int* __restrict__ cl, double* __restrict__ vl) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  const int row = bs * i + p;
  int o = rl[i];
  const double scale = w ? w[row] : 1.0;
  for (int j = r[row]; j < r[row + 1]; ++j)
    if (c[j] % bs == p) {
      cl[o] = c[j] / bs;
      vl[o] = w ? v[j] / scale : v[j];
      ++o;
    }
}

deco_csr* local_matrix(const deco_csr* A, int bs, int p, const double* w) {
  if (bs <= 0 || A->n % bs != 0 || A->m % bs != 0)
    fail(DECO_ERR_ARG, "n or m are not divisible by bs: n=%d m=%d bs=%d", A->n, A->m, bs);
  if (p < 0 || p >= bs) fail(DECO_ERR_ARG, "pressure id %d out of range for bs %d", p, bs);
  const int nl = A->n / bs;
  deco_csr* L = csr_new(nl, A->m / bs);
  DBuf<int> cnt(nl + 1);
  if (nl > 0) {
    hipLaunchKernelGGL(k_local_count, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, bs, p, A->r.p, A->c.p,
                       cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(L, cnt.p, nl);
  if (nl > 0) {
    hipLaunchKernelGGL(k_local_fill, dim3(grid_for(nl, 256)), dim3(256), 0, stream(), nl, bs, p, A->r.p, A->c.p,
                       A->v.p, w, L->r.p, L->c.p, L->v.p);
    DECO_CHECK_LAUNCH();
  }
  return L;
}

// ------------------------------------------------------------------ row subset --------
__global__ void k_subset_count(int ns, int stride, int off, const int* __restrict__ r, int* __restrict__ cnt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < ns) cnt[i] = r[i * stride + off + 1] - r[i * stride + off];
}
// one wavefront per subset row copies its entries
__global__ void k_subset_fill(int ns, int stride, int off, const int* __restrict__ r, const int* __restrict__ c,
                              const double* __restrict__ v, const int* __restrict__ rs, int* __restrict__ cs,
                              double* __restrict__ vs) {
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (i >= ns) return;
  const int b = r[i * stride + off], e = r[i * stride + off + 1], o = rs[i];
  for (int k = b + lane; k < e; k += 64) {
    cs[o + k - b] = c[k];
    vs[o + k - b] = v[k];
  }
}

deco_csr* csr_row_subset(const deco_csr* A, int stride, int off) {
  if (stride <= 0 || off < 0 || off >= stride || A->n % stride != 0)
    fail(DECO_ERR_ARG, "row subset stride %d offset %d does not fit n=%d", stride, off, A->n);
  const int ns = A->n / stride;
  deco_csr* S = csr_new(ns, A->m);
  DBuf<int> cnt(ns + 1);
  if (ns > 0) {
    hipLaunchKernelGGL(k_subset_count, dim3(grid_for(ns, 256)), dim3(256), 0, stream(), ns, stride, off, A->r.p,
                       cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(S, cnt.p, ns);
  if (ns > 0) {
    hipLaunchKernelGGL(k_subset_fill, dim3((ns + 3) / 4), dim3(256), 0, stream(), ns, stride, off, A->r.p, A->c.p,
                       A->v.p, S->r.p, S->c.p, S->v.p);
    DECO_CHECK_LAUNCH();
  }
  return S;
}

// per row: its entries in the columns stride*j + off
__global__ void k_colsub_count(int n, int stride, int off, const int* __restrict__ r, const int* __restrict__ c,
                               int* __restrict__ cnt) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int k = 0;
  for (int j = r[i]; j < r[i + 1]; ++j) k += (c[j] % stride == off);
  cnt[i] = k;
}
// one thread per row: those entries with column (c - off) / stride, in order
__global__ void k_colsub_fill(int n, int stride, int off, const int* __restrict__ r, const int* __restrict__ c,
                              const double* __restrict__ v, const int* __restrict__ r2, int* __restrict__ c2,
                              double* __restrict__ v2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int o = r2[i];
  for (int j = r[i]; j < r[i + 1]; ++j)
    if (c[j] % stride == off) {
      c2[o] = (c[j] - off) / stride;
      v2[o] = v[j];
      ++o;
    }
}

deco_csr* csr_col_subset(const deco_csr* A, int stride, int off) {
  if (stride <= 0 || off < 0 || off >= stride || A->m % stride != 0)
    fail(DECO_ERR_ARG, "column subset stride %d offset %d does not fit m=%d", stride, off, A->m);
  std::unique_ptr<deco_csr> S(csr_new(A->n, A->m / stride));
  DBuf<int> cnt(A->n + 1);
  if (A->n > 0) {
    hipLaunchKernelGGL(k_colsub_count, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, stride, off, A->r.p,
                       A->c.p, cnt.p);
    DECO_CHECK_LAUNCH();
  }
  build_row_ptr(S.get(), cnt.p, A->n);
  if (A->n > 0) {
    hipLaunchKernelGGL(k_colsub_fill, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, stride, off, A->r.p,
                       A->c.p, A->v.p, S->r.p, S->c.p, S->v.p);
    DECO_CHECK_LAUNCH();
  }
  return S.release();
}

// ------------------------------------------------------------------ in-cell reorder ---
// Unknown k of a cell moves to position reord[k] (rows: reord_row, columns: reord_col);
// new row i holds old row (i/bs)*bs + ord_row[i%bs] (adapter/csr2coo.cpp:31-37,56-67,82-86).
struct CellPerm {
  int ord_row[9];
  int reord_col[9];
};

__global__ void k_reorder_len(int n, int bs, CellPerm p, const int* __restrict__ r, int* __restrict__ len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = (i / bs) * bs + p.ord_row[i % bs];
  len[i] = r[o + 1] - r[o];
}

// One thread per new row: copy the old row with mapped columns. The old row is sorted and the
// map permutes columns only inside aligned groups of bs, so sorting the runs of one column cell
// gives the row sorted by column, the order of the reference's (row, col) sort
// (csr2coo.cpp:69-72).
__global__ void k_reorder_fill(int n, int bs, CellPerm p, const int* __restrict__ r, const int* __restrict__ c,
                               const double* __restrict__ v, const int* __restrict__ r2, int* __restrict__ c2,
                               double* __restrict__ v2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int o = (i / bs) * bs + p.ord_row[i % bs];
  const int b = r[o], e = r[o + 1], d = r2[i];
  int run = d;
  for (int k = b; k < e; ++k) {
    const int col = c[k], cell = col / bs, rest = col - cell * bs;
    const int nc = cell * bs + p.reord_col[rest];
    const double val = v[k];
    int q = d + (k - b);
    if (k > b && c[k - 1] / bs != cell) run = q;
    while (q > run && c2[q - 1] > nc) {   // insertion inside the current column cell
      c2[q] = c2[q - 1];
      v2[q] = v2[q - 1];
      --q;
    }
    c2[q] = nc;
    v2[q] = val;
  }
}

__global__ void k_reorder_vec(int n, int bs, CellPerm p, const double* __restrict__ b, double* __restrict__ b2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b2[i] = b[(i / bs) * bs + p.ord_row[i % bs]];
}

static void parse_order(const char* s, int bs, int* reord, int* ord, const char* what) {
  if (!s) fail(DECO_ERR_ARG, "%s is null", what);
  int seen[9] = {0};
  for (int i = 0; i < bs; ++i) {
    const int d = s[i] - '1';
    if (s[i] == 0 || d < 0 || d >= bs || seen[d])
      fail(DECO_ERR_ARG, "%s \"%s\" is not a permutation of 1..%d", what, s, bs);
    seen[d] = 1;
    reord[d] = i;
    if (ord) ord[i] = d;
  }
}

void block_reorder(deco_csr* A, int bs, const char* row_order, const char* col_order, double* b) {
  if (bs < 1 || bs > 9) fail(DECO_ERR_UNSUPPORTED, "block size %d (1..9 supported)", bs);
  if (A->n % bs != 0 || A->m % bs != 0) fail(DECO_ERR_ARG, "n=%d m=%d not divisible by bs=%d", A->n, A->m, bs);
  if (bs == 1) return;                                             // csr2coo.cpp:39
  CellPerm p{};
  int reord_row[9];
  parse_order(row_order, bs, reord_row, p.ord_row, "row_order");
  parse_order(col_order, bs, p.reord_col, nullptr, "col_order");
  if (bs == 2 && reord_row[0] == 0 && reord_row[1] == 1 && p.reord_col[0] == 0 && p.reord_col[1] == 1)
    return;                                                        // csr2coo.cpp:40-41
  const int n = A->n;
  if (n == 0) return;
  DBuf<int> len(n + 1), r2;
  hipLaunchKernelGGL(k_reorder_len, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, bs, p, A->r.p, len.p);
  DECO_CHECK_LAUNCH();
  r2.alloc(n + 1);
  scan_to_ptr_async(len.p, r2.p, n);
  DBuf<int> c2(A->nnz);
  DBuf<double> v2(A->nnz);
  hipLaunchKernelGGL(k_reorder_fill, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, bs, p, A->r.p, A->c.p,
                     A->v.p, r2.p, c2.p, v2.p);
  DECO_CHECK_LAUNCH();
  if (b) {
    DBuf<double> b2(n);
    hipLaunchKernelGGL(k_reorder_vec, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, bs, p, (const double*)b,
                       b2.p);
    DECO_CHECK_LAUNCH();
    copy_d2d(b, b2.p, n);
  }
  A->r = std::move(r2);
  A->c = std::move(c2);
  A->v = std::move(v2);
  A->invalidate();
  if (A->At) { delete A->At; A->At = nullptr; }
  sync();
}

void csr_filter(deco_csr* A, double tol) { filter_impl<0>(A, tol); }
void csr_filter_proper(deco_csr* A, double tol) { filter_impl<1>(A, tol); }

// ------------------------------------------------------------------ conversions --------
__global__ void k_d2f(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (float)a[i];
}

void to_f32(const double* a, float* b, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_d2f, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, a, b);
  DECO_CHECK_LAUNCH();
}

}  // namespace deco

deco_csr::~deco_csr() {
  delete At;
  delete b32;
  delete pc;
  delete cs;
}

void deco_csr::invalidate() {
  ++gen;
  vf_valid = false;
  rb_valid = false;
  sell_ok = -1;
  c16_ok = -1;
  delete b32;
  b32 = nullptr;
  delete pc;
  pc = nullptr;
  pc_ok = -1;
  delete cs;
  cs = nullptr;
  cs_ok = -1;
  gather_seg = -1.0;
}

namespace deco {
// number of distinct column blocks over the R rows of block row I (rows sorted by column)
__global__ void k_bsr_count(int nb, int R, int C, const int* __restrict__ r, const int* __restrict__ c,
                            int* __restrict__ cnt) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  int pos[3], end[3];
  for (int q = 0; q < R; ++q) { pos[q] = r[I * R + q]; end[q] = r[I * R + q + 1]; }
  int count = 0;
  while (true) {
    int mn = INT_MAX;
    for (int q = 0; q < R; ++q)
      if (pos[q] < end[q]) mn = min(mn, c[pos[q]] / C);
    if (mn == INT_MAX) break;
    ++count;
    for (int q = 0; q < R; ++q)
      while (pos[q] < end[q] && c[pos[q]] / C == mn) ++pos[q];
  }
  cnt[I] = count;
}

// block columns and block values (bv zeroed beforehand): row-major blocks one after another
// (ps == 0), or row-planar (element (rr, q) of block o at rr*ps + o*C + q)
template <class VT>
__global__ void k_bsr_fill(int nb, int R, int C, const int* __restrict__ r, const int* __restrict__ c,
                           const double* __restrict__ v, const int* __restrict__ br, int* __restrict__ bc,
                           VT* __restrict__ bv, size_t ps) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  int pos[3], end[3];
  for (int q = 0; q < R; ++q) { pos[q] = r[I * R + q]; end[q] = r[I * R + q + 1]; }
  int o = br[I];
  while (true) {
    int mn = INT_MAX;
    for (int q = 0; q < R; ++q)
      if (pos[q] < end[q]) mn = min(mn, c[pos[q]] / C);
    if (mn == INT_MAX) break;
    bc[o] = mn;
    for (int q = 0; q < R; ++q)
      for (; pos[q] < end[q] && c[pos[q]] / C == mn; ++pos[q])
        bv[ps ? q * ps + (size_t)o * C + (c[pos[q]] - mn * C) : (size_t)o * R * C + q * C + (c[pos[q]] - mn * C)] =
            (VT)v[pos[q]];
    ++o;
  }
}

// Largest padding (fraction of the blocks) for which block shadows take the sliced layout;
// negative: always the chunked one.
static double g_sell_padding = 0.10;
static bool g_sell_offsets16 = true;
static bool g_sell_dia = true;   // the uniform-offset form where it applies
int g_sell_pairs = 1;

// 16-bit offsets of the sliced layout's block columns from their block rows (bcd zeroed
// beforehand: padding entries and the lanes past the last block row keep 0); ovf is set when an
// offset does not fit
__global__ void k_sell_delta(int nb, const int* __restrict__ br, const int* __restrict__ soff,
                             const int* __restrict__ bc, short* __restrict__ bcd, int* __restrict__ ovf) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  const int cnt = br[I + 1] - br[I];
  size_t e = (size_t)soff[I / kSellRows] + (I % kSellRows);
  for (int k = 0; k < cnt; ++k, e += kSellRows) {
    const int d = bc[e] - I;
    if (d < -32768 || d > 32767) {
      atomicOr(ovf, 1);
      return;
    }
    bcd[e] = (short)d;
  }
}

// per slice of 64 block rows: its width (longest block row) and its entries (width * 64)
__global__ void k_sell_width(int nb, int nslices, const int* __restrict__ cnt, int* __restrict__ ent,
                             int* __restrict__ wid) {
  const int sl = blockIdx.x * blockDim.x + threadIdx.x;
  if (sl >= nslices) return;
  int w = 0;
  const int I1 = min(nb, (sl + 1) * kSellRows);
  for (int I = sl * kSellRows; I < I1; ++I) w = max(w, cnt[I]);
  wid[sl] = w;
  ent[sl] = w * kSellRows;
}

// sliced layout: block k of block row I at soff[I/64] + 64*k + I%64 (bc and bv zeroed beforehand)
template <class VT>
__global__ void k_sell_fill(int nb, int R, int C, const int* __restrict__ r, const int* __restrict__ c,
                            const double* __restrict__ v, const int* __restrict__ soff, size_t T,
                            int* __restrict__ bc, VT* __restrict__ bv) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  int pos[3], end[3];
  for (int q = 0; q < R; ++q) { pos[q] = r[I * R + q]; end[q] = r[I * R + q + 1]; }
  size_t e = (size_t)soff[I / kSellRows] + (I % kSellRows);
  while (true) {
    int mn = INT_MAX;
    for (int q = 0; q < R; ++q)
      if (pos[q] < end[q]) mn = min(mn, c[pos[q]] / C);
    if (mn == INT_MAX) break;
    bc[e] = mn;
    for (int q = 0; q < R; ++q)
      for (; pos[q] < end[q] && c[pos[q]] / C == mn; ++pos[q])
        bv[sell_vidx(C, T, q, c[pos[q]] - mn * C, e)] = (VT)v[pos[q]];
    e += kSellRows;
  }
}

// Distinct block-column offsets (block column - block row) of the R x C blocks of a CSR matrix,
// one row per thread: each workgroup collects its offsets in an LDS set of kDiaMax slots, then
// merges the set into the global table (kDiaMax slots, INT_MIN = free); ovf is set when either
// overflows.
__device__ __forceinline__ bool dia_insert(int* set, int d) {
  for (int q = 0; q < kDiaMax; ++q) {
    const int v = __atomic_load_n(&set[q], __ATOMIC_RELAXED);
    if (v == d) return true;
    if (v == INT_MIN) {
      const int old = atomicCAS(&set[q], INT_MIN, d);
      if (old == INT_MIN || old == d) return true;
    }
  }
  return false;
}
__global__ __launch_bounds__(256) void k_dia_offsets(int n, int R, int C, const int* __restrict__ r,
                                                     const int* __restrict__ c, int* __restrict__ table,
                                                     int* __restrict__ ovf) {
  __shared__ int set[kDiaMax];
  __shared__ int bad;
  if (threadIdx.x < kDiaMax) set[threadIdx.x] = INT_MIN;
  if (threadIdx.x == 0) bad = 0;
  __syncthreads();
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) {
    const int I = i / R;
    for (int j = r[i]; j < r[i + 1]; ++j)
      if (!dia_insert(set, c[j] / C - I)) {
        bad = 1;
        break;
      }
  }
  __syncthreads();
  if (threadIdx.x < kDiaMax) {
    const int d = set[threadIdx.x];
    if (d != INT_MIN && !dia_insert(table, d)) atomicOr(ovf, 1);
  }
  if (threadIdx.x == 0 && bad) atomicOr(ovf, 1);
}

// uniform-offset sliced layout: the values of block row I's block at block column I + doff[k] at
// entry soff[I/64] + 64*k + I%64 (bv zeroed beforehand)
template <class VT>
__global__ void k_dia_fill(int nb, int R, int C, int noff, const int* __restrict__ doff, const int* __restrict__ r,
                           const int* __restrict__ c, const double* __restrict__ v, const int* __restrict__ soff,
                           size_t T, VT* __restrict__ bv) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nb) return;
  const size_t e0 = (size_t)soff[I / kSellRows] + (I % kSellRows);
  for (int q = 0; q < R; ++q)
    for (int j = r[I * R + q]; j < r[I * R + q + 1]; ++j) {
      const int J = c[j] / C, d = J - I;
      int k = 0;
      while (k < noff - 1 && doff[k] != d) ++k;
      bv[sell_vidx(C, T, q, c[j] - J * C, e0 + (size_t)k * kSellRows)] = (VT)v[j];
    }
}

// *out = max(*out, longest row)
__global__ __launch_bounds__(256) void k_max_row(int n, const int* __restrict__ ptr, int* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int len = i < n ? ptr[i + 1] - ptr[i] : 0;
  for (int off = 32; off > 0; off >>= 1) len = max(len, __shfl_xor(len, off));
  if ((threadIdx.x & 63) == 0 && len > 0) atomicMax(out, len);
}

// rb[k] = (i, ptr[i]) for the first row i in [0, n] with ptr[i] >= k*cap; rb[nchunks] = (n, ptr[n])
__global__ void k_chunks_cap(int n, const int* __restrict__ ptr, int nchunks, int cap, int2* __restrict__ rb) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k > nchunks) return;
  if (k == nchunks) { rb[k] = make_int2(n, ptr[n]); return; }
  const int64_t target = (int64_t)k * cap;
  int lo = 0, hi = n;
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (ptr[mid] >= target) hi = mid; else lo = mid + 1;
  }
  rb[k] = make_int2(lo, ptr[lo]);
}
}  // namespace deco

namespace deco {
// The sorted block-column offsets of the R x C blocks of A when there are at most kDiaMax of them
// and the uniform-offset layout (nb block rows in slices of 64, noff entries each) stores at most
// (1 + padding) x nnzb blocks; else empty.
static std::vector<int> dia_offsets(const deco_csr* A, int R, int C, int nb, int64_t nnzb) {
  std::vector<int> off;
  if (!g_sell_dia || g_sell_padding < 0.0 || nb <= 0 || nnzb <= 0) return off;
  DBuf<int> table(kDiaMax), ovf(1);
  std::vector<int> h(kDiaMax, INT_MIN);
  table.upload(h.data(), kDiaMax);
  ovf.zero();
  hipLaunchKernelGGL(k_dia_offsets, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, R, C, A->r.p, A->c.p,
                     table.p, ovf.p);
  DECO_CHECK_LAUNCH();
  int o = 0;
  copy_d2h(&o, ovf.p, 1);
  if (o) return off;
  copy_d2h(h.data(), table.p, kDiaMax);
  for (int d : h)
    if (d != INT_MIN) off.push_back(d);
  std::sort(off.begin(), off.end());
  const int64_t ns = (nb + kSellRows - 1) / kSellRows;
  const int64_t T = ns * kSellRows * (int64_t)off.size();
  if (off.empty() || T > INT_MAX || (double)T > (double)nnzb * (1.0 + g_sell_padding)) off.clear();
  return off;
}
}  // namespace deco

const deco_bsr* deco_csr::block32(int R, int C) const {
  using namespace deco;
  if (b32 && b32->R == R && b32->C == C) return b32;
  delete b32;
  b32 = nullptr;
  if (R < 1 || R > 3 || C < 1 || C > 3 || n % R != 0 || m % C != 0) return nullptr;
  auto* B = new deco_bsr();
  B->R = R;
  B->C = C;
  B->nb = n / R;
  B->mb = m / C;
  const int nb = B->nb;
  DBuf<int> cnt(nb + 1);
  B->br.alloc(nb + 1);
  if (nb > 0) {
    hipLaunchKernelGGL(k_bsr_count, dim3(grid_for(nb, 256)), dim3(256), 0, stream(), nb, R, C, r.p, c.p, cnt.p);
    DECO_CHECK_LAUNCH();
  }
  const int64_t nnzb = scan_to_ptr(cnt.p, B->br.p, nb);
  B->nnzb = (int)nnzb;
  const std::vector<int> doff = m >= C ? dia_offsets(this, R, C, nb, nnzb) : std::vector<int>();
  if (!doff.empty()) {   // uniform-offset sliced layout
    const int ns = (nb + kSellRows - 1) / kSellRows, no = (int)doff.size();
    std::vector<int> hoff(ns + 1), hwid(ns, no);
    for (int sl = 0; sl <= ns; ++sl) hoff[sl] = sl * kSellRows * no;
    B->sell = B->dia = true;
    B->noff = no;
    B->doff.upload(doff.data(), no);
    B->nslices = ns;
    B->stot = (size_t)ns * kSellRows * no;
    B->soff.upload(hoff.data(), ns + 1);
    B->swid.upload(hwid.data(), ns);
    B->bv.alloc(B->stot * R * C);
    DECO_HIP(hipMemsetAsync(B->bv.p, 0, sizeof(float) * B->stot * R * C, stream()));
    hipLaunchKernelGGL(k_dia_fill<float>, dim3(grid_for(nb, 256)), dim3(256), 0, stream(), nb, R, C, no,
                       (const int*)B->doff.p, r.p, c.p, v.p, (const int*)B->soff.p, B->stot, B->bv.p);
    DECO_CHECK_LAUNCH();
    sync();   // the host arrays go out of scope here
    b32 = B;
    return B;
  }
  if (g_sell_padding >= 0.0 && nb > 0 && m >= C) {   // sliced layout when its padding is small
    const int ns = (nb + kSellRows - 1) / kSellRows;
    DBuf<int> ent(ns + 1);
    B->swid.alloc(ns);
    hipLaunchKernelGGL(k_sell_width, dim3(grid_for(ns, 256)), dim3(256), 0, stream(), nb, ns, cnt.p, ent.p, B->swid.p);
    DECO_CHECK_LAUNCH();
    B->soff.alloc(ns + 1);
    const int64_t T = scan_to_ptr(ent.p, B->soff.p, ns);
    if (T <= INT_MAX && (double)T <= (double)nnzb * (1.0 + g_sell_padding)) {
      B->sell = true;
      B->nslices = ns;
      B->stot = (size_t)T;
      B->bc.alloc(std::max<int64_t>(T, 1));
      B->bv.alloc(std::max<int64_t>(T, 1) * R * C);
      DECO_HIP(hipMemsetAsync(B->bc.p, 0, sizeof(int) * std::max<int64_t>(T, 1), stream()));
      DECO_HIP(hipMemsetAsync(B->bv.p, 0, sizeof(float) * std::max<int64_t>(T, 1) * R * C, stream()));
      hipLaunchKernelGGL(k_sell_fill<float>, dim3(grid_for(nb, 256)), dim3(256), 0, stream(), nb, R, C, r.p, c.p,
                         v.p, (const int*)B->soff.p, B->stot, B->bc.p, B->bv.p);
      DECO_CHECK_LAUNCH();
      if (g_sell_offsets16 && nb <= m / C) {   // an offset 0 then addresses a block column of x
        const size_t ne = std::max<size_t>(B->stot, 1);
        B->bcd.alloc(ne);
        DECO_HIP(hipMemsetAsync(B->bcd.p, 0, sizeof(short) * ne, stream()));
        DBuf<int> ovf(1);
        ovf.zero();
        hipLaunchKernelGGL(k_sell_delta, dim3(grid_for(nb, 256)), dim3(256), 0, stream(), nb, (const int*)B->br.p,
                           (const int*)B->soff.p, (const int*)B->bc.p, B->bcd.p, ovf.p);
        DECO_CHECK_LAUNCH();
        int o = 0;
        copy_d2h(&o, ovf.p, 1);
        B->d16 = (o == 0);
        if (!B->d16) B->bcd.release();
      }
      b32 = B;
      return B;
    }
    B->swid.release();
    B->soff.release();
  }
  B->bc.alloc(std::max<int64_t>(nnzb, 1));
  B->bv.alloc(std::max<int64_t>(nnzb, 1) * R * C);
  DECO_HIP(hipMemsetAsync(B->bv.p, 0, sizeof(float) * std::max<int64_t>(nnzb, 1) * R * C, stream()));
  if (nb > 0) {
    hipLaunchKernelGGL(k_bsr_fill<float>, dim3(grid_for(nb, 256)), dim3(256), 0, stream(), nb, R, C, r.p, c.p, v.p,
                       B->br.p, B->bc.p, B->bv.p, bsr_planar_rt(4, R, C) ? (size_t)nnzb * C : (size_t)0);
    DECO_CHECK_LAUNCH();
  }
  B->nchunks = (int)std::max<int64_t>(1, (nnzb + kBsrCap - 1) / kBsrCap);
  B->rb.alloc(B->nchunks + 1);
  hipLaunchKernelGGL(k_chunks_cap, dim3(grid_for(B->nchunks + 1, 256)), dim3(256), 0, stream(), nb, B->br.p,
                     B->nchunks, kBsrCap, B->rb.p);
  DECO_CHECK_LAUNCH();
  b32 = B;
  return B;
}

namespace deco {
// streaming read of n2 16-byte vectors (eight non-temporal loads in flight per lane), one
// partial sum per workgroup: the read bandwidth the products' read-dominated traffic is compared
// with (64 workgroups per CU and eight loads per lane measured fastest on MI355X)
__global__ __launch_bounds__(256) void k_read_stream(const f64x2* __restrict__ p, size_t n2, double* __restrict__ part) {
  constexpr int K = 8;
  __shared__ double red[4];
  double s = 0.0;
  const size_t stride = (size_t)gridDim.x * 256;
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  for (; i + (K - 1) * stride < n2; i += K * stride) {
    f64x2 t[K];
#pragma unroll
    for (int k = 0; k < K; ++k) t[k] = __builtin_nontemporal_load(p + i + k * stride);
#pragma unroll
    for (int k = 0; k < K; ++k) s += t[k].x + t[k].y;
  }
  for (; i < n2; i += stride) {
    const f64x2 a = __builtin_nontemporal_load(p + i);
    s += a.x + a.y;
  }
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}
}  // namespace deco

bool deco_csr::sell_scalar() const {
  using namespace deco;
  if (sell_ok < 0) {
    sell_ok = 0;
    if (g_sell_padding >= 0.0 && n > 0 && m > 0 && nnz > 0) {
      const int ns = (n + kSellRows - 1) / kSellRows;
      DBuf<int> len(n), ent(ns + 1), wid(ns), off(ns + 1);
      hipLaunchKernelGGL(k_row_len, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, r.p, len.p);
      DECO_CHECK_LAUNCH();
      hipLaunchKernelGGL(k_sell_width, dim3(grid_for(ns, 256)), dim3(256), 0, stream(), n, ns, (const int*)len.p, ent.p,
                         wid.p);
      DECO_CHECK_LAUNCH();
      const int64_t T = scan_to_ptr(ent.p, off.p, ns);
      sell_ok = (T <= INT_MAX && (double)T <= (double)nnz * (1.0 + g_sell_padding)) ? 1 : 0;
      if (!sell_ok && !dia_offsets(this, 1, 1, n, nnz).empty()) sell_ok = 1;
    }
  }
  return sell_ok == 1;
}

const deco_bsr* deco_csr::block64(int R, int C) const {
  using namespace deco;
  const deco_bsr* B0 = block32(R, C);
  if (!B0 || B0->has64) return B0;
  deco_bsr* B = b32;
  if (B->dia) {
    const size_t nv = B->stot * R * C;
    B->bv64.alloc(nv);
    DECO_HIP(hipMemsetAsync(B->bv64.p, 0, sizeof(double) * nv, stream()));
    hipLaunchKernelGGL(k_dia_fill<double>, dim3(grid_for(B->nb, 256)), dim3(256), 0, stream(), B->nb, R, C, B->noff,
                       (const int*)B->doff.p, r.p, c.p, v.p, (const int*)B->soff.p, B->stot, B->bv64.p);
    DECO_CHECK_LAUNCH();
    B->has64 = true;
    return B;
  }
  if (B->sell) {
    const size_t nv = std::max<size_t>(B->stot, 1) * R * C;
    B->bv64.alloc(nv);
    DECO_HIP(hipMemsetAsync(B->bv64.p, 0, sizeof(double) * nv, stream()));
    // rewrites the same block columns, adds the FP64 values
    hipLaunchKernelGGL(k_sell_fill<double>, dim3(grid_for(B->nb, 256)), dim3(256), 0, stream(), B->nb, R, C, r.p,
                       c.p, v.p, (const int*)B->soff.p, B->stot, B->bc.p, B->bv64.p);
    DECO_CHECK_LAUNCH();
    B->has64 = true;
    return B;
  }
  const size_t nv = (size_t)std::max(B->nnzb, 1) * R * C;
  B->bv64.alloc(nv);
  DECO_HIP(hipMemsetAsync(B->bv64.p, 0, sizeof(double) * nv, stream()));
  if (B->nb > 0) {
    // rewrites the same block columns, adds the FP64 values
    hipLaunchKernelGGL(k_bsr_fill<double>, dim3(grid_for(B->nb, 256)), dim3(256), 0, stream(), B->nb, R, C, r.p,
                       c.p, v.p, B->br.p, B->bc.p, B->bv64.p, bsr_planar_rt(8, R, C) ? (size_t)B->nnzb * C : (size_t)0);
    DECO_CHECK_LAUNCH();
  }
  B->has64 = true;
  return B;
}

const int2* deco_csr::row_chunks() const {
  if (!rb_valid) {
    // chunks of at most kStreamTile entries: the cap drops below kStreamCap when the longest row
    // would let a chunk overflow the tile (rows up to half the tile); the window a chunk's range
    // can span is cap + longest - 1
    int cap = deco::kStreamCap, win = deco::kStreamTile;
    chunks_fit = false;
    if (n > 0 && nnz > 0) {
      deco::DBuf<int> mx(1);
      mx.zero();
      hipLaunchKernelGGL(deco::k_max_row, dim3(deco::grid_for(n, 256)), dim3(256), 0, deco::stream(), n, r.p, mx.p);
      DECO_CHECK_LAUNCH();
      int longest = 0;
      deco::copy_d2h(&longest, mx.p, 1);
      if (longest <= deco::kStreamTile / 2) {
        if (longest > deco::kStreamTile - deco::kStreamCap + 1) cap = deco::kStreamTile - longest + 1;
        win = cap + std::max(longest, 1) - 1;
        chunks_fit = true;   // a chunk's last row starts below (k+1)*cap, so it ends within the window
      }
    }
    chunk_cap = cap;
    chunk_win = win;
    nchunks = (int)std::max<int64_t>(1, ((int64_t)nnz + cap - 1) / cap);
    rb.ensure(nchunks + 1);
    hipLaunchKernelGGL(deco::k_chunks_cap, dim3(deco::grid_for(nchunks + 1, 256)), dim3(256), 0, deco::stream(), n,
                       r.p, nchunks, cap, rb.p);
    DECO_CHECK_LAUNCH();
    rb_valid = true;
  }
  return rb.p;
}

namespace deco {
// per stream chunk: its smallest column (the base of its column offsets), and ovf set when its
// columns span more than maxspan
__global__ __launch_bounds__(256) void k_chunk_span(const int2* __restrict__ desc, const int* __restrict__ c,
                                                    int* __restrict__ cbase, int* __restrict__ ovf, int maxspan) {
  __shared__ int slo[4], shi[4];
  const int j0 = desc[blockIdx.x].y, j1 = desc[blockIdx.x + 1].y;
  int lo = INT_MAX, hi = INT_MIN;
  for (int j = j0 + (int)threadIdx.x; j < j1; j += 256) {
    lo = min(lo, c[j]);
    hi = max(hi, c[j]);
  }
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off));
    hi = max(hi, __shfl_xor(hi, off));
  }
  if ((threadIdx.x & 63) == 0) {
    slo[threadIdx.x >> 6] = lo;
    shi[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      lo = min(lo, slo[w]);
      hi = max(hi, shi[w]);
    }
    lo = min(lo, slo[0]);
    hi = max(hi, shi[0]);
    cbase[blockIdx.x] = j1 > j0 ? lo : 0;
    if (j1 > j0 && (int64_t)hi - lo > maxspan) atomicOr(ovf, 1);
  }
}

__global__ __launch_bounds__(256) void k_chunk_c16(const int2* __restrict__ desc, const int* __restrict__ c,
                                                   const int* __restrict__ cbase, unsigned short* __restrict__ c16) {
  const int j0 = desc[blockIdx.x].y, j1 = desc[blockIdx.x + 1].y, b = cbase[blockIdx.x];
  for (int j = j0 + (int)threadIdx.x; j < j1; j += 256) c16[j] = (unsigned short)(c[j] - b);
}
}  // namespace deco

namespace deco {
static bool g_stream_pieces = true;   // the row-piece form for the stream products where it fits
// Largest matrix (entries) that takes the row-piece form. Its products have a shorter dependency
// chain than the CSR-stream kernel's (no LDS staging or barrier, no row-range load before the
// sums), which pays where a product is latency-bound: measured per AMG operator of the
// 168x100x120 hierarchy inside the solves, 1-5 us faster per launch up to 2.7M entries, equal or
// slower above (level 1 and 2 operators, the level-0 interpolation).
constexpr int kPieceMaxNnz = 2800000;

// per row: its first and last column into the smallest / largest column of its slice (rows sorted)
__global__ void k_piece_span(int n, const int* __restrict__ r, const int* __restrict__ c,
                             const int* __restrict__ place, int* __restrict__ smin, int* __restrict__ smax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || r[i] == r[i + 1]) return;
  const int sl = place[i] >> 6;
  atomicMin(&smin[sl], c[r[i]]);
  atomicMax(&smax[sl], c[r[i + 1] - 1]);
}

// per row: its pieces at lanes place & 63 .. of slice place >> 6 (structure: the lanes' row
// offsets and the slots' columns too, padding slots at the row's last column, or the slice base)
template <class VT>
__global__ void k_piece_fill(int n, int W, const int* __restrict__ r, const int* __restrict__ c,
                             const double* __restrict__ v, const int* __restrict__ place,
                             const int* __restrict__ srow, const int* __restrict__ cbase,
                             unsigned char* __restrict__ pdel, unsigned short* __restrict__ c16,
                             int* __restrict__ c32, VT* __restrict__ vals, bool structure) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = r[i], e = r[i + 1];
  const int np = max(1, (e - b + W - 1) / W);
  const int sl = place[i] >> 6, l0 = place[i] & 63;
  const int last = e > b ? c[e - 1] : (cbase ? cbase[sl] : 0);
  for (int p = 0; p < np; ++p) {
    if (structure) pdel[(size_t)sl * 64 + l0 + p] = (unsigned char)(i - srow[sl]);
    for (int t = 0; t < W; ++t) {
      const int j = b + p * W + t;
      const size_t slot = ((size_t)sl * W + t) * 64 + l0 + p;
      if (structure) {
        const int cj = j < e ? c[j] : last;
        if (cbase) c16[slot] = (unsigned short)(cj - cbase[sl]);
        else c32[slot] = cj;
      }
      vals[slot] = j < e ? (VT)v[j] : VT(0);
    }
  }
}
}  // namespace deco

const deco_pieces* deco_csr::pieces(bool fp64) const {
  using namespace deco;
  if (pc_ok < 0) {
    pc_ok = 0;
    if (g_stream_pieces && n > 0 && nnz > 0 && nnz <= kPieceMaxNnz) {
      std::vector<int> hr(n + 1);
      copy_d2h(hr.data(), r.p, n + 1);
      int maxlen = 0;
      for (int i = 0; i < n; ++i) maxlen = std::max(maxlen, hr[i + 1] - hr[i]);
      // the piece width with the fewest bytes (values and 2-byte columns per slot, a row byte per
      // lane) among those that keep every row within one slice
      int bestW = 0;
      int64_t best_ns = 0;
      double best_cost = 0.0;
      for (int W : {1, 2, 4, 8}) {
        if ((maxlen + W - 1) / W > 64) continue;
        int64_t ns = 0;
        int lane = 64;
        for (int i = 0; i < n; ++i) {
          const int np = std::max(1, (hr[i + 1] - hr[i] + W - 1) / W);
          if (lane + np > 64) { ++ns; lane = 0; }
          lane += np;
        }
        const double cost = (double)ns * 64 * W * 6.0 + (double)ns * 64;
        if (!bestW || cost < best_cost) { bestW = W; best_ns = ns; best_cost = cost; }
      }
      if (bestW && (double)best_ns * 64 * bestW <= 1.3 * nnz && best_ns * 64 * bestW <= INT_MAX) {
        auto* P = new deco_pieces();
        const int W = bestW, ns = (int)best_ns;
        P->W = W;
        P->nslices = ns;
        P->nslot = (size_t)ns * 64 * W;
        std::vector<int> place(n), srow(ns);
        int sl = -1, lane = 64;
        for (int i = 0; i < n; ++i) {
          const int np = std::max(1, (hr[i + 1] - hr[i] + W - 1) / W);
          if (lane + np > 64) { ++sl; lane = 0; srow[sl] = i; }
          place[i] = sl * 64 + lane;
          lane += np;
        }
        DBuf<int> dplace;
        dplace.upload(place.data(), n);
        P->srow.upload(srow.data(), ns);
        std::vector<int> smin(ns, INT_MAX), smax(ns, INT_MIN);
        DBuf<int> dmin, dmax;
        dmin.upload(smin.data(), ns);
        dmax.upload(smax.data(), ns);
        hipLaunchKernelGGL(k_piece_span, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, r.p, c.p,
                           (const int*)dplace.p, dmin.p, dmax.p);
        DECO_CHECK_LAUNCH();
        copy_d2h(smin.data(), dmin.p, ns);
        copy_d2h(smax.data(), dmax.p, ns);
        bool fits = g_sell_offsets16;
        for (int q = 0; q < ns && fits; ++q) {
          if (smin[q] == INT_MAX) smin[q] = 0;
          else if ((int64_t)smax[q] - smin[q] > 65535) fits = false;
        }
        for (int q = 0; q < ns; ++q)
          if (smin[q] == INT_MAX) smin[q] = 0;
        P->d16 = fits;
        P->pdel.alloc((size_t)ns * 64);
        DECO_HIP(hipMemsetAsync(P->pdel.p, 0xFF, (size_t)ns * 64, stream()));
        // padding lanes keep offset / column 0 (a valid column) and value 0
        if (fits) {
          P->cbase.upload(smin.data(), ns);
          P->c16.alloc(P->nslot);
          DECO_HIP(hipMemsetAsync(P->c16.p, 0, sizeof(unsigned short) * P->nslot, stream()));
        } else {
          P->c32.alloc(P->nslot);
          DECO_HIP(hipMemsetAsync(P->c32.p, 0, sizeof(int) * P->nslot, stream()));
        }
        P->vf.alloc(P->nslot);
        DECO_HIP(hipMemsetAsync(P->vf.p, 0, sizeof(float) * P->nslot, stream()));
        hipLaunchKernelGGL(k_piece_fill<float>, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, W, r.p, c.p,
                           v.p, (const int*)dplace.p, (const int*)P->srow.p,
                           fits ? (const int*)P->cbase.p : (const int*)nullptr, P->pdel.p, P->c16.p, P->c32.p,
                           P->vf.p, true);
        DECO_CHECK_LAUNCH();
        sync();   // the host arrays go out of scope here
        pc = P;
        pc_ok = 1;
      }
    }
  }
  if (pc_ok != 1) return nullptr;
  if (fp64 && !pc->has64) {
    // the FP64 values in the same slots: the placement is recomputed from the row pointers
    std::vector<int> hr(n + 1);
    copy_d2h(hr.data(), r.p, n + 1);
    const int W = pc->W;
    std::vector<int> place(n);
    int sl = -1, lane = 64;
    for (int i = 0; i < n; ++i) {
      const int np = std::max(1, (hr[i + 1] - hr[i] + W - 1) / W);
      if (lane + np > 64) { ++sl; lane = 0; }
      place[i] = sl * 64 + lane;
      lane += np;
    }
    DBuf<int> dplace;
    dplace.upload(place.data(), n);
    pc->v64.alloc(pc->nslot);
    DECO_HIP(hipMemsetAsync(pc->v64.p, 0, sizeof(double) * pc->nslot, stream()));
    hipLaunchKernelGGL(k_piece_fill<double>, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, W, r.p, c.p, v.p,
                       (const int*)dplace.p, (const int*)pc->srow.p,
                       pc->d16 ? (const int*)pc->cbase.p : (const int*)nullptr, pc->pdel.p, pc->c16.p, pc->c32.p,
                       pc->v64.p, false);
    DECO_CHECK_LAUNCH();
    sync();
    pc->has64 = true;
  }
  return pc;
}

bool deco_csr::cols16() const {
  using namespace deco;
  if (c16_ok < 0) {
    c16_ok = 0;
    if (g_sell_offsets16 && n > 0 && nnz > 0) {
      const int2* desc = row_chunks();
      cbase.ensure(nchunks);
      DBuf<int> ovf(1);
      ovf.zero();
      hipLaunchKernelGGL(k_chunk_span, dim3(nchunks), dim3(256), 0, stream(), desc, (const int*)c.p, cbase.p, ovf.p,
                         65535);
      DECO_CHECK_LAUNCH();
      int o = 0;
      copy_d2h(&o, ovf.p, 1);
      if (o == 0) {
        c16.ensure(nnz);
        hipLaunchKernelGGL(k_chunk_c16, dim3(nchunks), dim3(256), 0, stream(), desc, (const int*)c.p,
                           (const int*)cbase.p, c16.p);
        DECO_CHECK_LAUNCH();
        c16_ok = 1;
      } else {
        cbase.release();
      }
    }
  }
  return c16_ok == 1;
}

namespace deco {
// the column-sorted chunk form for the stream products: 0 never, 1 by the rule of
// deco_csr::sorted (default), 2 wherever it applies
static int g_stream_sorted = 1;
// FP32 products of matrices with 16-bit chunk columns take the sorted form when the row-order
// gathers of a wavefront touch at least this many 64-byte segments of x on average (measured on
// the 168x100x120 AMG operators: the level-0 interpolation, 7.2, was 0.7 us slower sorted; the
// level-0 restriction, 10.8, 1.4 us faster; the level matrices, 19-33, 1-5 us faster)
constexpr double kSortMinSegments = 9.0;

// per 64-entry group of the column array (the stream kernel's wavefront loads): the number of
// distinct 64-byte segments of an FP32 x its gathers touch, summed into *total
__global__ __launch_bounds__(256) void k_gather_segments(int nnz, const int* __restrict__ c,
                                                         unsigned long long* __restrict__ total) {
  const int j = blockIdx.x * 256 + (int)threadIdx.x;
  const int lane = threadIdx.x & 63;
  const int sg = j < nnz ? (c[j] >> 4) : -1;
  bool first = j < nnz;
  for (int k = 0; k < 64; ++k) {
    const int o = __shfl(sg, k);
    if (k < lane && o == sg) first = false;
  }
  const unsigned long long b = __ballot(first);
  if (lane == 0) atomicAdd(total, (unsigned long long)__popcll(b));
}

// Chunk ch's own entries [j0, j1) sorted by column in LDS (a bitonic sort of 4096 keys, column
// offset << 12 | slot, ties in row order), written as column offset | slot << 20. Every chunk fits
// its window (at most kStreamTile < 4096 entries from its start), so its slots stay below 4096.
__global__ __launch_bounds__(256) void k_csort_build(const int2* __restrict__ desc, const int* __restrict__ c,
                                                     int cap, const int* __restrict__ cbase,
                                                     unsigned int* __restrict__ p) {
  constexpr int N = 4096;
  __shared__ unsigned int key[N];
  const int ch = blockIdx.x;
  const int j0 = desc[ch].y, j1 = desc[ch + 1].y, w0 = ch * cap, cb = cbase[ch];
  const int cnt = j1 - j0;
  for (int i = threadIdx.x; i < N; i += 256)
    key[i] = i < cnt ? ((unsigned int)(c[j0 + i] - cb) << 12) | (unsigned int)(j0 + i - w0) : 0xFFFFFFFFu;
  __syncthreads();
  for (int k = 2; k <= N; k <<= 1) {
    for (int h = k >> 1; h > 0; h >>= 1) {
      for (int i = threadIdx.x; i < N; i += 256) {
        const int o = i ^ h;
        if (o > i) {
          const unsigned int a = key[i], b = key[o];
          if (((i & k) == 0) ? (a > b) : (a < b)) {
            key[i] = b;
            key[o] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int i = threadIdx.x; i < cnt; i += 256) {
    const unsigned int q = key[i];
    p[j0 + i] = (q >> 12) | ((q & 0xFFFu) << kSortColBits);
  }
}

// values in the sorted order: dst[j] = src[window start + slot of entry j]
template <class VT>
__global__ __launch_bounds__(256) void k_csort_vals(const int2* __restrict__ desc, int cap,
                                                    const unsigned int* __restrict__ p, const VT* __restrict__ src,
                                                    VT* __restrict__ dst) {
  const int ch = blockIdx.x;
  const int j0 = desc[ch].y, j1 = desc[ch + 1].y, w0 = ch * cap;
  for (int j = j0 + (int)threadIdx.x; j < j1; j += 256) dst[j] = src[w0 + (int)(p[j] >> kSortColBits)];
}
}  // namespace deco

double deco_csr::gather_segments() const {
  using namespace deco;
  if (gather_seg < 0.0) {
    gather_seg = 0.0;
    if (nnz > 0) {
      DBuf<unsigned long long> tot(1);
      tot.zero();
      hipLaunchKernelGGL(k_gather_segments, dim3(grid_for(nnz, 256)), dim3(256), 0, stream(), nnz, (const int*)c.p,
                         tot.p);
      DECO_CHECK_LAUNCH();
      unsigned long long t = 0;
      copy_d2h(&t, tot.p, 1);
      gather_seg = (double)t / (double)((nnz + 63) / 64);
    }
  }
  return gather_seg;
}

const deco_csort* deco_csr::sorted(bool fp64) const {
  using namespace deco;
  static_assert(kStreamTile < 4096, "slots of the sorted form are 12 bits");
  if (g_stream_sorted == 0 || n <= 0 || nnz <= 0) return nullptr;
  // the rule: in place of 32-bit columns always (the same bytes: measured on A of the 168x100x120
  // system, FP32 108 -> 80 us, FP64 142 -> 141 us); in place of 16-bit columns (two bytes more per
  // entry) only for FP32 values whose row-order gathers scatter
  if (g_stream_sorted == 1 && cols16() && (fp64 || gather_segments() < kSortMinSegments)) return nullptr;
  if (cs_ok < 0) {
    cs_ok = 0;
    const int2* desc = row_chunks();
    if (chunks_fit) {
      auto* S = new deco_csort();
      S->cbase.alloc(nchunks);
      DBuf<int> ovf(1);
      ovf.zero();
      hipLaunchKernelGGL(k_chunk_span, dim3(nchunks), dim3(256), 0, stream(), desc, (const int*)c.p, S->cbase.p,
                         ovf.p, (int)kSortColMask);
      DECO_CHECK_LAUNCH();
      int o = 0;
      copy_d2h(&o, ovf.p, 1);
      if (o == 0) {
        S->p.alloc(nnz);
        hipLaunchKernelGGL(k_csort_build, dim3(nchunks), dim3(256), 0, stream(), desc, (const int*)c.p, chunk_cap,
                           (const int*)S->cbase.p, S->p.p);
        DECO_CHECK_LAUNCH();
        cs = S;
        cs_ok = 1;
      } else {
        delete S;
      }
    }
  }
  if (cs_ok != 1) return nullptr;
  const int2* desc = row_chunks();
  if (fp64 && !cs->has64) {
    cs->v64.alloc(nnz);
    hipLaunchKernelGGL(k_csort_vals<double>, dim3(nchunks), dim3(256), 0, stream(), desc, chunk_cap,
                       (const unsigned int*)cs->p.p, (const double*)v.p, cs->v64.p);
    DECO_CHECK_LAUNCH();
    cs->has64 = true;
  }
  if (!fp64 && !cs->has32) {
    // the FP32 values are the f32 shadow's, so the products equal those on the shadow bit for bit
    const float* f = const_cast<deco_csr*>(this)->values_f32();
    cs->vf.alloc(nnz);
    hipLaunchKernelGGL(k_csort_vals<float>, dim3(nchunks), dim3(256), 0, stream(), desc, chunk_cap,
                       (const unsigned int*)cs->p.p, f, cs->vf.p);
    DECO_CHECK_LAUNCH();
    cs->has32 = true;
  }
  return cs;
}

const float* deco_csr::values_f32() {
  if (!vf_valid) {
    vf.ensure(nnz);
    deco::to_f32(v.p, vf.p, nnz);
    vf_valid = true;
  }
  return vf.p;
}

using namespace deco;

// ====================================================================== C ABI ==========
extern "C" {

int deco_init(int device) { return guard([&] { init_device(device); }); }

const char* deco_last_error(void) { return t_last_error.c_str(); }

int deco_sync(void) { return guard([&] { sync(); }); }

void* deco_stream(void) {
  try { return (void*)stream(); } catch (...) { return nullptr; }
}

int deco_malloc(size_t bytes, void** out) {
  return guard([&] {
    if (!out) fail(DECO_ERR_ARG, "null output");
    ensure_init();
    *out = nullptr;
    if (bytes) DECO_HIP(hipMalloc(out, bytes));
  });
}

int deco_free(void* p) {
  return guard([&] { if (p) DECO_HIP(hipFree(p)); });
}

int deco_memcpy_h2d(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    need(dst, (int64_t)bytes, "destination");
    need(src, (int64_t)bytes, "source");
    if (bytes) DECO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream()));
    sync();
  });
}

int deco_memcpy_d2h(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    need(dst, (int64_t)bytes, "destination");
    need(src, (int64_t)bytes, "source");
    if (bytes) DECO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, stream()));
    sync();
  });
}

int deco_memcpy_d2d(void* dst, const void* src, size_t bytes) {
  return guard([&] {
    need(dst, (int64_t)bytes, "destination");
    need(src, (int64_t)bytes, "source");
    if (bytes) DECO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, stream()));
    sync();
  });
}

int deco_memset(void* dst, int value, size_t bytes) {
  return guard([&] {
    need(dst, (int64_t)bytes, "destination");
    if (bytes) DECO_HIP(hipMemsetAsync(dst, value, bytes, stream()));
    sync();
  });
}

static void check_csr_args(int n, int m, int nnz) {
  if (n < 0 || m < 0 || nnz < 0) fail(DECO_ERR_ARG, "negative CSR size n=%d m=%d nnz=%d", n, m, nnz);
}

namespace deco {
// structure check of an uploaded CSR: err |= 1 for a row pointer that decreases, leaves
// [0, nnz] or does not span it, |= 2 for a column outside [0, m) (the loops stay inside [0, nnz)
// either way)
__global__ void k_csr_check(int n, int m, int nnz, const int* __restrict__ r, const int* __restrict__ c,
                            int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (i == 0 && (r[0] != 0 || r[n] != nnz)) atomicOr(err, 1);
  const int b = r[i], e = r[i + 1];
  if (b < 0 || e < b || e > nnz) {
    atomicOr(err, 1);
    return;
  }
  for (int k = b; k < e; ++k)
    if (c[k] < 0 || c[k] >= m) {
      atomicOr(err, 2);
      return;
    }
}
}  // namespace deco

// rejects a structure the kernels cannot address safely (they index x by the columns and walk
// the rows by the pointers)
static void check_csr_structure(const deco_csr* A) {
  if (A->n == 0) return;
  DBuf<int> err(1);
  err.zero();
  hipLaunchKernelGGL(k_csr_check, dim3(grid_for(A->n, 256)), dim3(256), 0, stream(), A->n, A->m, A->nnz, A->r.p,
                     A->c.p, err.p);
  DECO_CHECK_LAUNCH();
  int e = 0;
  copy_d2h(&e, err.p, 1);
  if (e & 1) fail(DECO_ERR_ARG, "row pointer decreases or leaves [0, nnz]");
  if (e & 2) fail(DECO_ERR_ARG, "column index outside [0, %d)", A->m);
}

int deco_csr_create(int n, int m, int nnz, const int* r, const int* c, const double* v, deco_csr** out) {
  return guard([&] {
    check_csr_args(n, m, nnz);
    if (!out || !r || (nnz && (!c || !v))) fail(DECO_ERR_ARG, "null CSR array");
    if (r[0] != 0 || r[n] != nnz) fail(DECO_ERR_ARG, "row pointer does not span [0, nnz]");
    auto* A = csr_new(n, m);
    A->nnz = nnz;
    A->r.upload(r, n + 1);
    A->c.upload(c, nnz);
    A->v.upload(v, nnz);
    try {
      check_csr_structure(A);
    } catch (...) {
      delete A;
      throw;
    }
    sync();
    *out = A;
  });
}

int deco_csr_create_dev(int n, int m, int nnz, const int* r, const int* c, const double* v, deco_csr** out) {
  return guard([&] {
    check_csr_args(n, m, nnz);
    if (!out) fail(DECO_ERR_ARG, "null output");
    need(r, (int64_t)n + 1, "r");
    need(c, nnz, "c");
    need(v, nnz, "v");
    auto* A = csr_new(n, m);
    A->nnz = nnz;
    A->r.alloc(n + 1);
    A->c.alloc(nnz);
    A->v.alloc(nnz);
    copy_d2d(A->r.p, r, n + 1);
    copy_d2d(A->c.p, c, nnz);
    copy_d2d(A->v.p, v, nnz);
    try {
      check_csr_structure(A);
    } catch (...) {
      delete A;
      throw;
    }
    sync();
    *out = A;
  });
}

int deco_csr_destroy(deco_csr* A) {
  return guard([&] { delete A; });
}

int deco_csr_info(const deco_csr* A, int* n, int* m, int* nnz) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    if (n) *n = A->n;
    if (m) *m = A->m;
    if (nnz) *nnz = A->nnz;
  });
}

int deco_csr_download(const deco_csr* A, int* r, int* c, double* v) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    if (r) A->r.download(r, A->n + 1);
    if (c) A->c.download(c, A->nnz);
    if (v) A->v.download(v, A->nnz);
  });
}

int deco_csr_device_ptrs(const deco_csr* A, int** r, int** c, double** v) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    if (r) *r = A->r.p;
    if (c) *c = A->c.p;
    if (v) *v = A->v.p;
  });
}

int deco_csr_update_values(deco_csr* A, const double* v) {
  return guard([&] {
    if (!A || (A->nnz && !v)) fail(DECO_ERR_ARG, "null argument");
    A->v.upload(v, A->nnz);
    A->invalidate();
    if (A->At) { delete A->At; A->At = nullptr; }
    sync();
  });
}

int deco_csr_copy(const deco_csr* A, deco_csr** out) {
  return guard([&] {
    if (!A || !out) fail(DECO_ERR_ARG, "null argument");
    *out = csr_clone(A);
    sync();
  });
}

int deco_csr_filter(deco_csr* A, double tol) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    csr_filter(A, tol);
    sync();
  });
}

int deco_csr_filter_proper(deco_csr* A, double tol) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    csr_filter_proper(A, tol);
    sync();
  });
}

int deco_csr_local_matrix(const deco_csr* A, int bs, int p, const double* w, deco_csr** out) {
  return guard([&] {
    if (!A || !out) fail(DECO_ERR_ARG, "null argument");
    *out = local_matrix(A, bs, p, w);
    sync();
  });
}

int deco_block_reorder(deco_csr* A, int bs, const char* row_order, const char* col_order, double* b_dev) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    block_reorder(A, bs, row_order, col_order, b_dev);
  });
}

int deco_csr_transpose(const deco_csr* A, deco_csr** out) {
  return guard([&] {
    if (!A || !out) fail(DECO_ERR_ARG, "null argument");
    *out = csr_transpose(A);
    sync();
  });
}

int deco_spmv(const deco_csr* A, const double* x, double* y, int bs) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    if (bs < 1 || A->n % bs != 0) fail(DECO_ERR_ARG, "n=%d not divisible by bs=%d", A->n, bs);
    need(x, A->m, "x");
    need(y, A->n / bs, "y");
    if (bs == 1) spmv(A, x, y);
    else
      launch_spmv<EPI_SET, double, double, double, double>(csr_group(A), A->n / bs, bs, A->r.p, A->c.p, A->v.p, x, y,
                                                           nullptr, 0, 0);
    sync();
  });
}

int deco_spmv_residual(const deco_csr* A, const double* x, const double* z, double* y) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    need(x, A->m, "x");
    need(z, A->n, "z");
    need(y, A->n, "y");
    spmv_res(A, x, z, y);
    sync();
  });
}

int deco_spmv_transpose(deco_csr* A, const double* x, double* y) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    need(x, A->n, "x");
    need(y, A->m, "y");
    if (!A->At) A->At = csr_transpose(A);
    spmv(A->At, x, y);
    sync();
  });
}

int deco_spmv_f32(deco_csr* A, const float* x, float* y) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    need(x, A->m, "x");
    need(y, A->n, "y");
    const float* vf = A->values_f32();
    spmv_stream<EPI_SET, float, float, float, float>(A, vf, x, y, (const float*)nullptr);
    sync();
  });
}

int deco_spmv_block_f32(deco_csr* A, int R, int C, const float* x, float* y) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    need(x, A->m, "x");
    need(y, A->n, "y");
    const deco_bsr* B = A->block32(R, C);
    if (!B) fail(DECO_ERR_ARG, "%d x %d blocks do not tile a %d x %d matrix", R, C, A->n, A->m);
    bsr_stream_f(B, x, EpiStd<EPI_SET, float, float>{y, nullptr});
    sync();
  });
}

int deco_block_info(deco_csr* A, int R, int C, int* sliced, int* nnzb, long long* entries) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    const deco_bsr* B = A->block32(R, C);
    if (!B) fail(DECO_ERR_ARG, "%d x %d blocks do not tile a %d x %d matrix", R, C, A->n, A->m);
    if (sliced) *sliced = B->dia ? 3 : B->sell ? (B->d16 ? 2 : 1) : 0;
    if (nnzb) *nnzb = B->nnzb;
    if (entries) *entries = B->sell ? (long long)B->stot : (long long)B->nnzb;
  });
}

int deco_set_sell_padding(double max_fraction) {
  return guard([&] {
    if (!(max_fraction == max_fraction)) fail(DECO_ERR_ARG, "padding fraction is NaN");
    g_sell_padding = max_fraction;
  });
}

int deco_set_sell_offsets16(int enable) {
  return guard([&] { g_sell_offsets16 = enable != 0; });
}

int deco_set_sell_dia(int enable) {
  return guard([&] { g_sell_dia = enable != 0; });
}

int deco_set_sell_pairs(int mode) {
  return guard([&] {
    if (mode < 0 || mode > 2) fail(DECO_ERR_ARG, "sell pair mode %d (0, 1 or 2)", mode);
    g_sell_pairs = mode;
  });
}

int deco_set_stream_pieces(int enable) {
  return guard([&] { g_stream_pieces = enable != 0; });
}

int deco_set_stream_sorted(int mode) {
  return guard([&] {
    if (mode < 0 || mode > 2) fail(DECO_ERR_ARG, "stream sorted mode %d (0, 1 or 2)", mode);
    g_stream_sorted = mode;
  });
}

int deco_read_bandwidth(size_t bytes, int iters, double* gbps) {
  return guard([&] {
    if (!gbps || iters <= 0 || bytes < ((size_t)1 << 20)) fail(DECO_ERR_ARG, "bad argument");
    const size_t n2 = bytes / 16;
    int dev = 0, cus = 0;
    DECO_HIP(hipGetDevice(&dev));
    DECO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int G = std::max(cus, 1) * 64;
    DBuf<double> buf(2 * n2), part(G);
    buf.zero();
    hipLaunchKernelGGL(k_read_stream, dim3(G), dim3(256), 0, stream(), (const f64x2*)buf.p, n2, part.p);
    DECO_CHECK_LAUNCH();
    Event ev0, ev1;
    hipEvent_t e0 = ev0.e, e1 = ev1.e;
    DECO_HIP(hipEventRecord(e0, stream()));
    for (int k = 0; k < iters; ++k)
      hipLaunchKernelGGL(k_read_stream, dim3(G), dim3(256), 0, stream(), (const f64x2*)buf.p, n2, part.p);
    DECO_CHECK_LAUNCH();
    DECO_HIP(hipEventRecord(e1, stream()));
    DECO_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DECO_HIP(hipEventElapsedTime(&ms, e0, e1));
    *gbps = (double)(n2 * 16) * iters / ((double)ms * 1e-3) / 1e9;
  });
}

int deco_spmv_block(deco_csr* A, int R, int C, const double* x, double* y) {
  return guard([&] {
    if (!A) fail(DECO_ERR_ARG, "null CSR");
    need(x, A->m, "x");
    need(y, A->n, "y");
    const deco_bsr* B = A->block64(R, C);
    if (!B) fail(DECO_ERR_ARG, "%d x %d blocks do not tile a %d x %d matrix", R, C, A->n, A->m);
    bsr_stream_d(B, x, EpiOuter{y});
    sync();
  });
}

int deco_spmv_block_time(deco_csr* A, int R, int C, int precision, int iters, double* avg_ms, int* nnzb) {
  return guard([&] {
    if (!A || !avg_ms || iters <= 0) fail(DECO_ERR_ARG, "bad argument");
    if (precision != DECO_PREC_FP64 && precision != DECO_PREC_FP32) fail(DECO_ERR_ARG, "bad precision %d", precision);
    const deco_bsr* B = precision == DECO_PREC_FP32 ? A->block32(R, C) : A->block64(R, C);
    if (!B) fail(DECO_ERR_ARG, "%d x %d blocks do not tile a %d x %d matrix", R, C, A->n, A->m);
    if (nnzb) *nnzb = B->nnzb;
    Event ev0, ev1;
    hipEvent_t e0 = ev0.e, e1 = ev1.e;
    if (precision == DECO_PREC_FP32) {
      DBuf<float> x(A->m), y(A->n);
      x.zero();
      bsr_stream_f(B, x.p, EpiStd<EPI_SET, float, float>{y.p, nullptr});
      DECO_HIP(hipEventRecord(e0, stream()));
      for (int k = 0; k < iters; ++k) bsr_stream_f(B, x.p, EpiStd<EPI_SET, float, float>{y.p, nullptr});
      DECO_HIP(hipEventRecord(e1, stream()));
    } else {
      DBuf<double> x(A->m), y(A->n);
      x.zero();
      bsr_stream_d(B, x.p, EpiOuter{y.p});
      DECO_HIP(hipEventRecord(e0, stream()));
      for (int k = 0; k < iters; ++k) bsr_stream_d(B, x.p, EpiOuter{y.p});
      DECO_HIP(hipEventRecord(e1, stream()));
    }
    DECO_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DECO_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = (double)ms / iters;
  });
}

int deco_spmv_time(deco_csr* A, int precision, int iters, double* avg_ms) {
  return guard([&] {
    if (!A || !avg_ms || iters <= 0) fail(DECO_ERR_ARG, "bad argument");
    if (precision != DECO_PREC_FP64 && precision != DECO_PREC_FP32) fail(DECO_ERR_ARG, "bad precision %d", precision);
    Event ev0, ev1;
    hipEvent_t e0 = ev0.e, e1 = ev1.e;
    if (precision == DECO_PREC_FP32) {
      const float* vf = A->values_f32();
      DBuf<float> x(A->m), y(A->n);
      x.zero();
      spmv_stream<EPI_SET, float, float, float, float>(A, vf, x.p, y.p, (const float*)nullptr);
      DECO_HIP(hipEventRecord(e0, stream()));
      for (int k = 0; k < iters; ++k) spmv_stream<EPI_SET, float, float, float, float>(A, vf, x.p, y.p, (const float*)nullptr);
      DECO_HIP(hipEventRecord(e1, stream()));
    } else {
      DBuf<double> x(A->m), y(A->n);
      x.zero();
      spmv_stream_f(A, A->v.p, x.p, EpiOuter{y.p});
      DECO_HIP(hipEventRecord(e0, stream()));
      for (int k = 0; k < iters; ++k) spmv_stream_f(A, A->v.p, x.p, EpiOuter{y.p});
      DECO_HIP(hipEventRecord(e1, stream()));
    }
    DECO_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    DECO_HIP(hipEventElapsedTime(&ms, e0, e1));
    *avg_ms = (double)ms / iters;
  });
}

}  // extern "C"
