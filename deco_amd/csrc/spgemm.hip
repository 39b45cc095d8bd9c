// Sparse matrix-matrix product on the device (core/MxM.cpp:113-146) and the FSAI power-2
// pattern (precond/FSAI/symbolic.tpp:149-172).
//
// One wavefront per output row. Pass 1 (symbolic): the row's candidate keys are gathered into
// LDS, bitonic-sorted and deduplicated; the count gives the row pointer (device scan). Pass 2
// (numeric): the keys are rebuilt, written, and the values are accumulated by walking the
// entries of the A row in order; for each A entry the lanes take the B row's entries, find their
// slot by binary search in the sorted keys and add a_ij*b_jk -- one writer per slot per step,
// so the sum of every output entry follows the (j,k) order of MxM_compute (MxM.cpp:229-240).
// Rows are binned by their upper bound of keys (table sizes 256 / 1024 / 4096).
#include "csr.hpp"
#include <climits>

namespace deco {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

// upper bound of keys per row: sum over j in row i of |row j of B| (MxM.cpp:9-31)
__global__ void k_spgemm_ub(int n, const int* __restrict__ ra, const int* __restrict__ ca, const int* __restrict__ rb,
                            int* __restrict__ ub) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int s = 0;
  for (int j = ra[i]; j < ra[i + 1]; ++j) s += rb[ca[j] + 1] - rb[ca[j]];
  ub[i] = s;
}

// bin rows by upper bound into lists (order inside a list is irrelevant for the result)
__global__ void k_spgemm_bin(int n, const int* __restrict__ ub, int* __restrict__ cnt, int* __restrict__ lists) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = -1;
  if (i < n) {
    const int u = ub[i];
    b = (u <= 256) ? 0 : (u <= 1024) ? 1 : (u <= 4096) ? 2 : 3;
  }
  wave_bin_append<4>(b, i, cnt, lists, (size_t)n);
}

// Wave-level bitonic sort of P (power of two) keys in LDS, then dedup. Returns unique count
// (keys[0..u) sorted ascending, valid in every lane).
template <int TS>
__device__ int wave_sort_unique(int* keys, int cnt) {
  const int lane = threadIdx.x & 63;
  int P = 1;
  while (P < cnt) P <<= 1;
  for (int k = cnt + lane; k < P; k += 64) keys[k] = INT_MAX;
  wave_sync();
  for (int size = 2; size <= P; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = lane; k < P; k += 64) {
        const int partner = k ^ stride;
        if (partner > k) {
          const bool up = ((k & size) == 0);
          const int a = keys[k], b = keys[partner];
          if ((a > b) == up) { keys[k] = b; keys[partner] = a; }
        }
      }
      wave_sync();
    }
  }
  // dedup: mark heads, prefix over the wave in chunks
  int total = 0;
  for (int base = 0; base < cnt; base += 64) {
    const int k = base + lane;
    int key = 0;
    bool head = false;
    if (k < cnt) {
      key = keys[k];
      head = (k == 0) || (keys[k - 1] != key);
    }
    const unsigned long long mask = __ballot(head);
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    wave_sync();
    if (head) keys[total + before] = key;  // total + before <= k, reads of this chunk are done
    wave_sync();
    total += __popcll(mask);
  }
  return total;
}

// TRIL: keep keys <= i only and stop each B row at the first key > i (symbolic.tpp:136,162)
template <int TS, int WPB, bool TRIL, bool NUMERIC>
__global__ __launch_bounds__(64 * WPB) void k_spgemm_rows(const int* __restrict__ list, int nrows,
                                                         const int* __restrict__ ra, const int* __restrict__ ca,
                                                         const double* __restrict__ va, const int* __restrict__ rb,
                                                         const int* __restrict__ cb, const double* __restrict__ vb,
                                                         int* __restrict__ count, const int* __restrict__ rc,
                                                         int* __restrict__ cc, double* __restrict__ vc) {
  __shared__ int s_keys[WPB][TS];
  __shared__ double s_vals[NUMERIC ? WPB : 1][NUMERIC ? TS : 1];
  __shared__ int s_fill[WPB];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int idx = blockIdx.x * WPB + w;
  if (idx >= nrows) return;
  const int i = list[idx];
  int* keys = s_keys[w];
  if (lane == 0) s_fill[w] = 0;
  wave_sync();
  // gather candidate keys: row j of B for each entry j of row i of A
  for (int j = ra[i]; j < ra[i + 1]; ++j) {
    const int col = ca[j];
    const int b0 = rb[col], b1 = rb[col + 1];
    const int base = s_fill[w];
    int taken = 0;
    for (int kk = b0; kk < b1; kk += 64) {
      const int k = kk + lane;
      int key = INT_MAX;
      bool ok = false;
      if (k < b1) {
        key = cb[k];
        ok = !TRIL || key <= i;
      }
      const unsigned long long mask = __ballot(ok);
      const int before = __popcll(mask & ((1ull << lane) - 1ull));
      if (ok) keys[base + taken + before] = key;
      taken += __popcll(mask);
    }
    wave_sync();
    if (lane == 0) s_fill[w] = base + taken;
    wave_sync();
  }
  const int cnt = s_fill[w];
  const int u = wave_sort_unique<TS>(keys, cnt);
  if (!NUMERIC) {
    if (lane == 0) count[i] = u;
    return;
  }
  double* vals = s_vals[w];
  for (int k = lane; k < u; k += 64) vals[k] = 0.0;
  wave_sync();
  for (int j = ra[i]; j < ra[i + 1]; ++j) {
    const int col = ca[j];
    const double a = va[j];
    const int b0 = rb[col], b1 = rb[col + 1];
    for (int kk = b0 + lane; kk < b1; kk += 64) {
      const int key = cb[kk];
      if (TRIL && key > i) continue;
      int lo = 0, hi = u - 1;
      while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (keys[mid] < key) lo = mid + 1; else hi = mid;
      }
      vals[lo] += vb[kk] * a;  // keys within one B row are distinct: one writer per slot
    }
    wave_sync();
  }
  const int o = rc[i];
  for (int k = lane; k < u; k += 64) {
    cc[o + k] = keys[k];
    vc[o + k] = vals[k];
  }
}

// Rows with more than 4096 candidate keys: one block per row with a dense flag/value
// accumulator of width B->m in global memory (slot per block), filled in the order of the A
// row's entries (same summation order as the LDS path), compacted in column order.
template <bool TRIL, bool NUMERIC>
__global__ __launch_bounds__(256) void k_spgemm_big(const int* __restrict__ list, int nrows, int m,
                                                   const int* __restrict__ ra, const int* __restrict__ ca,
                                                   const double* __restrict__ va, const int* __restrict__ rb,
                                                   const int* __restrict__ cb, const double* __restrict__ vb,
                                                   int* __restrict__ flags, double* __restrict__ vals,
                                                   int* __restrict__ count, const int* __restrict__ rc,
                                                   int* __restrict__ cc, double* __restrict__ vc) {
  __shared__ int sm[4];
  const int idx = blockIdx.x;
  if (idx >= nrows) return;
  const int i = list[idx];
  int* fl = flags + (size_t)idx * m;
  double* vl = vals + (size_t)idx * m;
  for (int k = threadIdx.x; k < m; k += 256) {
    fl[k] = 0;
    if (NUMERIC) vl[k] = 0.0;
  }
  __syncthreads();
  for (int j = ra[i]; j < ra[i + 1]; ++j) {
    const int col = ca[j];
    const double a = va[j];
    for (int kk = rb[col] + threadIdx.x; kk < rb[col + 1]; kk += 256) {
      const int key = cb[kk];
      if (TRIL && key > i) continue;
      fl[key] = 1;
      if (NUMERIC) vl[key] += vb[kk] * a;
    }
    __syncthreads();
  }
  // ordered compaction over the m columns
  int base = 0;
  for (int t0 = 0; t0 < m; t0 += 256) {
    const int k = t0 + threadIdx.x;
    const int f = (k < m) ? fl[k] : 0;
    int v = f;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
      const int t = __shfl_up(v, off, 64);
      if (lane >= off) v += t;
    }
    if (lane == 63) sm[wid] = v;
    __syncthreads();
    int add = 0, tot = 0;
    for (int w = 0; w < 4; ++w) {
      if (w < wid) add += sm[w];
      tot += sm[w];
    }
    if (NUMERIC && f) {
      const int pos = rc[i] + base + add + v - 1;
      cc[pos] = k;
      vc[pos] = vl[k];
    }
    base += tot;
    __syncthreads();
  }
  if (!NUMERIC && threadIdx.x == 0) count[i] = base;
}

template <bool TRIL>
static deco_csr* spgemm_impl(const deco_csr* A, const deco_csr* B, bool numeric) {
  if (A->m != B->n) fail(DECO_ERR_ARG, "SpGEMM shape mismatch: %dx%d times %dx%d", A->n, A->m, B->n, B->m);
  const int n = A->n;
  deco_csr* C = csr_new(n, B->m);
  if (n == 0) {
    C->r.alloc(1);
    C->r.zero();
    return C;
  }
  DBuf<int> ub(n), cnt(4), lists((size_t)4 * n), counts(n + 1);
  hipLaunchKernelGGL(k_spgemm_ub, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, A->r.p, A->c.p, B->r.p, ub.p);
  DECO_CHECK_LAUNCH();
  cnt.zero();
  hipLaunchKernelGGL(k_spgemm_bin, dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, (const int*)ub.p, cnt.p,
                     lists.p);
  DECO_CHECK_LAUNCH();
  int hc[4];
  copy_d2h(hc, cnt.p, 4);
  const int mB = B->m;
  const int chunk = 64;
  DBuf<int> bflags;
  DBuf<double> bvals;
  if (hc[3] > 0) {
    const int slots = std::min(hc[3], chunk);
    bflags.alloc((size_t)slots * mB);
    bvals.alloc((size_t)slots * mB);
  }
  auto run_big = [&](bool num, const int* rc, int* cc, double* vc) {
    for (int s0 = 0; s0 < hc[3]; s0 += chunk) {
      const int cntb = std::min(chunk, hc[3] - s0);
      const int* L = lists.p + 3 * (size_t)n + s0;
      if (num)
        hipLaunchKernelGGL((k_spgemm_big<TRIL, true>), dim3(cntb), dim3(256), 0, stream(), L, cntb, mB, A->r.p, A->c.p,
                           A->v.p, B->r.p, B->c.p, B->v.p, bflags.p, bvals.p, counts.p, rc, cc, vc);
      else
        hipLaunchKernelGGL((k_spgemm_big<TRIL, false>), dim3(cntb), dim3(256), 0, stream(), L, cntb, mB, A->r.p,
                           A->c.p, A->v.p, B->r.p, B->c.p, B->v.p, bflags.p, bvals.p, counts.p, rc, cc, vc);
      DECO_CHECK_LAUNCH();
    }
  };
  auto run = [&](bool num, const int* rc, int* cc, double* vc) {
    if (hc[0])
      hipLaunchKernelGGL((num ? k_spgemm_rows<256, 4, TRIL, true> : k_spgemm_rows<256, 4, TRIL, false>),
                         dim3(grid_for(hc[0], 4)), dim3(256), 0, stream(), (const int*)lists.p, hc[0], A->r.p, A->c.p,
                         A->v.p, B->r.p, B->c.p, B->v.p, counts.p, rc, cc, vc);
    if (hc[1])
      hipLaunchKernelGGL((num ? k_spgemm_rows<1024, 2, TRIL, true> : k_spgemm_rows<1024, 2, TRIL, false>),
                         dim3(grid_for(hc[1], 2)), dim3(128), 0, stream(), (const int*)(lists.p + n), hc[1], A->r.p,
                         A->c.p, A->v.p, B->r.p, B->c.p, B->v.p, counts.p, rc, cc, vc);
    if (hc[2])
      hipLaunchKernelGGL((num ? k_spgemm_rows<4096, 1, TRIL, true> : k_spgemm_rows<4096, 1, TRIL, false>),
                         dim3(hc[2]), dim3(64), 0, stream(), (const int*)(lists.p + 2 * (size_t)n), hc[2], A->r.p,
                         A->c.p, A->v.p, B->r.p, B->c.p, B->v.p, counts.p, rc, cc, vc);
    DECO_CHECK_LAUNCH();
  };
  run(false, nullptr, nullptr, nullptr);
  run_big(false, nullptr, nullptr, nullptr);
  build_row_ptr(C, counts.p, n);
  (void)numeric;  // the pattern-only product also writes values; the caller zeroes them
  run(true, C->r.p, C->c.p, C->v.p);
  run_big(true, C->r.p, C->c.p, C->v.p);
  return C;
}

deco_csr* spgemm(const deco_csr* A, const deco_csr* B) { return spgemm_impl<false>(A, B, true); }

deco_csr* spgemm_tril_pattern(const deco_csr* A) {
  deco_csr* F = spgemm_impl<true>(A, A, false);
  if (F->nnz) DECO_HIP(hipMemsetAsync(F->v.p, 0, sizeof(double) * F->nnz, stream()));
  return F;
}

}  // namespace deco

using namespace deco;

extern "C" {

int deco_spgemm(const deco_csr* A, const deco_csr* B, deco_csr** C) {
  return guard([&] {
    if (!A || !B || !C) fail(DECO_ERR_ARG, "null argument");
    *C = spgemm(A, B);
    sync();
  });
}

int deco_spgemm_tril_pattern(const deco_csr* A, deco_csr** F) {
  return guard([&] {
    if (!A || !F) fail(DECO_ERR_ARG, "null argument");
    if (A->n != A->m) fail(DECO_ERR_ARG, "matrix is not square");
    *F = spgemm_tril_pattern(A);
    sync();
  });
}

}  // extern "C"
