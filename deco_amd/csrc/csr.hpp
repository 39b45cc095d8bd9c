// Device CSR storage and the CSR kernels shared by the whole library.
//
// Layout in HBM: r int32[n+1], c int32[nnz] (sorted per row), v f64[nnz], plus an optional
// f32 shadow of v for the FP32 preconditioner. Unknown index = cell*bs + k (pressure k = 0).
#pragma once
#include "common.hpp"

struct deco_csr {
  int n = 0, m = 0, nnz = 0;
  deco::DBuf<int> r, c;
  deco::DBuf<double> v;
  deco::DBuf<float> vf;      // f32 copy of v, valid when vf_valid
  bool vf_valid = false;
  deco_csr* At = nullptr;    // cached transpose (adjoint SpMV)
  // SpMV row chunks: chunk k holds the rows whose first entry lies in [k*CAP, (k+1)*CAP)
  mutable deco::DBuf<int> rb;
  mutable int nchunks = 0;
  mutable bool rb_valid = false;
  ~deco_csr();
  const float* values_f32();  // creates / refreshes the shadow
  const int* row_chunks() const;  // creates the chunk table on first use
  void invalidate() { vf_valid = false; rb_valid = false; }
};

namespace deco {

// ------------------------------------------------------------------ SpMV ---------------
// One lane group of G lanes per output row o; A row = o*rs. Epilogues:
enum Epi : int {
  EPI_SET = 0,   // y[o] = sum
  EPI_RES = 1,   // y[o] = z[o*zs+zoff] - sum
  EPI_ACC = 2,   // y[o] = y[o] + sum
  EPI_ADDZ = 3,  // y[o] = z[o] + sum
};

template <class A, class B>
using acc_t = typename std::conditional<std::is_same<A, double>::value || std::is_same<B, double>::value,
                                        double, float>::type;

template <int G, int EPI, class VT, class XT, class YT, class ZT>
__global__ __launch_bounds__(256) void k_spmv(int nout, int rs, const int* __restrict__ r,
                                              const int* __restrict__ c, const VT* __restrict__ v,
                                              const XT* __restrict__ x, YT* __restrict__ y,
                                              const ZT* __restrict__ z, int zs, int zoff) {
  using AT = acc_t<VT, XT>;
  const int o = (int)((blockIdx.x * 256u + threadIdx.x) / G);
  const int lane = threadIdx.x & (G - 1);
  if (o >= nout) return;
  const int row = o * rs;
  const int b = r[row], e = r[row + 1];
  AT acc = AT(0);
  for (int j = b + lane; j < e; j += G) {
    const int col = __builtin_nontemporal_load(c + j);
    const VT a = __builtin_nontemporal_load(v + j);
    acc += (AT)a * (AT)x[col];
  }
  acc = group_sum<G>(acc);
  if (lane == 0) {
    if constexpr (EPI == EPI_SET) y[o] = (YT)acc;
    else if constexpr (EPI == EPI_RES) y[o] = (YT)((AT)z[(int64_t)o * zs + zoff] - acc);
    else if constexpr (EPI == EPI_ACC) y[o] = (YT)((AT)y[o] + acc);
    else y[o] = (YT)((AT)z[o] + acc);
  }
}

// Lane-group width for a given average row length.
inline int spmv_group(double avg) {
  if (avg <= 2.5) return 2;
  if (avg <= 5.0) return 4;
  if (avg <= 10.0) return 8;
  if (avg <= 24.0) return 16;
  if (avg <= 48.0) return 32;
  return 64;
}

template <int EPI, class VT, class XT, class YT, class ZT>
void launch_spmv(int G, int nout, int rs, const int* r, const int* c, const VT* v, const XT* x,
                 YT* y, const ZT* z, int zs, int zoff) {
  if (nout <= 0) return;
  hipStream_t s = stream();
  const unsigned grid = grid_for((int64_t)nout * G, 256);
#define DECO_SPMV_CASE(GG)                                                                 \
  case GG:                                                                                 \
    hipLaunchKernelGGL((k_spmv<GG, EPI, VT, XT, YT, ZT>), dim3(grid), dim3(256), 0, s, nout, \
                       rs, r, c, v, x, y, z, zs, zoff);                                    \
    break;
  switch (G) {
    DECO_SPMV_CASE(2)
    DECO_SPMV_CASE(4)
    DECO_SPMV_CASE(8)
    DECO_SPMV_CASE(16)
    DECO_SPMV_CASE(32)
    DECO_SPMV_CASE(64)
    default: fail(DECO_ERR_ARG, "bad SpMV group %d", G);
  }
#undef DECO_SPMV_CASE
  DECO_CHECK_LAUNCH();
}

// ------------------------------------------------------------------ CSR-stream SpMV ------
// Chunk k holds the rows whose first entry lies in [k*kStreamCap, (k+1)*kStreamCap). A block
// streams the contiguous entry range of its chunk(s) into LDS: every lane issues 6 (FP64) or 12
// (FP32) independent (column, value) loads and then the x gathers, so many bytes are in flight
// per wave; the products are then summed per row in column order (the same order as alg::spmv,
// core/alg.hpp:59-61). Ranges that exceed the LDS tile (rows over 129 entries) fall back to a
// block reduction per row.
constexpr int kStreamCap = 3072;                 // target entries per chunk
constexpr int kStreamBlock = 512;

// Epilogue functors: called once per row with the row sum (accumulator type AT).
template <int EPI, class YT, class ZT>
struct EpiStd {
  YT* __restrict__ y;
  const ZT* __restrict__ z;
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc) const {
    if constexpr (EPI == EPI_SET) y[o] = (YT)acc;
    else if constexpr (EPI == EPI_RES) y[o] = (YT)((AT)z[o] - acc);
    else if constexpr (EPI == EPI_ACC) y[o] = (YT)((AT)y[o] + acc);
    else y[o] = (YT)((AT)z[o] + acc);
  }
};

// y[o] = sum for the Krylov solvers' product with the system matrix: its own kernel
// instantiation, so a profile separates it from the preconditioner's products
struct EpiOuter {
  double* __restrict__ y;
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc) const { y[o] = (double)acc; }
};

// y[o] = z[o*zs] - sum: residual of a row subset (rows zs*o of the full matrix)
template <class YT, class ZT>
struct EpiStrideRes {
  YT* __restrict__ y;
  const ZT* __restrict__ z;
  int zs;
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc) const { y[o] = (YT)((AT)z[(size_t)o * zs] - acc); }
};

// Chunks per block (one: two chunks per block for FP32, 12 loads per lane, measured 3-7% slower
// on the config-3 operators and slower still on small AMG levels with half the blocks).
template <class VT, class XT>
constexpr int stream_cpb() {
  return 1;
}

template <class VT, class XT, class E>
__global__ __launch_bounds__(kStreamBlock) void k_spmv_stream(int nch, const int* __restrict__ rb,
                                                             const int* __restrict__ r, const int* __restrict__ c,
                                                             const VT* __restrict__ v, const XT* __restrict__ x,
                                                             E epi) {
  using AT = acc_t<VT, XT>;
  constexpr int CPB = stream_cpb<VT, XT>();
  constexpr int CAP = CPB * kStreamCap;
  constexpr int LDSN = CAP + 128;      // a group of CPB chunks holds <= CAP + longest row - 1
  constexpr int U = CAP / kStreamBlock;
  __shared__ AT prod[LDSN];
  const int grp = xcd_block(blockIdx.x, gridDim.x);
  const int r0 = rb[grp * CPB], r1 = rb[min(grp * CPB + CPB, nch)];
  if (r0 >= r1) return;
  const int j0 = r[r0], j1 = r[r1];
  const int cnt = j1 - j0;
  const int tid = threadIdx.x;
  if (cnt <= LDSN) {
    // first tile: U independent loads per lane
    int cc[U];
    VT vv[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = tid + u * kStreamBlock;
      if (t < cnt) {
        cc[u] = __builtin_nontemporal_load(c + j0 + t);
        vv[u] = __builtin_nontemporal_load(v + j0 + t);
      }
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int t = tid + u * kStreamBlock;
      if (t < cnt) prod[t] = (AT)vv[u] * (AT)x[cc[u]];
    }
    for (int t = tid + CAP; t < cnt; t += kStreamBlock)
      prod[t] = (AT)__builtin_nontemporal_load(v + j0 + t) * (AT)x[__builtin_nontemporal_load(c + j0 + t)];
    __syncthreads();
    for (int row = r0 + tid; row < r1; row += kStreamBlock) {
      const int b = r[row] - j0, e = r[row + 1] - j0;
      AT acc = AT(0);
      for (int k = b; k < e; ++k) acc += prod[k];
      epi(row, acc);
    }
  } else {
    __shared__ AT red[kStreamBlock / 64];
    for (int row = r0; row < r1; ++row) {
      const int b = r[row], e = r[row + 1];
      AT acc = AT(0);
      for (int j = b + tid; j < e; j += kStreamBlock) acc += (AT)v[j] * (AT)x[c[j]];
      acc = block_sum<kStreamBlock>(acc, red);
      if (tid == 0) epi(row, acc);
    }
  }
}

template <class VT, class XT, class E>
void spmv_stream_f(const deco_csr* A, const VT* v, const XT* x, E epi) {
  if (A->n <= 0) return;
  const int* rb = A->row_chunks();
  constexpr int CPB = stream_cpb<VT, XT>();
  hipLaunchKernelGGL((k_spmv_stream<VT, XT, E>), dim3((A->nchunks + CPB - 1) / CPB), dim3(kStreamBlock), 0, stream(),
                     A->nchunks, rb, A->r.p, A->c.p, v, x, epi);
  DECO_CHECK_LAUNCH();
}

template <int EPI, class VT, class XT, class YT, class ZT>
void spmv_stream(const deco_csr* A, const VT* v, const XT* x, YT* y, const ZT* z) {
  spmv_stream_f(A, v, x, EpiStd<EPI, YT, ZT>{y, z});
}

// Convenience wrappers on a deco_csr (values f64 or the f32 shadow).
inline int csr_group(const deco_csr* A, int rs = 1) {
  const int rows = A->n / rs;
  return spmv_group(rows > 0 ? (double)A->nnz / (double)A->n : 1.0);
}

template <class XT, class YT>
void spmv(const deco_csr* A, const XT* x, YT* y) {
  spmv_stream<EPI_SET, double, XT, YT, double>(A, A->v.p, x, y, nullptr);
}
template <class XT, class YT, class ZT>
void spmv_res(const deco_csr* A, const XT* x, const ZT* z, YT* y) {
  spmv_stream<EPI_RES, double, XT, YT, ZT>(A, A->v.p, x, y, z);
}

// ------------------------------------------------------------------ scans --------------
// out[0..n] = exclusive prefix sum of in[0..n-1] (out[n] = total); in and out may alias.
// Returns the total (read back to the host).
int64_t scan_to_ptr(const int* in, int* out, int n);
// Same without the host readback.
void scan_to_ptr_async(const int* in, int* out, int n);

// Builds a CSR in `dst` from per-row counts (device, length n): r = scan(counts), allocates
// c and v with the total. Returns nnz.
int build_row_ptr(deco_csr* dst, const int* counts, int n);

// Sort every row of (c, v) by column (rows of any length up to 4096 entries).
void sort_rows(int n, const int* r, int* c, double* v, int max_row);

int max_row_length(const deco_csr* A);

deco_csr* csr_transpose(const deco_csr* A);
// rows stride*i + off of A (all columns), as an (n/stride) x m CSR
deco_csr* csr_row_subset(const deco_csr* A, int stride, int off);
deco_csr* csr_clone(const deco_csr* A);
deco_csr* csr_new(int n, int m);

}  // namespace deco
