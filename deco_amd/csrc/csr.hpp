// Device CSR storage and the CSR kernels shared by the whole library.
//
// Layout in HBM: r int32[n+1], c int32[nnz] (sorted per row), v f64[nnz], plus an optional
// f32 shadow of v for the FP32 preconditioner. Unknown index = cell*bs + k (pressure k = 0).
#pragma once
#include "common.hpp"

// Block-CSR shadow of a CSR matrix whose rows and columns come in groups (R rows per block
// row, C columns per block column): block row I covers rows I*R .. I*R+R-1, block (I, J) holds
// the R x C values of columns J*C .. J*C+C-1, row-major, zeros where the CSR has no entry. The
// FP32 values are made with the structure; the FP64 values on first FP64 use.
// Two layouts. Chunked block CSR: chunk table of the block rows whose first block lies in
// [k*kBsrCap, (k+1)*kBsrCap), as (first block row, its first block) pairs. Sliced (SELL): block
// rows in slices of 64, one lane per block row; entry k of block row 64*s + l sits at
// soff[s] + 64*k + l (slices padded to their longest block row with zero blocks of column 0),
// chosen when that padding stays within the library's threshold (deco_set_sell_padding).
struct deco_bsr {
  int R = 1, C = 1, nb = 0, nnzb = 0, nchunks = 0;
  int mb = 0;                  // block columns
  deco::DBuf<int> br, bc;
  deco::DBuf<int2> rb;
  deco::DBuf<float> bv;
  deco::DBuf<double> bv64;
  bool has64 = false;
  bool sell = false;
  int nslices = 0;
  size_t stot = 0;             // entries of the sliced layout, padding included
  deco::DBuf<int> soff, swid;  // per slice: first entry, width (blocks per block row)
  // sliced layout: each entry's block column as a 16-bit offset from its block row (padding and
  // lanes past the last block row: 0), read instead of bc when every offset fits and the block
  // rows do not outnumber the block columns (so an offset 0 stays inside x)
  deco::DBuf<short> bcd;
  bool d16 = false;
  // uniform-offset form of the sliced layout (stencil matrices): the k-th entry of every block row
  // I is the block at block column I + doff[k] (noff sorted offsets shared by all block rows; zero
  // blocks where a row has none, read at a column clamped into x), so no column is stored and
  // every slice has width noff. Chosen before the plain sliced layout when the matrix has at most
  // kDiaMax distinct offsets and the zero blocks stay within the padding threshold.
  bool dia = false;
  int noff = 0;
  deco::DBuf<int> doff;
};

// Row-piece sliced form of a CSR matrix for the stream products of the AMG levels: every row is
// cut into pieces of W entries (the last one padded with zero values; an empty row gets one
// padding piece), the pieces of consecutive rows fill slices of 64 (a row's pieces never cross a
// slice: the slice is closed with padding lanes first), and slot t of lane l of slice s holds
// entry t of that lane's piece at (s * W + t) * 64 + l. A wavefront takes a slice: each lane sums
// its piece, the lanes of one row add their sums in a segmented shuffle tree, and the row's first
// lane applies the epilogue. No LDS, no barrier, one lane per W entries.
struct deco_pieces {
  int W = 1;                         // entries per piece: 1, 2, 4 or 8
  int nslices = 0;
  size_t nslot = 0;                  // nslices * 64 * W
  deco::DBuf<int> srow;              // per slice: the row of its first piece
  deco::DBuf<unsigned char> pdel;    // per lane: its row - srow (255: padding lane)
  deco::DBuf<int> cbase;             // per slice: its smallest column (16-bit columns)
  deco::DBuf<unsigned short> c16;    // per slot: column - cbase
  deco::DBuf<int> c32;               // per slot: column, when a slice spans more than 65535 columns
  deco::DBuf<float> vf;              // per slot: FP32 value
  deco::DBuf<double> v64;            // per slot: FP64 value (filled on first FP64 use)
  bool d16 = false, has64 = false;
};

// Column-sorted form of the CSR-stream chunks: the entries of each chunk's own rows, [j0, j1),
// stored by ascending column, each with the slot of its row-order position in the chunk's window
// (j - k*cap, below 4096), so that the stream kernel's x gathers of one wavefront instruction touch
// a few adjacent cache lines while the products still land in row order in LDS (the same sums, bit
// for bit). Packed per entry: (column - cbase[chunk]) in the low 20 bits, the slot in the high 12.
struct deco_csort {
  deco::DBuf<int> cbase;             // per chunk: its smallest column
  deco::DBuf<unsigned int> p;        // per entry: column offset | slot << 20
  deco::DBuf<float> vf;              // per entry: FP32 value in the same order
  deco::DBuf<double> v64;            // per entry: FP64 value (filled on first FP64 use)
  bool has32 = false, has64 = false;
};
constexpr int kSortColBits = 20;
constexpr unsigned int kSortColMask = (1u << kSortColBits) - 1u;

struct deco_csr {
  int n = 0, m = 0, nnz = 0;
  deco::DBuf<int> r, c;
  deco::DBuf<double> v;
  deco::DBuf<float> vf;      // f32 copy of v, valid when vf_valid
  bool vf_valid = false;
  deco_csr* At = nullptr;    // cached transpose (adjoint SpMV)
  // SpMV row chunks: chunk k holds the rows whose first entry lies in [k*cap, (k+1)*cap)
  // as (first row, its first entry) pairs, nchunks + 1 of them
  mutable deco::DBuf<int2> rb;
  mutable int nchunks = 0;
  mutable int chunk_cap = 0;
  mutable int chunk_win = 0;   // entries a chunk's range can span: cap + longest row - 1 (at most kStreamTile)
  mutable bool chunks_fit = false;   // every chunk's range lies in its window (no row above half the tile)
  mutable bool rb_valid = false;
  ~deco_csr();
  const float* values_f32();  // creates / refreshes the shadow
  const int2* row_chunks() const;  // creates the chunk table on first use
  mutable deco_bsr* b32 = nullptr;  // block shadow (dropped with the FP32 shadow by invalidate())
  // the block shadow with R x C blocks and FP32 values (created on first use); null when n or
  // m do not divide into such blocks
  const deco_bsr* block32(int R, int C) const;
  // the same with its FP64 values filled
  const deco_bsr* block64(int R, int C) const;
  // unknowns per cell of the system this matrix belongs to (set by the CPR setup): above 1 the
  // solvers' FP64 product with the matrix runs on the cell x cell block shadow
  int cell = 1;
  // whether the rows fit the sliced layout within the padding threshold (decided on first call)
  bool sell_scalar() const;
  mutable signed char sell_ok = -1;
  // Columns as 16-bit offsets from a per-chunk base (the smallest column of each stream chunk),
  // for the stream products of the AMG levels: built on first call when every chunk's columns
  // span at most 65535, else false (32-bit columns). Dropped by invalidate().
  bool cols16() const;
  mutable deco::DBuf<unsigned short> c16;
  mutable deco::DBuf<int> cbase;
  mutable signed char c16_ok = -1;
  // the row-piece sliced form (deco_pieces), built on first call when its slots stay within 1.3
  // times the entries for some W, else null; fp64: with its FP64 values. Dropped by invalidate().
  const deco_pieces* pieces(bool fp64) const;
  mutable deco_pieces* pc = nullptr;
  mutable signed char pc_ok = -1;
  // the column-sorted chunk form (deco_csort) for products with FP64 (fp64) or FP32 values (from
  // the f32 shadow, created if needed), built on first call when every chunk fits its window and
  // spans fewer than 2^20 columns; null where it does not apply or, by default, where the
  // row-order form has 16-bit columns and is expected to be faster (FP64 values; FP32 values with
  // row-order gathers on few segments of x). Dropped by invalidate().
  const deco_csort* sorted(bool fp64) const;
  mutable deco_csort* cs = nullptr;
  mutable signed char cs_ok = -1;
  // distinct 64-byte segments of an FP32 x per 64 row-order gathers, on average (computed once)
  double gather_segments() const;
  mutable double gather_seg = -1.0;
  // count of invalidate() calls: matrices derived from this one compare it to see a change
  unsigned long long gen = 0;
  void invalidate();
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
// Chunk k holds the rows whose first entry lies in [k*cap, (k+1)*cap), so its entries lie in the
// window [k*cap, k*cap + win), win = cap + longest row - 1 (deco_csr::row_chunks lowers the cap
// below kStreamCap when win would exceed kStreamTile). A workgroup loads the whole window -- every
// lane 13 (FP64, 256 lanes) or 7 (FP32, 512 lanes) independent (column, value) loads, indices
// past the window clamped to its last entry, then all of its x gathers -- without waiting for the
// chunk's row range, which loads in parallel and only drives the sums: one dependent memory round
// fewer than loading the range first. Entries of the window outside the chunk's rows (the
// previous chunk's last row, the next chunk's first rows) are gathered at a clamped column and
// their products dropped. Rows shorter than 8 entries on average are summed in column order (the
// same order as alg::spmv, core/alg.hpp:59-61); longer rows by a lane group and a fixed xor tree.
// A chunk whose range exceeds the window (a row longer than half the tile) falls back to a block
// reduction per row.
constexpr int kStreamCap = 3072;                 // target entries per chunk
constexpr int kStreamTile = kStreamCap + 128;    // window entries a chunk's range must fit
constexpr int kStreamPre = 3;                    // a lane's further rows whose operands load together
template <class VT>
constexpr int stream_block() {
  return sizeof(VT) == 8 ? 256 : 512;
}

// Epilogue functors. pre(o) loads the operands row o needs besides its sum; the stream kernels
// issue it before the tile's gathers, so its latency is hidden behind them.
// operator()(o, acc, pre) writes row o from its sum (accumulator type AT).
template <int EPI, class YT, class ZT>
struct EpiStd {
  YT* __restrict__ y;
  const ZT* __restrict__ z;
  using PT = typename std::conditional<EPI == EPI_ACC, YT, ZT>::type;
  struct Pre { PT a; };
  __device__ __forceinline__ Pre pre(int o) const {
    if constexpr (EPI == EPI_SET) return Pre{PT(0)};
    else if constexpr (EPI == EPI_ACC) return Pre{y[o]};
    else return Pre{z[o]};
  }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    if constexpr (EPI == EPI_SET) y[o] = (YT)acc;
    else if constexpr (EPI == EPI_RES) y[o] = (YT)((AT)p.a - acc);
    else y[o] = (YT)((AT)p.a + acc);                    // EPI_ACC, EPI_ADDZ
  }
  // vector bytes per row (the byte count of deco_count_bytes)
  size_t row_bytes() const {
    if (EPI == EPI_SET) return sizeof(YT);
    if (EPI == EPI_ACC) return 2 * sizeof(YT);
    return sizeof(ZT) + sizeof(YT);
  }
};

// y[o] = z[o] - sum, and for every bs-th row (the pressure rows of a system of bs unknowns per cell)
// also yp[o / bs] = y[o]: a residual and its pressure restriction in one pass
template <class YT, class ZT>
struct EpiResP {
  YT* __restrict__ y;
  const ZT* __restrict__ z;
  YT* __restrict__ yp;
  int bs;
  struct Pre { ZT a; };
  __device__ __forceinline__ Pre pre(int o) const { return Pre{z[o]}; }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    const YT r = (YT)((AT)p.a - acc);
    y[o] = r;
    if (o % bs == 0) yp[o / bs] = r;
  }
  size_t row_bytes() const { return sizeof(ZT) + sizeof(YT) + sizeof(YT) / bs; }
};

// y[o] = sum for the Krylov solvers' product with the system matrix: its own kernel
// instantiation, so a profile separates it from the preconditioner's products
struct EpiOuter {
  double* __restrict__ y;
  struct Pre {};
  __device__ __forceinline__ Pre pre(int) const { return Pre{}; }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre&) const { y[o] = (double)acc; }
  size_t row_bytes() const { return sizeof(double); }
};

// y[o] = z[o*zs] - sum: residual of a row subset (rows zs*o of the full matrix)
template <class YT, class ZT>
struct EpiStrideRes {
  YT* __restrict__ y;
  const ZT* __restrict__ z;
  int zs;
  struct Pre { ZT a; };
  __device__ __forceinline__ Pre pre(int o) const { return Pre{z[(size_t)o * zs]}; }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const { y[o] = (YT)((AT)p.a - acc); }
  size_t row_bytes() const { return sizeof(ZT) + sizeof(YT); }
};

// EpiStrideRes that also writes x0[o] = 0 + (y[o] - 0)/d[o]: the zero-guess Jacobi-type sweep on
// the residual it produces (the AMG's k_relax_zero arithmetic), so the cycle starts one launch later
template <class YT, class ZT>
struct EpiStrideResJ0 {
  YT* __restrict__ y;
  const ZT* __restrict__ z;
  int zs;
  const YT* __restrict__ d;
  YT* __restrict__ x0;
  struct Pre { ZT a; YT d; };
  __device__ __forceinline__ Pre pre(int o) const { return Pre{z[(size_t)o * zs], d[o]}; }
  template <class AT>
  __device__ __forceinline__ void operator()(int o, AT acc, const Pre& p) const {
    const YT bo = (YT)((AT)p.a - acc);
    y[o] = bo;
    x0[o] = (YT)0 + (bo - (YT)0) / p.d;
  }
  size_t row_bytes() const { return sizeof(ZT) + 3 * sizeof(YT); }
};

// x access of the stream kernel: a plain vector, or a functor that computes x_j (fused steps)
template <class XT>
struct PtrGather {
  using type = XT;
  const XT* __restrict__ x;
  __device__ __forceinline__ XT operator()(int j) const { return x[j]; }
};

// Sum of prod[k], prod[k + G], ... below e: in that order when inorder (rows averaging under 8
// entries, summed as alg::spmv sums them), else in four interleaved partial sums added pairwise,
// so four LDS reads are in flight instead of one dependent chain
template <class AT>
__device__ __forceinline__ AT stream_row_sum(const AT* prod, int k, int e, int G, bool inorder) {
  if (inorder) {
    AT a = AT(0);
    for (; k < e; k += G) a += prod[k];
    return a;
  }
  AT a0 = AT(0), a1 = AT(0), a2 = AT(0), a3 = AT(0);
  for (; k + 3 * G < e; k += 4 * G) {
    a0 += prod[k];
    a1 += prod[k + G];
    a2 += prod[k + 2 * G];
    a3 += prod[k + 3 * G];
  }
  for (; k < e; k += G) a0 += prod[k];
  return (a0 + a1) + (a2 + a3);
}

// CT int: c holds the columns. CT unsigned short: c holds 16-bit offsets from cbase[chunk]
// (deco_csr::cols16). CT unsigned int: the column-sorted form (deco_csort; every chunk fits its
// window): c holds column offset | slot << 20 and v the values in the same order; an entry of the
// chunk's own range writes its product to its slot, the window's other entries to a slot no row
// reads.
template <class VT, class XG, class E, class CT = int>
__global__ __launch_bounds__(stream_block<VT>()) void k_spmv_stream(const int2* __restrict__ desc,
                                                                   const int* __restrict__ r,
                                                                   const CT* __restrict__ c,
                                                                   const VT* __restrict__ v, XG x, E epi,
                                                                   const int* __restrict__ cbase, int cap, int win,
                                                                   int nnz, int m) {
  using XT = typename XG::type;
  using AT = acc_t<VT, XT>;
  constexpr int BLK = stream_block<VT>();
  constexpr int U = (kStreamTile + BLK - 1) / BLK;
  constexpr int WIN = U * BLK;
  constexpr bool SORTED = std::is_same<CT, unsigned int>::value;
  static_assert(WIN - 1 >= kStreamTile && WIN <= (1 << (32 - kSortColBits)), "slot range");
  __shared__ AT prod[WIN];
  const int ch = xcd_block(blockIdx.x, gridDim.x);
  const int w0 = ch * cap;   // window start
  const int tid = threadIdx.x;
  int cb = 0;
  if constexpr (!std::is_same<CT, int>::value) cb = cbase[ch];
  // The window's columns, decoded as they load; in the sorted form with FP64 values the packed words
  // stay raw (one register each) and are decoded at the gather and at the LDS write: decoded at
  // load, the 13 column and slot pairs made the compiler wait on each packed word before the next
  // value load (the FP64 product with A 141 -> 157 us), while FP32 products are faster decoded early.
  constexpr bool RAW = SORTED && sizeof(VT) == 8;
  int cc[U];
  VT vv[U];
  int slot[(SORTED && !RAW) ? U : 1];
  const int jl = min(w0 + win, nnz) - 1;   // last entry of the window
  if (nnz > 0) {
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int j = min(w0 + tid + u * BLK, jl);
      if constexpr (RAW) {
        cc[u] = (int)__builtin_nontemporal_load(c + j);
      } else if constexpr (SORTED) {
        const unsigned int p = __builtin_nontemporal_load(c + j);
        cc[u] = min(cb + (int)(p & kSortColMask), m - 1);
        slot[u] = (int)(p >> kSortColBits);
      } else {
        cc[u] = min(cb + (int)__builtin_nontemporal_load(c + j), m - 1);
      }
      vv[u] = __builtin_nontemporal_load(v + j);
    }
  }
  const int2 d0 = desc[ch], d1 = desc[ch + 1];   // (first row, its first entry) of chunks ch, ch+1
  const int r0 = d0.x, r1 = d1.x;
  const int j0 = d0.y, j1 = d1.y;
  const int cnt = j1 - j0;
  const bool fits = j1 - w0 <= win;
  // Lanes per row, 2^lg: 1 (the row summed in column order) unless the rows average 8 or more
  // entries, then the widest group (at most a wavefront) that gives every row of the chunk its
  // own group, so a long row's sum is not one lane's chain of LDS reads.
  const int nr = r1 - r0;
  int lg = 0;
  if (cnt >= 8 * nr)
    while (lg < 6 && (nr << (lg + 1)) <= BLK) ++lg;
  const int sub = tid & ((1 << lg) - 1);
  // this lane's first row: bounds and epilogue operands load behind the window's gathers
  const int row = r0 + (tid >> lg);
  int rbeg = 0, rend = 0;
  typename E::Pre pre{};
  if (row < r1 && fits) {
    rbeg = r[row];
    rend = r[row + 1];
    pre = epi.pre(row);
  }
  if (nnz > 0) {
    XT xx[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (RAW) xx[u] = x(min(cb + (int)((unsigned int)cc[u] & kSortColMask), m - 1));
      else xx[u] = x(cc[u]);
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      if constexpr (SORTED) {
        const int j = w0 + tid + u * BLK;
        int sl;
        if constexpr (RAW) sl = (int)((unsigned int)cc[u] >> kSortColBits);
        else sl = slot[u];
        prod[(j >= j0 && j < j1) ? sl : WIN - 1] = (AT)vv[u] * (AT)xx[u];
      } else {
        prod[tid + u * BLK] = (AT)vv[u] * (AT)xx[u];
      }
    }
  }
  __syncthreads();
  if (fits) {
    if (row < r1) {   // a lane group is active or inactive as a whole
      const int G = 1 << lg, S = BLK >> lg;
      const bool inorder = cnt < 8 * nr;   // short rows: column order
      // A chunk of more rows than lane groups gives a lane further rows, row + S, row + 2S, ...:
      // the bounds and epilogue operands of the next kStreamPre of them load together here, after
      // the window's registers are free, so their memory rounds overlap instead of following one
      // another in the loop below.
      int pb[kStreamPre], pe[kStreamPre];
      typename E::Pre pp[kStreamPre];
#pragma unroll
      for (int k = 0; k < kStreamPre; ++k) {
        const int rw = row + (k + 1) * S;
        pb[k] = 0;
        pe[k] = 0;
        pp[k] = typename E::Pre{};
        if (rw < r1) {
          pb[k] = r[rw];
          pe[k] = r[rw + 1];
          pp[k] = epi.pre(rw);
        }
      }
      AT acc = stream_row_sum<AT>(prod, rbeg - w0 + sub, rend - w0, G, inorder);
      for (int off = G >> 1; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
      if (sub == 0) epi(row, acc, pre);
#pragma unroll
      for (int k = 0; k < kStreamPre; ++k) {
        const int rw = row + (k + 1) * S;
        if (rw < r1) {
          AT a2 = stream_row_sum<AT>(prod, pb[k] - w0 + sub, pe[k] - w0, G, inorder);
          for (int off = G >> 1; off > 0; off >>= 1) a2 += __shfl_xor(a2, off);
          if (sub == 0) epi(rw, a2, pp[k]);
        }
      }
      for (int rw = row + (kStreamPre + 1) * S; rw < r1; rw += S) {
        const int b = r[rw] - w0, e = r[rw + 1] - w0;
        AT a2 = stream_row_sum<AT>(prod, b + sub, e, G, inorder);
        for (int off = G >> 1; off > 0; off >>= 1) a2 += __shfl_xor(a2, off);
        if (sub == 0) epi(rw, a2, epi.pre(rw));
      }
    }
  } else if constexpr (!SORTED) {   // (the sorted form exists only when every chunk fits)
    __shared__ AT red[BLK / 64];
    for (int rw = r0; rw < r1; ++rw) {
      const int b = r[rw], e = r[rw + 1];
      AT acc = AT(0);
      for (int j = b + tid; j < e; j += BLK) acc += (AT)v[j] * (AT)x(cb + (int)c[j]);
      acc = block_sum<BLK>(acc, red);
      if (tid == 0) epi(rw, acc, epi.pre(rw));
    }
  }
}

// Row-piece product (deco_pieces): one wavefront per slice of 64 pieces, W entries per lane with
// all their loads in flight before the gathers; a row's sum is its pieces' sums in a fixed
// segmented tree (lane l adds lane l + 2^k's partial while both hold the same row), so it is
// deterministic; a row of one piece sums its entries in column order.
template <class VT, class XG, class E, class CT, int W>
__global__ __launch_bounds__(256) void k_spmv_piece(int nslices, const int* __restrict__ srow,
                                                    const unsigned char* __restrict__ pdel,
                                                    const int* __restrict__ cbase, const CT* __restrict__ col,
                                                    const VT* __restrict__ val, XG x, E epi) {
  using XT = typename XG::type;
  using AT = acc_t<VT, XT>;
  const int s = xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
  if (s >= nslices) return;
  const int lane = threadIdx.x & 63;
  const int rd = pdel[(size_t)s * 64 + lane];
  const int r0 = srow[s];
  int cb = 0;
  if constexpr (!std::is_same<CT, int>::value) cb = cbase[s];
  int cc[W];
  VT vv[W];
#pragma unroll
  for (int t = 0; t < W; ++t) {
    const size_t e = ((size_t)s * W + t) * 64 + lane;
    cc[t] = cb + (int)__builtin_nontemporal_load(col + e);
    vv[t] = __builtin_nontemporal_load(val + e);
  }
  const int prev = __shfl_up(rd, 1);
  const bool live = rd != 255;
  const bool lead = live && (lane == 0 || prev != rd);
  const int row = r0 + rd;
  typename E::Pre pre{};
  if (lead) pre = epi.pre(row);
  XT xx[W];
#pragma unroll
  for (int t = 0; t < W; ++t) xx[t] = x(cc[t]);
  AT acc = AT(0);
#pragma unroll
  for (int t = 0; t < W; ++t) acc += (AT)vv[t] * (AT)xx[t];
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const AT o = __shfl_down(acc, off);
    const int od = __shfl_down(rd, off);
    if (lane + off < 64 && od == rd) acc += o;
  }
  if (lead) epi(row, acc, pre);
}

template <class VT, class XG, class E, class CT>
void spmv_piece_w(const deco_pieces* P, const CT* col, const VT* val, XG x, E epi) {
  const dim3 grid((P->nslices + 3) / 4), blk(256);
  switch (P->W) {
    case 1: hipLaunchKernelGGL((k_spmv_piece<VT, XG, E, CT, 1>), grid, blk, 0, stream(), P->nslices, P->srow.p,
                               P->pdel.p, P->cbase.p, col, val, x, epi); break;
    case 2: hipLaunchKernelGGL((k_spmv_piece<VT, XG, E, CT, 2>), grid, blk, 0, stream(), P->nslices, P->srow.p,
                               P->pdel.p, P->cbase.p, col, val, x, epi); break;
    case 4: hipLaunchKernelGGL((k_spmv_piece<VT, XG, E, CT, 4>), grid, blk, 0, stream(), P->nslices, P->srow.p,
                               P->pdel.p, P->cbase.p, col, val, x, epi); break;
    default: hipLaunchKernelGGL((k_spmv_piece<VT, XG, E, CT, 8>), grid, blk, 0, stream(), P->nslices, P->srow.p,
                                P->pdel.p, P->cbase.p, col, val, x, epi); break;
  }
  DECO_CHECK_LAUNCH();
}

// the product on A's row-piece form when it has one (values of A in VT), true if launched
template <class VT, class XG, class E>
bool spmv_piece(const deco_csr* A, XG x, E epi) {
  const deco_pieces* P = A->pieces(sizeof(VT) == 8);
  if (!P) return false;
  const VT* val;
  if constexpr (sizeof(VT) == 8) val = P->v64.p;
  else val = P->vf.p;
  if (P->d16) spmv_piece_w(P, (const unsigned short*)P->c16.p, val, x, epi);
  else spmv_piece_w(P, (const int*)P->c32.p, val, x, epi);
  count_bytes((double)A->nnz * (sizeof(VT) + (P->d16 ? 2 : 4)) + 65.0 * P->nslices + 4.0 * P->nslices +
              (double)sizeof(typename XG::type) * A->m + (double)A->n * epi.row_bytes());
  return true;
}

// the product on A's column-sorted chunk form when it has one (values of A in VT), true if launched
template <class VT, class XG, class E>
bool spmv_sorted(const deco_csr* A, XG x, E epi) {
  const deco_csort* S = A->sorted(sizeof(VT) == 8);
  if (!S) return false;
  const VT* val;
  if constexpr (sizeof(VT) == 8) val = S->v64.p;
  else val = S->vf.p;
  const int2* desc = A->row_chunks();
  hipLaunchKernelGGL((k_spmv_stream<VT, XG, E, unsigned int>), dim3(A->nchunks), dim3(stream_block<VT>()), 0,
                     stream(), desc, A->r.p, (const unsigned int*)S->p.p, val, x, epi, (const int*)S->cbase.p,
                     A->chunk_cap, A->chunk_win, A->nnz, A->m);
  DECO_CHECK_LAUNCH();
  count_bytes((double)A->nnz * (sizeof(VT) + 4) + 4.0 * (A->n + 1) + 12.0 * A->nchunks + 8.0 +
              (double)sizeof(typename XG::type) * A->m + (double)A->n * epi.row_bytes());
  return true;
}

// the CSR-stream product on A's row-order arrays with 32-bit columns
template <class VT, class XG, class E>
void spmv_stream_i32(const deco_csr* A, const VT* v, XG x, E epi) {
  if (A->n <= 0) return;
  const int2* desc = A->row_chunks();
  hipLaunchKernelGGL((k_spmv_stream<VT, XG, E>), dim3(A->nchunks), dim3(stream_block<VT>()), 0, stream(), desc,
                     A->r.p, A->c.p, v, x, epi, (const int*)nullptr, A->chunk_cap, A->chunk_win, A->nnz, A->m);
  DECO_CHECK_LAUNCH();
  count_bytes((double)A->nnz * (sizeof(VT) + 4) + 4.0 * (A->n + 1) + 8.0 * (A->nchunks + 1) +
              (double)sizeof(typename XG::type) * A->m + (double)A->n * epi.row_bytes());
}

// whether v is A's own values (FP64, or the current f32 shadow), which its other forms copy
template <class VT>
bool own_values(const deco_csr* A, const VT* v) {
  return v == (const VT*)A->v.p || (A->vf_valid && v == (const VT*)A->vf.p);
}

// the CSR-stream product: on A's column-sorted chunk form where deco_csr::sorted takes it (A's
// own values), else on the row-order arrays with 32-bit columns
template <class VT, class XG, class E>
void spmv_stream_g(const deco_csr* A, const VT* v, XG x, E epi) {
  if (A->n <= 0) return;
  if (own_values(A, v) && spmv_sorted<VT>(A, x, epi)) return;
  spmv_stream_i32(A, v, x, epi);
}

// the product of the AMG operators: A's row-piece form where it has one, else its column-sorted
// chunk form (both for A's own values), else the row-order arrays with 16-bit chunk columns when
// A has them (deco_csr::cols16), else with 32-bit columns
template <class VT, class XG, class E>
void spmv_stream_c16(const deco_csr* A, const VT* v, XG x, E epi) {
  if (A->n <= 0) return;
  if (own_values(A, v) && (spmv_piece<VT>(A, x, epi) || spmv_sorted<VT>(A, x, epi))) return;
  if (!A->cols16()) {
    spmv_stream_i32(A, v, x, epi);
    return;
  }
  const int2* desc = A->row_chunks();
  hipLaunchKernelGGL((k_spmv_stream<VT, XG, E, unsigned short>), dim3(A->nchunks), dim3(stream_block<VT>()), 0,
                     stream(), desc, A->r.p, (const unsigned short*)A->c16.p, v, x, epi, (const int*)A->cbase.p,
                     A->chunk_cap, A->chunk_win, A->nnz, A->m);
  DECO_CHECK_LAUNCH();
  count_bytes((double)A->nnz * (sizeof(VT) + 2) + 4.0 * (A->n + 1) + 12.0 * A->nchunks + 8.0 +
              (double)sizeof(typename XG::type) * A->m + (double)A->n * epi.row_bytes());
}

template <class VT, class XT, class E>
void spmv_stream_f(const deco_csr* A, const VT* v, const XT* x, E epi) {
  spmv_stream_g(A, v, PtrGather<XT>{x}, epi);
}

template <int EPI, class VT, class XT, class YT, class ZT>
void spmv_stream(const deco_csr* A, const VT* v, const XT* x, YT* y, const ZT* z) {
  spmv_stream_f(A, v, x, EpiStd<EPI, YT, ZT>{y, z});
}

// ------------------------------------------------------------------ FP32 block SpMV ----
// Same scheme as the CSR stream kernel on blocks: a workgroup takes the block rows of one chunk,
// every lane loads kBsrUnroll (column, R x C values) blocks and the C x entries of each, the R
// block products go to LDS, and each block row's R sums run over its blocks in column order.
// Per 2 x 2 block: 4 + 16 bytes instead of 4 x (4 + 4) in FP32 CSR.
constexpr int kBsrCap = 1536;
constexpr int kBsrBlock = 512;
constexpr int kBsrUnroll = kBsrCap / kBsrBlock;

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <class VT, int N>
struct BsrVals { VT a[N]; };

// Value layouts of the block shadow. Interleaved: block j's R*C values at [j*R*C, (j+1)*R*C).
// Row-planar: plane rr (stride ps elements) holds row rr of every block, C values per block, so a
// wavefront's load of one row of 64 consecutive blocks is one contiguous range.
template <class VT, int R, int C>
constexpr bool bsr_planar() {
  return !(sizeof(VT) == 4 && (R * C == 2 || R * C == 4));
}
inline bool bsr_planar_rt(int value_size, int R, int C) { return !(value_size == 4 && (R * C == 2 || R * C == 4)); }

template <class VT, int R, int C>
__device__ __forceinline__ BsrVals<VT, R * C> bsr_load(const VT* __restrict__ bv, int j, size_t ps) {
  BsrVals<VT, R * C> o;
  if constexpr (!bsr_planar<VT, R, C>()) {
    if constexpr (R * C == 4) {
      const f32x4 t = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(bv) + j);
      o.a[0] = t.x; o.a[1] = t.y; o.a[2] = t.z; o.a[3] = t.w;
    } else {
      const f32x2 t = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(bv) + j);
      o.a[0] = t.x; o.a[1] = t.y;
    }
  } else if constexpr (sizeof(VT) == 8 && C == 2) {
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      const f64x2 t = __builtin_nontemporal_load(reinterpret_cast<const f64x2*>(bv + rr * ps) + j);
      o.a[rr * 2] = t.x; o.a[rr * 2 + 1] = t.y;
    }
  } else if constexpr (sizeof(VT) == 4 && C == 2) {
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      const f32x2 t = __builtin_nontemporal_load(reinterpret_cast<const f32x2*>(bv + rr * ps) + j);
      o.a[rr * 2] = t.x; o.a[rr * 2 + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int rr = 0; rr < R; ++rr)
#pragma unroll
      for (int q = 0; q < C; ++q) o.a[rr * C + q] = __builtin_nontemporal_load(bv + rr * ps + (size_t)j * C + q);
  }
  return o;
}

template <int C, class XT>
__device__ __forceinline__ void bsr_gather(const XT* __restrict__ x, int col, XT (&xv)[C]) {
  if constexpr (C == 2 && sizeof(XT) == 4) {
    const f32x2 t = *(reinterpret_cast<const f32x2*>(x) + col);
    xv[0] = t.x; xv[1] = t.y;
  } else if constexpr (C == 2 && sizeof(XT) == 8) {
    const f64x2 t = *(reinterpret_cast<const f64x2*>(x) + col);
    xv[0] = t.x; xv[1] = t.y;
  } else {
#pragma unroll
    for (int q = 0; q < C; ++q) xv[q] = x[(size_t)col * C + q];
  }
}

template <class VT, int R, int C, class XT, class E, int BLK = kBsrBlock, int UNR = kBsrUnroll>
__global__ __launch_bounds__(BLK) void k_bsr_stream(const int2* __restrict__ desc, const int* __restrict__ br,
                                                   const int* __restrict__ bc, const VT* __restrict__ bv, size_t ps,
                                                   const XT* __restrict__ x, E epi) {
  using AT = acc_t<VT, XT>;
  constexpr int CAP = BLK * UNR;
  constexpr int LDSB = CAP + 64;              // blocks: a chunk holds <= cap + longest block row - 1
  __shared__ AT prod[LDSB * R];
  const int ch = xcd_block(blockIdx.x, gridDim.x);
  const int2 d0 = desc[ch], d1 = desc[ch + 1];   // (first block row, its first block) of chunks ch, ch+1
  const int r0 = d0.x, r1 = d1.x;
  if (r0 >= r1) return;
  const int j0 = d0.y, cnt = d1.y - j0;
  const int tid = threadIdx.x;
  // this lane's first block row: bounds and epilogue operands load together with the tile
  const int row = r0 + tid;
  int rbeg = 0, rend = 0;
  typename E::Pre pre[R];
  if (row < r1) {
    rbeg = br[row];
    rend = br[row + 1];
#pragma unroll
    for (int rr = 0; rr < R; ++rr) pre[rr] = epi.pre(row * R + rr);
  }
  if (cnt <= LDSB) {
    if (cnt > 0) {   // branch-free tile: clamped indices, every gather in flight before the products
      const int jl = j0 + cnt - 1;
      int cc[UNR];
      BsrVals<VT, R * C> vv[UNR];
      XT xv[UNR][C];
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int j = min(j0 + tid + u * BLK, jl);
        cc[u] = __builtin_nontemporal_load(bc + j);
        vv[u] = bsr_load<VT, R, C>(bv, j, ps);
      }
#pragma unroll
      for (int u = 0; u < UNR; ++u) bsr_gather<C>(x, cc[u], xv[u]);
#pragma unroll
      for (int u = 0; u < UNR; ++u) {
        const int t = tid + u * BLK;
        if (t < cnt) {
#pragma unroll
          for (int rr = 0; rr < R; ++rr) {
            AT p = (AT)vv[u].a[rr * C] * (AT)xv[u][0];
#pragma unroll
            for (int q = 1; q < C; ++q) p += (AT)vv[u].a[rr * C + q] * (AT)xv[u][q];
            prod[t * R + rr] = p;
          }
        }
      }
    }
    for (int t = tid + CAP; t < cnt; t += BLK) {
      const BsrVals<VT, R * C> w = bsr_load<VT, R, C>(bv, j0 + t, ps);
      XT xv[C];
      bsr_gather<C>(x, bc[j0 + t], xv);
#pragma unroll
      for (int rr = 0; rr < R; ++rr) {
        AT p = (AT)w.a[rr * C] * (AT)xv[0];
#pragma unroll
        for (int q = 1; q < C; ++q) p += (AT)w.a[rr * C + q] * (AT)xv[q];
        prod[t * R + rr] = p;
      }
    }
    __syncthreads();
    if (row < r1) {
#pragma unroll
      for (int rr = 0; rr < R; ++rr) {
        AT acc = AT(0);
        for (int k = rbeg - j0; k < rend - j0; ++k) acc += prod[k * R + rr];
        epi(row * R + rr, acc, pre[rr]);
      }
      for (int rw = row + BLK; rw < r1; rw += BLK) {
        const int b = br[rw] - j0, e = br[rw + 1] - j0;
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
          AT acc = AT(0);
          for (int k = b; k < e; ++k) acc += prod[k * R + rr];
          epi(rw * R + rr, acc, epi.pre(rw * R + rr));
        }
      }
    }
  } else {
    __shared__ AT red[BLK / 64];
    for (int rw = r0; rw < r1; ++rw) {
#pragma unroll
      for (int rr = 0; rr < R; ++rr) {
        AT acc = AT(0);
        for (int j = br[rw] + tid; j < br[rw + 1]; j += BLK) {
          XT xv[C];
          bsr_gather<C>(x, bc[j], xv);
          const BsrVals<VT, R * C> w = bsr_load<VT, R, C>(bv, j, ps);
#pragma unroll
          for (int q = 0; q < C; ++q) acc += (AT)w.a[rr * C + q] * (AT)xv[q];
        }
        acc = block_sum<BLK>(acc, red);
        if (tid == 0) epi(rw * R + rr, acc, epi.pre(rw * R + rr));
      }
    }
  }
}

// ------------------------------------------------------------------ sliced (SELL) products --
// One lane per block row, 64 block rows per slice (a wavefront), 4 slices per workgroup. Step k
// of a wavefront loads entry k of 64 consecutive block rows: contiguous columns and values, and
// for stencil-like patterns x gathers that fall into a few cache lines, with no LDS staging and
// no barrier. Each row sums its blocks in column order, the C products of a block paired first
// (the chunked kernel's order). Values: C == 2, one plane per block row holding the pair
// (one 8- or 16-byte load); otherwise one plane per block element.
constexpr int kSellRows = 64;

__host__ __device__ inline size_t sell_vidx(int C, size_t T, int rr, int q, size_t e) {
  return C == 2 ? (size_t)rr * 2 * T + 2 * e + q : (size_t)(rr * C + q) * T + e;
}

template <class VT, int R, int C>
__device__ __forceinline__ BsrVals<VT, R * C> sell_load(const VT* __restrict__ bv, size_t e, size_t T) {
  BsrVals<VT, R * C> o;
  if constexpr (C == 2) {
    using V2 = typename std::conditional<sizeof(VT) == 8, f64x2, f32x2>::type;
#pragma unroll
    for (int rr = 0; rr < R; ++rr) {
      const V2 t = __builtin_nontemporal_load(reinterpret_cast<const V2*>(bv + (size_t)rr * 2 * T) + e);
      o.a[rr * 2] = t.x; o.a[rr * 2 + 1] = t.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < R * C; ++q) o.a[q] = __builtin_nontemporal_load(bv + (size_t)q * T + e);
  }
  return o;
}

// column tag of the uniform-offset sliced layout (deco_bsr::dia): no stored columns
struct DiaCols {};
constexpr int kDiaMax = 16;   // largest offset set of the uniform-offset layout

// block column of sliced entry e: stored (int) or as an offset from the lane's block row base
// (short; base 0 for the lanes past the last block row, whose entries hold offset 0)
template <class CT>
__device__ __forceinline__ int sell_col(const CT* __restrict__ bc, size_t e, int base) {
  if constexpr (std::is_same<CT, short>::value) return base + (int)__builtin_nontemporal_load(bc + e);
  else return __builtin_nontemporal_load(bc + e);
}

// Each lane sums its block row's entries in column order, KU at a time: a batch's loads are all
// in flight before its gathers. The launcher picks KU from the value type and the layout's mean
// width: 8 for scalar FP32 entries; 7 for other blocks of up to 4 values when the slices average
// more than 5 blocks per row (the 7-block cell rows of A, the 13-block rows of an FSAI factor on
// the A^2 pattern: one or two rounds); 4 otherwise (the FSAI factors' 4-block rows on A's
// pattern, 3 x 3 blocks). FP32 batches and batches of 7 are clamped: the last one of a row
// re-reads the row's last entry with zero values, so a slice of width w takes ceil(w/KU)
// dependent rounds. FP64 batches of 4 keep a remainder loop of single entries (measured: clamped
// or larger batches cost those products more in registers than they gain).
// CT DiaCols: the uniform-offset layout, entry k of block row I at block column I + doff[k]
// (clamped into the mb block columns; its values are zero where the row has no such block).
// SPW slices per wavefront (uniform-offset layout only): the wave issues the loads of all SPW
// slices' batches together, SPW times the bytes in flight per wave for the short FP32 rows. A
// slice past the end, or one of width 0, reads the first slice's first entries with zero values.
// Each row's sum takes the same terms in the same order as with one slice per wave.
template <class VT, int R, int C, class XT, class E, class CT, int KU, int SPW = 1>
__global__ __launch_bounds__(256) void k_sell_stream(int nb, int nslices, const int* __restrict__ soff,
                                                    const int* __restrict__ swid, const CT* __restrict__ bc,
                                                    const VT* __restrict__ bv, size_t T, const XT* __restrict__ x,
                                                    E epi, const int* __restrict__ doff, int mb) {
  constexpr bool DIA = std::is_same<CT, DiaCols>::value;
  static_assert(SPW == 1 || DIA, "several slices per wave need the uniform-offset layout");
  using AT = acc_t<VT, XT>;
  const int sl0 = (xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6)) * SPW;
  if (sl0 >= nslices) return;
  const int lane = threadIdx.x & 63;
  int I[SPW], base[SPW], o[SPW], w[SPW];
  bool live[SPW];
  typename E::Pre pre[SPW][R];
  AT acc[SPW][R];
  int wmax = 0;
#pragma unroll
  for (int s = 0; s < SPW; ++s) {
    const int sl = sl0 + s;
    const bool in = sl < nslices;
    I[s] = sl * kSellRows + lane;
    live[s] = in && I[s] < nb;
    base[s] = live[s] ? I[s] : 0;
    o[s] = in ? soff[sl] : 0;
    w[s] = in ? swid[sl] : 0;
    wmax = max(wmax, w[s]);
    if (live[s]) {
#pragma unroll
      for (int rr = 0; rr < R; ++rr) pre[s][rr] = epi.pre(I[s] * R + rr);
    }
#pragma unroll
    for (int rr = 0; rr < R; ++rr) acc[s][rr] = AT(0);
  }
  if constexpr (SPW > 1) {
#pragma unroll
    for (int s = 1; s < SPW; ++s)
      if (w[s] == 0) o[s] = o[0];     // slice 0 is in range; width 0 only with an all-empty matrix
  }
  constexpr bool clamped = (sizeof(VT) == 4) || KU == 7 || DIA;
  static_assert(SPW == 1 || clamped, "several slices per wave need clamped batches");
  int k = 0;
  for (; clamped ? k < wmax : k + KU <= w[0]; k += KU) {
    int cc[SPW][KU];
    BsrVals<VT, R * C> vv[SPW][KU];
    XT xv[SPW][KU][C];
#pragma unroll
    for (int u = 0; u < KU; ++u)
#pragma unroll
      for (int s = 0; s < SPW; ++s) {
        const int kk = max(min(k + u, w[s] - 1), 0);
        const size_t e = (size_t)o[s] + (size_t)kk * kSellRows + lane;
        if constexpr (DIA) cc[s][u] = min(max(base[s] + doff[kk], 0), mb - 1);
        else cc[s][u] = sell_col(bc, e, base[s]);
        vv[s][u] = sell_load<VT, R, C>(bv, e, T);
        if (clamped && k + u >= w[s]) {
#pragma unroll
          for (int q = 0; q < R * C; ++q) vv[s][u].a[q] = VT(0);
        }
      }
#pragma unroll
    for (int u = 0; u < KU; ++u)
#pragma unroll
      for (int s = 0; s < SPW; ++s) bsr_gather<C>(x, cc[s][u], xv[s][u]);
#pragma unroll
    for (int u = 0; u < KU; ++u)
#pragma unroll
      for (int s = 0; s < SPW; ++s)
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
          AT p = (AT)vv[s][u].a[rr * C] * (AT)xv[s][u][0];
#pragma unroll
          for (int q = 1; q < C; ++q) p += (AT)vv[s][u].a[rr * C + q] * (AT)xv[s][u][q];
          acc[s][rr] += p;
        }
  }
  if constexpr (!clamped) {
    for (; k < w[0]; ++k) {   // remainder of the unclamped batches (one slice per wave)
      const size_t e = (size_t)o[0] + (size_t)k * kSellRows + lane;
      const BsrVals<VT, R * C> vv = sell_load<VT, R, C>(bv, e, T);
      XT xv[C];
      const int col = sell_col(bc, e, base[0]);
      bsr_gather<C>(x, col, xv);
#pragma unroll
      for (int rr = 0; rr < R; ++rr) {
        AT p = (AT)vv.a[rr * C] * (AT)xv[0];
#pragma unroll
        for (int q = 1; q < C; ++q) p += (AT)vv.a[rr * C + q] * (AT)xv[q];
        acc[0][rr] += p;
      }
    }
  }
#pragma unroll
  for (int s = 0; s < SPW; ++s)
    if (live[s]) {
#pragma unroll
      for (int rr = 0; rr < R; ++rr) epi(I[s] * R + rr, acc[s][rr], pre[s][rr]);
    }
}

// FP32 uniform-offset products with two slices per wavefront: 0 never, 1 when a batch holds at
// most 8 values per lane (default: the scalar rows; measured per kernel, the 2 x 2 rows of 4 blocks
// and the 1 x 2 rows of 7 run slower with two slices), 2 always
extern int g_sell_pairs;

template <class VT, class XT, class E>
void bsr_stream_v(const deco_bsr* B, const VT* bv, const XT* x, E epi) {
  if (B->nb <= 0) return;
  if (B->sell) {
    const unsigned grid = (unsigned)((B->nslices + 3) / 4);
    const bool wide = B->stot > (size_t)5 * kSellRows * B->nslices;     // mean width over 5 blocks
    const bool narrow = B->stot <= (size_t)4 * kSellRows * B->nslices;  // at most 4 on average
#define DECO_SELL_LAUNCH_KU(RR, CC, KU)                                                                          \
  do {                                                                                                           \
    if (B->dia && sizeof(VT) == 4 && (g_sell_pairs == 2 || (g_sell_pairs == 1 && KU * RR * CC <= 8)))          \
      hipLaunchKernelGGL((k_sell_stream<VT, RR, CC, XT, E, DiaCols, KU, 2>), dim3((B->nslices + 7) / 8), dim3(256), \
                         0, stream(), B->nb, B->nslices, B->soff.p, B->swid.p, (const DiaCols*)nullptr, bv,      \
                         B->stot, x, epi, (const int*)B->doff.p, B->mb);                                         \
    else if (B->dia)                                                                                             \
      hipLaunchKernelGGL((k_sell_stream<VT, RR, CC, XT, E, DiaCols, KU>), dim3(grid), dim3(256), 0, stream(),     \
                         B->nb, B->nslices, B->soff.p, B->swid.p, (const DiaCols*)nullptr, bv, B->stot, x, epi,   \
                         (const int*)B->doff.p, B->mb);                                                          \
    else if (B->d16)                                                                                             \
      hipLaunchKernelGGL((k_sell_stream<VT, RR, CC, XT, E, short, KU>), dim3(grid), dim3(256), 0, stream(), B->nb, \
                         B->nslices, B->soff.p, B->swid.p, (const short*)B->bcd.p, bv, B->stot, x, epi,           \
                         (const int*)nullptr, B->mb);                                                            \
    else                                                                                                         \
      hipLaunchKernelGGL((k_sell_stream<VT, RR, CC, XT, E, int, KU>), dim3(grid), dim3(256), 0, stream(), B->nb,   \
                         B->nslices, B->soff.p, B->swid.p, (const int*)B->bc.p, bv, B->stot, x, epi,              \
                         (const int*)nullptr, B->mb);                                                            \
  } while (0)
#define DECO_SELL_LAUNCH(RR, CC)                                                        \
  do {                                                                                  \
    if constexpr (RR * CC * sizeof(VT) <= 4) {                                          \
      if (narrow) DECO_SELL_LAUNCH_KU(RR, CC, 4);                                       \
      else DECO_SELL_LAUNCH_KU(RR, CC, 8);                                              \
    }                                                                                   \
    else if constexpr (RR * CC <= 4) {                                                  \
      if (wide) DECO_SELL_LAUNCH_KU(RR, CC, 7);                                         \
      else DECO_SELL_LAUNCH_KU(RR, CC, 4);                                              \
    } else DECO_SELL_LAUNCH_KU(RR, CC, 4);                                              \
  } while (0)
    if (B->R == 2 && B->C == 2) DECO_SELL_LAUNCH(2, 2);
    else if (B->R == 1 && B->C == 1) DECO_SELL_LAUNCH(1, 1);
    else if (B->R == 1 && B->C == 2) DECO_SELL_LAUNCH(1, 2);
    else if (B->R == 3 && B->C == 3) DECO_SELL_LAUNCH(3, 3);
    else if (B->R == 1 && B->C == 3) DECO_SELL_LAUNCH(1, 3);
    else if (B->R == 2 && B->C == 1) DECO_SELL_LAUNCH(2, 1);
    else if (B->R == 3 && B->C == 1) DECO_SELL_LAUNCH(3, 1);
    else fail(DECO_ERR_UNSUPPORTED, "block shape %d x %d", B->R, B->C);
#undef DECO_SELL_LAUNCH
#undef DECO_SELL_LAUNCH_KU
    DECO_CHECK_LAUNCH();
    count_bytes((double)B->nnzb * (B->R * B->C * sizeof(VT) + (B->dia ? 0 : B->d16 ? 2 : 4)) + 8.0 * B->nslices +
                (double)sizeof(XT) * B->mb * B->C + (double)B->nb * B->R * epi.row_bytes());
    return;
  }
#define DECO_BSR_LAUNCH(RR, CC)                                                                                  \
  hipLaunchKernelGGL((k_bsr_stream<VT, RR, CC, XT, E>), dim3(B->nchunks), dim3(kBsrBlock), 0, stream(), B->rb.p, \
                     B->br.p, B->bc.p, bv, (size_t)B->nnzb * CC, x, epi)
  if (B->R == 2 && B->C == 2) DECO_BSR_LAUNCH(2, 2);
  else if (B->R == 1 && B->C == 1) DECO_BSR_LAUNCH(1, 1);
  else if (B->R == 1 && B->C == 2) DECO_BSR_LAUNCH(1, 2);
  else if (B->R == 3 && B->C == 3) DECO_BSR_LAUNCH(3, 3);
  else if (B->R == 1 && B->C == 3) DECO_BSR_LAUNCH(1, 3);
  else if (B->R == 2 && B->C == 1) DECO_BSR_LAUNCH(2, 1);
  else if (B->R == 3 && B->C == 1) DECO_BSR_LAUNCH(3, 1);
  else fail(DECO_ERR_UNSUPPORTED, "block shape %d x %d", B->R, B->C);
#undef DECO_BSR_LAUNCH
  DECO_CHECK_LAUNCH();
  count_bytes((double)B->nnzb * (B->R * B->C * sizeof(VT) + 4) + 4.0 * (B->nb + 1) + 8.0 * (B->nchunks + 1) +
              (double)sizeof(XT) * B->mb * B->C + (double)B->nb * B->R * epi.row_bytes());
}

// Scalar product y = A x (epilogue e) with VT values: through A's sliced 1 x 1 shadow when its
// padding is within the threshold (built on first use), else the CSR-stream kernel on 16-bit
// chunk columns where they fit. Each row sums in column order in the sliced layout.
template <class VT, class XT, class E>
void spmv_auto(const deco_csr* A, const XT* x, E epi) {
  const deco_bsr* B = !A->sell_scalar() ? nullptr
                      : std::is_same<VT, double>::value ? A->block64(1, 1) : A->block32(1, 1);
  if (B && B->sell) {
    if constexpr (std::is_same<VT, double>::value) bsr_stream_v<double>(B, B->bv64.p, x, epi);
    else bsr_stream_v<float>(B, B->bv.p, x, epi);
    return;
  }
  if constexpr (std::is_same<VT, double>::value) spmv_stream_c16(A, A->v.p, PtrGather<XT>{x}, epi);
  else spmv_stream_c16(A, const_cast<deco_csr*>(A)->values_f32(), PtrGather<XT>{x}, epi);
}

// product with the FP32 block values
template <class XT, class E>
void bsr_stream_f(const deco_bsr* B, const XT* x, E epi) {
  bsr_stream_v<float>(B, B->bv.p, x, epi);
}

// product with the FP64 block values (block64() filled them)
template <class E>
void bsr_stream_d(const deco_bsr* B, const double* x, E epi) {
  if (!B->has64) fail(DECO_ERR_STATE, "FP64 block values missing");
  bsr_stream_v<double>(B, B->bv64.p, x, epi);
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
// columns stride*j + off of A (all rows), as an n x (m/stride) CSR with column j
deco_csr* csr_col_subset(const deco_csr* A, int stride, int off);
deco_csr* csr_clone(const deco_csr* A);
deco_csr* csr_new(int n, int m);

}  // namespace deco
