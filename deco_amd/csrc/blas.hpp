// BLAS-1 kernels with deterministic reductions (core/alg.hpp:185-396 restated for the GPU).
//
// A reduction over n elements runs on a fixed grid red_grid(n): block partial sums are
// written to `part`, and any later kernel that needs the scalar re-reduces the partials in a
// fixed order (reduce_partials), so no atomics are used and results are bit-reproducible.
#pragma once
#include "common.hpp"

namespace deco {

// Device scratch owned by a solver/handle: partial sums and device scalars.
struct RedScratch {
  DBuf<double> part;  // 4 * kRedMaxGrid partials (four independent slots)
  DBuf<double> scal;  // device scalars
  RedScratch() : part(4 * kRedMaxGrid), scal(256) {}
  double* slot(int k) { return part.p + (size_t)k * kRedMaxGrid; }
};

// part[block] = sum_i x[i]*y[i] over the block's grid-stride share
template <class T>
__global__ __launch_bounds__(kRedBlock) void k_dot_part(int n, const T* __restrict__ x, const T* __restrict__ y,
                                                       double* __restrict__ part) {
  __shared__ double sm[kRedBlock / 64];
  double s = 0.0;
  for (int i = blockIdx.x * kRedBlock + threadIdx.x; i < n; i += gridDim.x * kRedBlock)
    s += (double)x[i] * (double)y[i];
  s = block_sum<kRedBlock>(s, sm);
  if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[0] = reduce(part) (optionally its square root)
__global__ void k_finish(const double* __restrict__ part, int count, double* __restrict__ out, int take_sqrt);

// y = b*y + a*x
__global__ void k_axpby(int n, double a, const double* __restrict__ x, double b, double* __restrict__ y);

// Host-side helpers (queue on the library stream).
void dot_partials(int n, const double* x, const double* y, double* part);
// reduce partials to a device scalar (dst) without host sync
void finish(const double* part, int count, double* dst, bool take_sqrt);
// host scalar results (synchronise)
double dot_host(int n, const double* x, const double* y, RedScratch& rs);
double nrm2_host(int n, const double* x, RedScratch& rs);

}  // namespace deco
