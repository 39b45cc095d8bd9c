// Preconditioner interface shared by CPR, FSAI, Jacobi, AMG and identity.
// apply(r, x): x = M^{-1} r on device buffers, queued on the library stream
// (CPRBase::apply, precond/CPR.hpp:3-8; the Mglobal/Mlocal apply members).
#pragma once
#include "csr.hpp"

struct deco_prec {
  enum Kind { IDENTITY = 0, JACOBI = 1, FSAI = 2, AMG = 3, CPR = 4 };
  int kind = IDENTITY;
  int n = 0;
  int precision = DECO_PREC_FP64;
  int layout = DECO_LAYOUT_BLOCK;
  virtual ~deco_prec() = default;
  virtual void apply(const double* r, double* x) = 0;
  virtual void apply_t(const double* r, double* x) {
    (void)r; (void)x;
    deco::fail(DECO_ERR_UNSUPPORTED, "transposed apply not available for this preconditioner");
  }
  // FP32 entry points used inside CPR (input may be f64 or f32, output f32)
  virtual void apply32(const float* r, float* x) {
    (void)r; (void)x;
    deco::fail(DECO_ERR_UNSUPPORTED, "FP32 apply not available for this preconditioner");
  }
  virtual void apply32_t(const float* r, float* x) {
    (void)r; (void)x;
    deco::fail(DECO_ERR_UNSUPPORTED, "FP32 transposed apply not available for this preconditioner");
  }
  // x = M^{-1} r (M^{-T} r) written in FP32: the flexible GMRES directions of an FP32
  // preconditioner, whose results are FP32 values already. Default: the FP64-output apply and a
  // conversion; CPR and FSAI write FP32 directly.
  virtual void apply_f32out(const double* r, float* x);
  virtual void apply_t_f32out(const double* r, float* x);
  virtual void set_precision(int p) { precision = p; }
  virtual void set_layout(int l) { layout = l; }
  // products on cell-block shadows (else on the CSR arrays), FP32 and FP64 alike
  bool blocks32() const { return layout == DECO_LAYOUT_BLOCK; }
  bool blocks64() const { return layout == DECO_LAYOUT_BLOCK; }
  // Pressure stage of CPR with two steps folded in (AMG only; false: not available, nothing
  // done). d0 receives the diagonal of the first zero-guess sweep x = b/d, which the caller's
  // residual epilogue then applies itself; stage() runs the rest of the cycle on (b, x) and adds
  // the result to add_to[stride*i] from its last kernel (x is work space: it does not receive
  // the result).
  virtual bool stage_diag(const double** d0) { (void)d0; return false; }
  virtual bool stage_diag32(const float** d0) { (void)d0; return false; }
  virtual void stage(const double* b, double* x, double* add_to, int stride) {
    (void)b; (void)x; (void)add_to; (void)stride;
    deco::fail(DECO_ERR_STATE, "fused pressure stage not available");
  }
  virtual void stage32(const float* b, float* x, float* add_to, int stride) {
    (void)b; (void)x; (void)add_to; (void)stride;
    deco::fail(DECO_ERR_STATE, "fused pressure stage not available");
  }
};

namespace deco {

// elementwise helpers used across preconditioners
template <class A, class B>
__global__ void k_convert(int64_t n, const A* __restrict__ a, B* __restrict__ b) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = (B)a[i];
}
template <class A, class B>
void convert(const A* a, B* b, int64_t n) {
  if (n <= 0) return;
  hipLaunchKernelGGL((k_convert<A, B>), dim3(grid_for(n, 256)), dim3(256), 0, stream(), n, a, b);
  DECO_CHECK_LAUNCH();
  count_bytes((double)n * (sizeof(A) + sizeof(B)));
}

struct Identity : deco_prec {
  explicit Identity(int n_) { kind = IDENTITY; n = n_; }
  void apply(const double* r, double* x) override { copy_d2d(x, r, n); }
  void apply_t(const double* r, double* x) override { copy_d2d(x, r, n); }
  void apply32(const float* r, float* x) override { copy_d2d(x, r, n); }
  void apply32_t(const float* r, float* x) override { copy_d2d(x, r, n); }
};

// Point / block Jacobi (precond/Jacobi/Jacobi.hpp:15-68, Jacobi.tpp:3-69).
struct Jacobi : deco_prec {
  int bs = 1;
  DBuf<double> J;     // bs==1: 1/a_ii ; else per cell bs*bs row-major inverse
  DBuf<float> Jf;
  Jacobi(const deco_csr* A, int bs_);
  void apply(const double* r, double* x) override;
  void apply_t(const double* r, double* x) override;
  void apply32(const float* r, float* x) override;
  void apply32_t(const float* r, float* x) override;
};

inline void convert_apply(deco_prec* M, bool tr, const double* r, float* x) {
  DBuf<double> t(M->n);
  if (tr) M->apply_t(r, t.p);
  else M->apply(r, t.p);
  convert((const double*)t.p, x, M->n);
}

deco_prec* make_fsai(const deco_cpr_params& p, const deco_csr* A);
deco_prec* make_amg(const deco_cpr_params& p, const deco_csr* Ap);
deco_prec* make_cpr(const deco_cpr_params& p, deco_csr* A, double* b);

}  // namespace deco

inline void deco_prec::apply_f32out(const double* r, float* x) { deco::convert_apply(this, false, r, x); }
inline void deco_prec::apply_t_f32out(const double* r, float* x) { deco::convert_apply(this, true, r, x); }
