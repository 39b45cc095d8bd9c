// Drop-in C++ facade of deco's solve path over libdeco_hip.so.
//
// Keeps the reference's class templates and signatures so a cpr.cpp-style driver compiles
// unchanged apart from its includes:
//   CSR<R,C,V>                          core/CSR.hpp:3-75   (host arrays + a device mirror)
//   CPRBase<V>, CPR<R,C,V,Mglo,Mloc>    precond/CPR.hpp:3-14,42
//   FSAI<R,C,V>, Jacobi<R,C,V>          precond/FSAI/FSAI.hpp:14, precond/Jacobi/Jacobi.hpp:15
//   hypre_AMG<R,C,V>                    precond/hypre_AMG/hypre_AMG.hpp:16 (the device AMG)
//   SOL<R,C,V>                          solver/SOL.hpp:19-49
// All numerical work runs on the GPU through the C ABI. Host pointers are accepted where the
// reference takes them and copied to/from the device. Numerical failures exit like the
// reference (CPR.hpp:416-419, rgmres.tpp:96-99); bad types throw std::invalid_argument.
#pragma once
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "deco_hip.h"
#include "ptree.hpp"

#if defined(DECO_USE_BOOST_PTREE) && __has_include(<boost/property_tree/ptree.hpp>)
#include <boost/property_tree/ptree.hpp>
using deco_ptree = boost::property_tree::ptree;
#else
using deco_ptree = deco::ptree;
#endif

namespace deco_facade {

inline void check(int code, const char* what) {
  if (code == DECO_OK) return;
  std::string msg = std::string(what) + ": " + deco_last_error();
  if (code == DECO_ERR_SINGULAR || code == DECO_ERR_BREAKDOWN) {
    std::fprintf(stderr, "%s\n", msg.c_str());
    std::exit(EXIT_FAILURE);
  }
  if (code == DECO_ERR_ARG || code == DECO_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// device scratch buffer of doubles
struct DevBuf {
  double* p = nullptr;
  size_t n = 0;
  void ensure(size_t k) {
    if (k <= n) return;
    if (p) deco_free(p);
    void* q = nullptr;
    check(deco_malloc(k * sizeof(double) + 8, &q), "deco_malloc");
    p = static_cast<double*>(q);
    n = k;
  }
  ~DevBuf() { if (p) deco_free(p); }
};

inline deco_cpr_params params_from(const deco_ptree& prec) {
  deco_cpr_params p{};
  p.block_size = prec.get<int>("block_size");
  p.pressure_id = prec.get<int>("pressure_id", 0);
  p.seq = prec.get<int>("seq", 3);
  const std::string w = prec.get<std::string>("weights", std::string("quasiimpes_GE"));
  if (w == "none") p.weights = DECO_W_NONE;
  else if (w == "quasiimpes") p.weights = DECO_W_QUASIIMPES;
  else if (w == "quasiimpes_GE") p.weights = DECO_W_QUASIIMPES_GE;
  else throw std::invalid_argument("Unknown weights type " + w);
  p.use_drs = prec.get<int>("use_drs", 0);
  p.drs_e_dd = prec.get<double>("drs.e_dd", 0.2);
  p.drs_e_ps = prec.get<double>("drs.e_ps", 0.0);
  p.cpd = prec.get<int>("CPD", 1);
  p.tol_zero = prec.get<double>("tol_zero", 0.0);
  const std::string gt = prec.get<std::string>("prec_glo.type", std::string("FSAI"));
  if (gt == "FSAI") p.glo_type = DECO_GLO_FSAI;
  else if (gt == "Jacobi") p.glo_type = DECO_GLO_JACOBI;
  else throw std::invalid_argument("Unknown global prec.prec_glo.type: " + gt);
  p.glo_tol_rel = prec.get<double>("prec_glo.tol_rel", 0.0);
  p.glo_tol_abs = prec.get<double>("prec_glo.tol_abs", 0.0);
  p.glo_power = prec.get<int>("prec_glo.power", 1);
  p.glo_block_size = prec.get<int>("prec_glo.block_size", 1);
  p.loc_tol_rel = prec.get<double>("prec_loc.tol_rel", 0.0);
  p.amg_coarse_type = prec.get<int>("prec_loc.amg_coarse_type", 10);
  p.amg_relax_type = prec.get<int>("prec_loc.amg_relax_type", -1);
  p.amg_agg_levs = prec.get<int>("prec_loc.amg_agg_levs", 0);
  p.amg_agg_path = prec.get<int>("prec_loc.amg_agg_path", 1);
  p.amg_A_drop_tol = prec.get<double>("prec_loc.amg_A_drop_tol", 0.0);
  p.amg_prol_trunc_fac = prec.get<double>("prec_loc.amg_prol_trunc_fac", 0.0);
  p.amg_SoC_alpha = prec.get<double>("prec_loc.amg_SoC_alpha", 0.25);
  p.amg_max_coarse = prec.get<int>("prec_loc.amg_max_coarse", 300);
  p.amg_max_levels = prec.get<int>("prec_loc.amg_max_levels", 25);
  p.amg_P_max_elmts = prec.get<int>("prec_loc.amg_P_max_elmts", 4);
  p.amg_max_row_sum = prec.get<double>("prec_loc.amg_max_row_sum", 0.9);
  p.amg_relax_weight = prec.get<double>("prec_loc.amg_relax_weight", 0.0);
  p.precision = prec.get<std::string>("precision", std::string("fp64")) == "fp32" ? DECO_PREC_FP32 : DECO_PREC_FP64;
  const std::string lay = prec.get<std::string>("layout", std::string("block"));
  if (lay == "block") p.layout = DECO_LAYOUT_BLOCK;
  else if (lay == "csr") p.layout = DECO_LAYOUT_CSR;
  else throw std::invalid_argument("Unknown layout: " + lay);
  return p;
}

// A preconditioner callable that carries its device handle, so SOL can run on the device.
template <typename V>
struct DevicePrecond {
  deco_prec* h = nullptr;
  int n = 0;
  std::shared_ptr<DevBuf> in, out;
  void run(const V* r, V* x) const {
    in->ensure(n);
    out->ensure(n);
    check(deco_memcpy_h2d(in->p, r, sizeof(V) * n), "h2d");
    check(deco_prec_apply(h, in->p, out->p), "apply");
    check(deco_memcpy_d2h(x, out->p, sizeof(V) * n), "d2h");
  }
  void operator()(V* r, V*& x) const { run(r, x); }
  void operator()(const V* r, V*& x) const { run(r, x); }
};

template <typename V>
DevicePrecond<V> make_callable(deco_prec* h, int n) {
  DevicePrecond<V> f;
  f.h = h;
  f.n = n;
  f.in = std::make_shared<DevBuf>();
  f.out = std::make_shared<DevBuf>();
  return f;
}

}  // namespace deco_facade

// ------------------------------------------------------------------------------------- CSR
template <typename R = int, typename C = int, typename V = double>
class CSR : public std::enable_shared_from_this<CSR<R, C, V>> {
  static_assert(std::is_same<R, int>::value && std::is_same<C, int>::value && std::is_same<V, double>::value,
                "the device path stores int32 indices and FP64 values");

 public:
  bool owns_data = 1;
  C n = 0, m = 0;
  R nz = 0;
  R* r = nullptr;
  C* c = nullptr;
  V* v = nullptr;
  int np = 1, mid = 0;
  C n_glo = 0, shift_nr = 0;
  C* displs_nr = nullptr;
  C* counts_nr = nullptr;
  deco_csr* dev = nullptr;   // device mirror (created from the host arrays on first use)

  CSR() = default;
  CSR(C n_, C m_, R nz_) : n(n_), m(m_), nz(nz_) {
    if (nz_ != 0) {
      r = (R*)std::malloc(sizeof(R) * (n_ + 1));
      c = (C*)std::malloc(sizeof(C) * nz_);
      v = (V*)std::malloc(sizeof(V) * nz_);
    }
  }
  CSR(C n_, C m_, R nz_, R* r_, C* c_, V* v_) : n(n_), m(m_), nz(nz_), r(r_), c(c_), v(v_) {}

  deco_csr* device() {
    if (!dev) deco_facade::check(deco_csr_create(n, m, nz, r, c, v, &dev), "deco_csr_create");
    return dev;
  }
  // copy the device mirror back into the host arrays (after in-place device operations)
  void pull() {
    int nn, mm, nnz;
    deco_facade::check(deco_csr_info(dev, &nn, &mm, &nnz), "deco_csr_info");
    if (nnz != nz) {
      c = (C*)std::realloc(c, sizeof(C) * (nnz ? nnz : 1));
      v = (V*)std::realloc(v, sizeof(V) * (nnz ? nnz : 1));
      nz = nnz;
    }
    deco_facade::check(deco_csr_download(dev, r, c, v), "deco_csr_download");
  }
  void setup_spmv_v2() {}   // one rank: no halo (core/CSR.hpp:97-130)

  // CSR::apply(x, b, bs) (core/CSR.hpp:132-136) with host vectors
  void apply(const V* x, V*& b, const int bs = 1) {
    xbuf.ensure(m);
    ybuf.ensure(n);
    deco_facade::check(deco_memcpy_h2d(xbuf.p, x, sizeof(V) * m), "h2d");
    deco_facade::check(deco_spmv(device(), xbuf.p, ybuf.p, bs), "deco_spmv");
    deco_facade::check(deco_memcpy_d2h(b, ybuf.p, sizeof(V) * (n / bs)), "d2h");
  }
  std::shared_ptr<CSR> copy() {
    auto A = std::make_shared<CSR>(n, m, nz);
    std::memcpy(A->r, r, sizeof(R) * (n + 1));
    std::memcpy(A->c, c, sizeof(C) * nz);
    std::memcpy(A->v, v, sizeof(V) * nz);
    return A;
  }
  void filter(const double tol) {
    deco_facade::check(deco_csr_filter(device(), tol), "deco_csr_filter");
    pull();
  }
  void filter_proper(double tol) {
    deco_facade::check(deco_csr_filter_proper(device(), tol), "deco_csr_filter_proper");
    pull();
  }
  std::shared_ptr<CSR> local_matrix(const int bs = 3, const int p = 0, const V* weights = nullptr) {
    deco_csr* L = nullptr;
    deco_facade::DevBuf w;
    if (weights) {
      w.ensure(n);
      deco_facade::check(deco_memcpy_h2d(w.p, weights, sizeof(V) * n), "h2d");
    }
    deco_facade::check(deco_csr_local_matrix(device(), bs, p, weights ? w.p : nullptr, &L), "local_matrix");
    return from_device(L);
  }
  void trans_mat_full() {
    deco_csr* T = nullptr;
    deco_facade::check(deco_csr_transpose(device(), &T), "deco_csr_transpose");
    deco_csr_destroy(dev);
    dev = T;
    std::swap(n, m);
    r = (R*)std::realloc(r, sizeof(R) * (n + 1));
    pull();
  }
  static std::shared_ptr<CSR> from_device(deco_csr* D) {
    int nn, mm, nnz;
    deco_facade::check(deco_csr_info(D, &nn, &mm, &nnz), "deco_csr_info");
    auto A = std::make_shared<CSR>(nn, mm, nnz);
    if (!A->r) A->r = (R*)std::malloc(sizeof(R) * (nn + 1));
    if (!A->c) { A->c = (C*)std::malloc(sizeof(C)); A->v = (V*)std::malloc(sizeof(V)); }
    A->dev = D;
    deco_facade::check(deco_csr_download(D, A->r, A->c, A->v), "deco_csr_download");
    return A;
  }
  ~CSR() {
    if (owns_data) { std::free(r); std::free(c); std::free(v); }
    std::free(displs_nr);
    std::free(counts_nr);
    if (dev) deco_csr_destroy(dev);
  }

 private:
  deco_facade::DevBuf xbuf, ybuf;
};

// --------------------------------------------------------------------- preconditioners
template <typename V>
class CPRBase {
 public:
  virtual ~CPRBase() = default;
  std::function<void(V* __restrict__, V*& __restrict__)> apply;
  deco_prec* handle = nullptr;
};

// Mglobal / Mlocal classes: they can also be used on their own (precond/FSAI/FSAI.hpp:14-47).
template <typename R = int, typename C = int, typename V = double>
class FSAI {
 public:
  std::string name = "FSAI";
  const deco_ptree& param;
  deco_prec* handle = nullptr;
  bool owner = true;
  std::function<void(const V* __restrict__, V*& __restrict__)> apply;
  FSAI(const deco_ptree& pt, std::shared_ptr<CSR<R, C, V>> A) : param(pt) {
    deco_ptree wrap;
    wrap.put("block_size", "1");
    for (auto& kv : pt.children) wrap.put("prec_glo." + kv.first, kv.second.data);
    deco_cpr_params p = deco_facade::params_from(wrap);
    deco_facade::check(deco_fsai_create(&p, A->device(), &handle), "FSAI");
    apply = deco_facade::make_callable<V>(handle, A->n);
  }
  ~FSAI() { if (owner && handle) deco_prec_destroy(handle); }
};

template <typename R = int, typename C = int, typename V = double>
class Jacobi {
 public:
  std::string name = "Jacobi";
  const deco_ptree& param;
  deco_prec* handle = nullptr;
  std::function<void(const V* __restrict__, V*& __restrict__)> apply;
  Jacobi(const deco_ptree& pt, std::shared_ptr<CSR<R, C, V>>& A) : param(pt) {
    deco_cpr_params p{};
    p.glo_block_size = pt.template get<int>("block_size", 1);
    deco_facade::check(deco_jacobi_create(&p, A->device(), &handle), "Jacobi");
    apply = deco_facade::make_callable<V>(handle, A->n);
  }
  ~Jacobi() { if (handle) deco_prec_destroy(handle); }
};

template <typename R = int, typename C = int, typename V = double>
class hypre_AMG {
 public:
  const deco_ptree& param;
  R n;
  deco_prec* handle = nullptr;
  std::function<void(const V*, V*&)> apply;
  hypre_AMG(const deco_ptree& pt, std::shared_ptr<CSR<R, C, V>>& A) : param(pt), n(A->n) {
    deco_ptree wrap;
    wrap.put("block_size", "1");
    for (auto& kv : pt.children) wrap.put("prec_loc." + kv.first, kv.second.data);
    deco_cpr_params p = deco_facade::params_from(wrap);
    deco_facade::check(deco_amg_create(&p, A->device(), &handle), "hypre_AMG");
    apply = deco_facade::make_callable<V>(handle, A->n);
  }
  ~hypre_AMG() { if (handle) deco_prec_destroy(handle); }
};

template <template <typename, typename, typename> class M>
struct glo_kind;
template <>
struct glo_kind<FSAI> { static constexpr int value = DECO_GLO_FSAI; };
template <>
struct glo_kind<Jacobi> { static constexpr int value = DECO_GLO_JACOBI; };

// CPR<R,C,V,Mglobal,Mlocal>(param, A, b) (precond/CPR.hpp:11-172): decouples A and b in place.
template <typename R = int, typename C = int, typename V = double,
          template <typename, typename, typename> class Mglobal = Jacobi,
          template <typename, typename, typename> class Mlocal = hypre_AMG>
class CPR : public CPRBase<V> {
 public:
  std::shared_ptr<CSR<R, C, V>> A;
  std::shared_ptr<CSR<R, C, V>> Ap;
  R n, sn;
  int bs, p;
  const deco_ptree& param;

  CPR(const deco_ptree& pt, std::shared_ptr<CSR<R, C, V>> A_, V*& b) : A(A_), param(pt) {
    deco_cpr_params prm = deco_facade::params_from(pt);
    prm.glo_type = glo_kind<Mglobal>::value;
    n = A->n;
    bs = prm.block_size;
    p = prm.pressure_id;
    sn = bs > 0 ? n / bs : 0;
    deco_facade::DevBuf bd;
    bd.ensure(n);
    if (b) deco_facade::check(deco_memcpy_h2d(bd.p, b, sizeof(V) * n), "h2d");
    deco_facade::check(deco_cpr_create(&prm, A->device(), b ? bd.p : nullptr, &this->handle), "CPR");
    A->pull();   // the host arrays see the decoupled system: the reference decouples A in place (CPR.hpp:91-114)
    if (b) deco_facade::check(deco_memcpy_d2h(b, bd.p, sizeof(V) * n), "d2h");
    const deco_csr* ap = nullptr;
    deco_facade::check(deco_cpr_pressure_matrix(this->handle, &ap), "pressure matrix");
    deco_csr* apc = nullptr;
    deco_facade::check(deco_csr_copy(ap, &apc), "copy");
    Ap = CSR<R, C, V>::from_device(apc);
    this->apply = deco_facade::make_callable<V>(this->handle, n);
  }
  ~CPR() override { if (this->handle) deco_prec_destroy(this->handle); }
};

// --------------------------------------------------------------------------------- SOL
// SOL(param, A, precond, b, x) (solver/SOL.hpp:19-49): runs the solve in the constructor on the
// device; takes ownership of b and x (frees both in the destructor, SOL.hpp:46-49).
template <typename R = int, typename C = int, typename V = double>
class SOL {
 public:
  const deco_ptree& param;
  std::shared_ptr<CSR<R, C, V>> A;
  V* b;
  V* x;
  const std::function<void(V*, V*&)>& precond;
  int iter = 0;
  double res_abs = 0, res_rel = 0;
  int flag = 1;
  deco_sol_info info{};

  SOL(const deco_ptree& param_, std::shared_ptr<CSR<R, C, V>> A_, const std::function<void(V*, V*&)>& precond_,
      V* b_, V* x_ = nullptr, std::shared_ptr<CSR<R, C, V>> Aorg_ = nullptr, V* borg_ = nullptr)
      : param(param_), A(A_), b(b_), x(x_), precond(precond_) {
    (void)Aorg_;
    (void)borg_;
    deco_sol_params sp{};
    const std::string type = param.template get<std::string>("type");
    if (type == "rgmres") sp.type = DECO_SOL_RGMRES;
    else if (type == "lgmres") sp.type = DECO_SOL_LGMRES;
    else if (type == "bicgstab") sp.type = DECO_SOL_BICGSTAB;
    else if (type == "lbicgstab") sp.type = DECO_SOL_LBICGSTAB;
    else if (type == "pcg") sp.type = DECO_SOL_PCG;
    else throw std::invalid_argument("Invalid solver type: " + type);
    sp.reltol = param.template get<double>("reltol");
    sp.restart = param.template get<int>("restart", 30);
    sp.itmax = param.template get<int>("itmax");
    sp.verbosity = param.template get<int>("verbosity", 0);
    sp.use_x0 = x ? 1 : 0;
    const std::string ortho = param.template get<std::string>("ortho", "dcgs2");
    if (ortho == "dcgs2") sp.ortho = DECO_ORTHO_DCGS2;
    else if (ortho == "mgs") sp.ortho = DECO_ORTHO_MGS;
    else throw std::invalid_argument("Invalid Gram-Schmidt variant: " + ortho);
    sp.flexible = param.template get<int>("flexible", 0);
    deco_prec* M = nullptr;
    if (precond) {
      auto* dp = precond.template target<deco_facade::DevicePrecond<V>>();
      if (!dp) throw std::invalid_argument("SOL runs on the device and needs a deco preconditioner");
      M = dp->h;
    }
    const int n = A->n;
    deco_facade::DevBuf bd, xd;
    bd.ensure(n);
    xd.ensure(n);
    deco_facade::check(deco_memcpy_h2d(bd.p, b, sizeof(V) * n), "h2d");
    if (x) deco_facade::check(deco_memcpy_h2d(xd.p, x, sizeof(V) * n), "h2d");
    else x = (V*)std::calloc(n, sizeof(V));
    deco_facade::check(deco_solve(&sp, A->device(), M, bd.p, xd.p, &info), "SOL");
    deco_facade::check(deco_memcpy_d2h(x, xd.p, sizeof(V) * n), "d2h");
    iter = info.iter;
    res_abs = info.res_abs;
    res_rel = info.res_rel;
    flag = info.flag;
  }
  V* get_result() const { return x; }
  ~SOL() {
    std::free(x);
    std::free(b);
  }
};

// AD::block_reorder (adapter/csr2coo.cpp:19-88): the driver's step before the solve
// (cpr.cpp:41), run on the device; the host arrays of A and b are refreshed from it.
namespace AD {
template <class R, class C, class V>
void block_reorder(std::shared_ptr<CSR<R, C, V>> A, V*& b, const deco_ptree& param) {
  const int bs = param.template get<int>("prec.block_size");
  const std::string row_order = param.template get<std::string>("prec.row_order");
  const std::string col_order = param.template get<std::string>("prec.col_order");
  deco_facade::DevBuf bd;
  bd.ensure(A->n);
  deco_facade::check(deco_memcpy_h2d(bd.p, b, sizeof(V) * A->n), "h2d");
  deco_facade::check(deco_block_reorder(A->device(), bs, row_order.c_str(), col_order.c_str(), bd.p),
                     "block_reorder");
  deco_facade::check(deco_memcpy_d2h(b, bd.p, sizeof(V) * A->n), "d2h");
  A->pull();
}
}  // namespace AD

// The reference's global profiler (PROFILE profiler(PROFILE::MPI), core/profiler.hpp:178): the
// same start/stop/report/getTime calls, recorded in the library's phase tree so the caller's
// phases ("IO", "setup & solve", "Prec", "Solve") and the library's own ("decoupling",
// "Global precond", "AMG apply", ...) form one tree. The first start() turns timing on.
namespace deco_facade {
struct PROFILE {
  void start(const std::string& name) {
    deco_facade::check(deco_profile_enable(1), "profile");
    deco_facade::check(deco_profile_start(name.c_str()), "profile");
  }
  void stop() { deco_facade::check(deco_profile_stop(), "profile"); }
  double getTime(const std::string& name) const {
    double t = 0.0;
    deco_facade::check(deco_profile_time(name.c_str(), &t), "profile");
    return t;
  }
  void reset() { deco_facade::check(deco_profile_reset(), "profile"); }
  // PROFILE::post_process (core/profiler.hpp:159-175) subtracts the matrix-format conversions of
  // hypre ("Hypre_CSR2IJMatrix") and Chronos from the setup phases; this library records no such
  // phase (its AMG reads the device CSR directly), so there is nothing to subtract
  void post_process() const {
    const double conv = getTime("Hypre_CSR2IJMatrix") + getTime("Chronos_build_dsmat");
    if (conv != 0.0) throw std::logic_error("profile: unexpected matrix-format conversion phase");
  }
  void report() const {
    size_t need = 0;
    deco_facade::check(deco_profile_report(nullptr, 0, &need), "profile");
    std::string s(need, '\0');
    deco_facade::check(deco_profile_report(&s[0], need, nullptr), "profile");
    std::fputs(s.c_str(), stdout);
  }
};
}  // namespace deco_facade
inline deco_facade::PROFILE profiler;
