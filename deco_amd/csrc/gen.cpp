// Seeded synthetic two-/three-phase finite-volume Jacobians on a 3-D grid.
//
// Layout (the conventions the reference's CPR code assumes, precond/CPR.hpp:321-447):
//   * cells in natural x-fastest order, id = x + nx*(y + ny*z); inactive cells (mask) are
//     dropped and the active ones renumbered in the same order;
//   * bs unknowns per cell, interleaved, pressure first (unknown index = cell*bs + k, k=0 is
//     pressure);
//   * every connection (7-point neighbours plus optional non-neighbour "fault" links) emits a
//     full bs x bs block in both rows, explicit zeros kept, so all rows of a cell share one
//     sorted column set and the diagonal block is dense;
//   * column indices sorted, a diagonal in every row.
//
// Random numbers are a pure function of (seed, stream, key) (splitmix64 hashing), so rows can
// be generated in parallel and the result does not depend on the thread count.
//
// Values. For cell i with neighbour set N(i) and transmissibility T_ij = 2 k_i k_j /
// (k_i + k_j) * f_ij (f_ij = kz_kx on z faces, 1 otherwise; fault links use T ~ lognormal):
//   up(ij)   = i or j with probability 1/2 each (seeded per undirected link),
//   lam(c,k) = mobility of phase k in cell c, U[lam_lo, lam_hi],
//   acc(i)   = phi*c_t*V/dt, U[acc_lo, acc_hi],
//   sacc(i,k)= saturation accumulation of phase k, U[sacc_lo, sacc_hi],
//   cpl(...) = coupling scale U[cpl_lo, cpl_hi], an independent draw per entry.
//   pressure row (k=0):
//     A[p_i, p_i] = sum_j T_ij + acc(i)          A[p_i, p_j] = -T_ij
//     A[p_i, s_i^l] = cpl * (sum_j T_ij + acc(i))  A[p_i, s_j^l] = -cpl * T_ij * lam(up,l)
//   (every coupling entry is the matching entry of its own row's physics times cpl, so the
//   quasi-IMPES Schur complement of each cell stays >= 0.88 of the pressure diagonal)
//   saturation row k >= 1 (transport by the upwinded mobilities; the transmissibility scale is
//   carried by the pressure equation):
//     A[s_i^k, s_i^k] = sacc(i,k) + sum_{j: up(ij)=i} lam(i,k)
//     A[s_i^k, s_j^k] = -lam(j,k) if up(ij)=j else 0 (explicit zero)
//     A[s_i^k, p_j]   = -cpl_ij * lam(up(ij),k) * T_ij / Tbar_i   (Tbar_i = mean T of cell i's faces)
//     A[s_i^k, p_i]   = sum_j cpl_ij * lam(up(ij),k) * T_ij / Tbar_i
//     A[s_i^k, s_i^l] = cpl * sacc(i,l)  (l != k)   A[s_i^k, s_j^l] = -cpl * lam(up,l)
//   The pressure couplings of a saturation row are its fluxes' pressure derivatives: they follow
//   the face transmissibilities (normalised to the row's O(1) scale) and sum to zero, because a
//   flux depends on pressure differences only. Then the decoupled pressure matrix keeps the
//   pressure block's M-matrix structure and positive row sums. With couplings that ignore T_ij,
//   or whose diagonal is an independent draw, the quasi-IMPES correction (~cpl^2 sum_j T_ij)
//   puts positive off-diagonals and negative row sums into it wherever T varies strongly
//   between faces (sigma 3), and CPR stagnated.
// (With T_ij * lam as the saturation coefficient the saturation block is transport dominated,
// T lam / sacc up to ~1e3 for the sigma-2 layered case, and CPR iteration counts grow without
// bound with the grid; that variant is selected by sat_scale_T.)
// The matrix is nonsymmetric. x* ~ U[-1,1] per unknown and b = A x* (row sums in column order).
//
// Configurations (1000+cfg is the default seed):
//   1   32x32x16, bs 2, log k ~ N(0,1), kz/kx 1, acc U[1e-4,1e-2]   (as written: 32,768 unknowns)
//   11  32x32x32, same parameters                                    (as counted: 65,536 unknowns)
//   2   46x112x22, bs 3, 60% active mask, +2% fault links T ~ logN(0,1), log k ~ N(0,1.5),
//       kz/kx 0.1, acc U[1e-4,1e-2]
//   3   168x100x120, bs 2, 12 layers in z with log-mean U[-2,2], sigma 2, kz/kx 0.1,
//       acc U[1e-6,1e-4]
//   4   16x16x8, bs 2, config-1 parameters (adjoint finite-difference case)
//   5   64x118x263, bs 2, one layer per z-slice with log-mean U[-2,2], sigma 3, kz/kx 1e-3,
//       acc U[1e-4,1e-2]
#include "deco_hip.h"
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <algorithm>
#include <unordered_set>
#include <string>
#include <stdexcept>
#include <omp.h>

namespace deco_gen_impl {

static inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// uniform in [0,1)
static inline double u01(uint64_t seed, uint64_t stream, uint64_t key) {
  uint64_t h = splitmix64(seed ^ splitmix64(stream * 0xD1B54A32D192ED03ull + 0x8CB92BA72F3D8DD7ull));
  h = splitmix64(h ^ splitmix64(key));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}
static inline double uab(uint64_t seed, uint64_t stream, uint64_t key, double a, double b) {
  return a + (b - a) * u01(seed, stream, key);
}
static inline double normal(uint64_t seed, uint64_t stream, uint64_t key) {
  double u1 = 1.0 - u01(seed, stream, 2 * key);
  double u2 = u01(seed, stream, 2 * key + 1);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

enum : uint64_t {
  S_MASK = 1, S_PERM = 2, S_LAYER = 3, S_ACC = 4, S_SACC = 5, S_LAM = 6, S_UP = 7,
  S_CPL_DIAG = 8, S_CPL_OFF = 9, S_FAULT_I = 10, S_FAULT_J = 11, S_FAULT_T = 12, S_X = 13
};

struct Spec {
  int nx, ny, nz, bs;
  double active_frac = 1.0;
  double fault_frac = 0.0, fault_sigma = 1.0;
  int n_layers = 0;            // 0: homogeneous log-mean
  double perm_mu = 0.0, perm_sigma = 1.0, layer_lo = 0.0, layer_hi = 0.0;
  double kz_kx = 1.0;
  double acc_lo = 1e-4, acc_hi = 1e-2;
  double lam_lo = 0.1, lam_hi = 1.0;
  double sacc_lo = 0.5, sacc_hi = 2.0;
  double cpl_lo = 0.01, cpl_hi = 0.1;
  bool sat_scale_T = false;    // saturation transport T*lam instead of lam
  bool sat_p_flux = true;      // saturation-row pressure couplings weighted by T_ij / mean T
  uint64_t seed = 0;
};

static Spec spec_for(int cfg) {
  Spec s;
  switch (cfg) {
    case 1:  s.nx = 32; s.ny = 32; s.nz = 16; s.bs = 2; break;
    case 11: s.nx = 32; s.ny = 32; s.nz = 32; s.bs = 2; break;
    case 2:
      s.nx = 46; s.ny = 112; s.nz = 22; s.bs = 3; s.active_frac = 0.6; s.fault_frac = 0.02;
      s.fault_sigma = 1.0; s.perm_sigma = 1.5; s.kz_kx = 0.1; break;
    case 3:
      s.nx = 168; s.ny = 100; s.nz = 120; s.bs = 2; s.n_layers = 12; s.perm_sigma = 2.0;
      s.layer_lo = -2.0; s.layer_hi = 2.0; s.kz_kx = 0.1; s.acc_lo = 1e-6; s.acc_hi = 1e-4; break;
    case 4:  s.nx = 16; s.ny = 16; s.nz = 8; s.bs = 2; break;
    case 5:
      s.nx = 64; s.ny = 118; s.nz = 263; s.bs = 2; s.n_layers = 263; s.perm_sigma = 3.0;
      s.layer_lo = -2.0; s.layer_hi = 2.0; s.kz_kx = 1e-3; break;
    default: throw std::invalid_argument("unknown synthetic configuration");
  }
  s.seed = 1000 + (uint64_t)cfg;
  return s;
}

}  // namespace deco_gen_impl

struct deco_gen {
  int n = 0, nnz = 0, bs = 0, ncells = 0;
  std::vector<int> r, c;
  std::vector<double> v, b, xs;
};

using namespace deco_gen_impl;

static void build_system(const Spec& sp, deco_gen* g) {
  const int nx = sp.nx, ny = sp.ny, nz = sp.nz, bs = sp.bs;
  const int64_t ngrid = (int64_t)nx * ny * nz;
  const uint64_t seed = sp.seed;

  // active cells and renumbering
  std::vector<int> gid2cell(ngrid, -1);
  std::vector<int64_t> cell2gid;
  cell2gid.reserve(ngrid);
  for (int64_t q = 0; q < ngrid; ++q) {
    bool act = sp.active_frac >= 1.0 || u01(seed, S_MASK, (uint64_t)q) < sp.active_frac;
    if (act) { gid2cell[q] = (int)cell2gid.size(); cell2gid.push_back(q); }
  }
  const int nc = (int)cell2gid.size();

  // permeability per active cell (indexed by grid id for determinism)
  std::vector<double> perm(nc);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nc; ++i) {
    int64_t q = cell2gid[i];
    int z = (int)(q / ((int64_t)nx * ny));
    double mu = sp.perm_mu;
    if (sp.n_layers > 0) {
      int layer = (int)((int64_t)z * sp.n_layers / nz);
      mu = uab(seed, S_LAYER, (uint64_t)layer, sp.layer_lo, sp.layer_hi);
    }
    perm[i] = std::exp(mu + sp.perm_sigma * normal(seed, S_PERM, (uint64_t)q));
  }

  // 7-point neighbour lists with transmissibilities
  std::vector<std::vector<std::pair<int, double>>> nbr(nc);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nc; ++i) {
    int64_t q = cell2gid[i];
    int x = (int)(q % nx), y = (int)((q / nx) % ny), z = (int)(q / ((int64_t)nx * ny));
    const int dx[6] = {-1, 1, 0, 0, 0, 0}, dy[6] = {0, 0, -1, 1, 0, 0}, dz[6] = {0, 0, 0, 0, -1, 1};
    for (int d = 0; d < 6; ++d) {
      int xx = x + dx[d], yy = y + dy[d], zz = z + dz[d];
      if (xx < 0 || xx >= nx || yy < 0 || yy >= ny || zz < 0 || zz >= nz) continue;
      int j = gid2cell[xx + (int64_t)nx * (yy + (int64_t)ny * zz)];
      if (j < 0) continue;
      double T = 2.0 * perm[i] * perm[j] / (perm[i] + perm[j]);
      if (dz[d] != 0) T *= sp.kz_kx;
      nbr[i].push_back({j, T});
    }
  }

  // non-neighbour fault links (sequential, seeded, rejection of duplicates/neighbours)
  if (sp.fault_frac > 0.0 && nc > 2) {
    int64_t nlinks = 0;
    for (int i = 0; i < nc; ++i) nlinks += (int64_t)nbr[i].size();
    nlinks /= 2;
    int64_t nf = (int64_t)std::llround(sp.fault_frac * (double)nlinks);
    std::unordered_set<uint64_t> used;
    for (int i = 0; i < nc; ++i)
      for (auto& e : nbr[i]) used.insert((uint64_t)std::min(i, e.first) * (uint64_t)nc + std::max(i, e.first));
    uint64_t trial = 0;
    for (int64_t f = 0; f < nf; ++trial) {
      int i = (int)(u01(seed, S_FAULT_I, trial) * nc);
      int j = (int)(u01(seed, S_FAULT_J, trial) * nc);
      if (i == j) continue;
      uint64_t key = (uint64_t)std::min(i, j) * (uint64_t)nc + std::max(i, j);
      if (used.count(key)) continue;
      used.insert(key);
      double T = std::exp(sp.fault_sigma * normal(seed, S_FAULT_T, key));
      nbr[i].push_back({j, T});
      nbr[j].push_back({i, T});
      ++f;
    }
  }
  for (int i = 0; i < nc; ++i) std::sort(nbr[i].begin(), nbr[i].end());

  // row pointers: every row of cell i has bs*(1+|N(i)|) entries
  const int n = nc * bs;
  g->n = n; g->bs = bs; g->ncells = nc;
  g->r.assign(n + 1, 0);
  for (int i = 0; i < nc; ++i)
    for (int k = 0; k < bs; ++k) g->r[i * bs + k + 1] = bs * (1 + (int)nbr[i].size());
  for (int i = 0; i < n; ++i) g->r[i + 1] += g->r[i];
  g->nnz = g->r[n];
  g->c.assign(g->nnz, 0);
  g->v.assign(g->nnz, 0.0);

  auto lam = [&](int cell, int k) { return uab(seed, S_LAM, (uint64_t)cell * 8 + k, sp.lam_lo, sp.lam_hi); };
  auto sacc = [&](int cell, int k) { return uab(seed, S_SACC, (uint64_t)cell * 8 + k, sp.sacc_lo, sp.sacc_hi); };
  auto link_key = [&](int i, int j) { return (uint64_t)std::min(i, j) * (uint64_t)nc + (uint64_t)std::max(i, j); };
  auto upwind = [&](int i, int j) { return u01(seed, S_UP, link_key(i, j)) < 0.5 ? std::min(i, j) : std::max(i, j); };
  // coupling draw for the entry (row unknown k of cell i, column unknown l of cell j)
  auto cpl_off = [&](int i, int j, int k, int l) {
    uint64_t key = (((uint64_t)i * (uint64_t)nc + (uint64_t)j) * 4 + (uint64_t)k) * 4 + (uint64_t)l;
    return uab(seed, S_CPL_OFF, key, sp.cpl_lo, sp.cpl_hi);
  };
  auto cpl_diag = [&](int i, int k, int l) {
    return uab(seed, S_CPL_DIAG, ((uint64_t)i * 4 + k) * 4 + l, sp.cpl_lo, sp.cpl_hi);
  };

#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < nc; ++i) {
    const auto& N = nbr[i];
    // column cells in sorted order with the diagonal cell inserted
    std::vector<int> cols;
    cols.reserve(N.size() + 1);
    bool placed = false;
    for (auto& e : N) {
      if (!placed && e.first > i) { cols.push_back(i); placed = true; }
      cols.push_back(e.first);
    }
    if (!placed) cols.push_back(i);
    double sumT = 0.0;
    for (auto& e : N) sumT += e.second;
    // face weights of the saturation rows' pressure couplings: T_ij over the cell's mean T
    const double tbar = N.empty() ? 1.0 : sumT / (double)N.size();
    auto fw = [&](double T) { return sp.sat_p_flux && !sp.sat_scale_T ? T / tbar : 1.0; };
    const double acc_i = uab(seed, S_ACC, (uint64_t)i, sp.acc_lo, sp.acc_hi);
    for (int k = 0; k < bs; ++k) {
      const int row = i * bs + k;
      int pos = g->r[row];
      // per-row sums needed on the diagonal
      double out_k = 0.0, flux_k = 0.0;
      if (k >= 1)
        for (auto& e : N) {
          int up = upwind(i, e.first);
          const double ts = sp.sat_scale_T ? e.second : 1.0;
          if (up == i) out_k += ts * lam(i, k);
          flux_k += cpl_off(i, e.first, k, 0) * ts * fw(e.second) * lam(up, k);
        }
      size_t ni = 0;
      for (int cj : cols) {
        const bool diag = (cj == i);
        double T = 0.0;
        int up = i;
        if (!diag) {
          while (N[ni].first != cj) ++ni;
          T = N[ni].second;
          up = upwind(i, cj);
        }
        for (int l = 0; l < bs; ++l) {
          double val;
          if (k == 0) {
            if (l == 0) val = diag ? sumT + acc_i : -T;
            else val = diag ? cpl_diag(i, 0, l) * (sumT + acc_i) : -cpl_off(i, cj, 0, l) * T * lam(up, l);
          } else {
            const double ts = sp.sat_scale_T ? T : 1.0;
            if (l == k) val = diag ? sacc(i, k) + out_k : (up == cj ? -ts * lam(cj, k) : 0.0);
            else if (l == 0) val = diag ? flux_k : -cpl_off(i, cj, k, 0) * ts * fw(T) * lam(up, k);
            else val = diag ? cpl_diag(i, k, l) * sacc(i, l) : -cpl_off(i, cj, k, l) * ts * lam(up, l);
          }
          g->c[pos] = cj * bs + l;
          g->v[pos] = val;
          ++pos;
        }
      }
    }
  }

  g->xs.resize(n);
  g->b.resize(n);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) g->xs[i] = 2.0 * u01(seed, S_X, (uint64_t)i) - 1.0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int j = g->r[i]; j < g->r[i + 1]; ++j) s += g->v[j] * g->xs[g->c[j]];
    g->b[i] = s;
  }
}

static thread_local std::string g_gen_err;

extern "C" {

int deco_gen_create(int cfg, uint64_t seed, deco_gen** out) {
  if (!out) return DECO_ERR_ARG;
  try {
    Spec sp = spec_for(cfg);
    if (seed != 0) sp.seed = seed;
    auto* g = new deco_gen();
    build_system(sp, g);
    *out = g;
    return DECO_OK;
  } catch (const std::exception& e) {
    g_gen_err = e.what();
    return DECO_ERR_ARG;
  }
}

int deco_gen_custom(int nx, int ny, int nz, int bs, uint64_t seed, deco_gen** out) {
  if (!out || nx <= 0 || ny <= 0 || nz <= 0 || bs < 1 || bs > 3) return DECO_ERR_ARG;
  Spec sp = spec_for(1);
  sp.nx = nx; sp.ny = ny; sp.nz = nz; sp.bs = bs;
  sp.seed = seed ? seed : 1000;
  auto* g = new deco_gen();
  build_system(sp, g);
  *out = g;
  return DECO_OK;
}

int deco_gen_scaled(int cfg, int nx, int ny, int nz, uint64_t seed, deco_gen** out) {
  if (!out || nx <= 0 || ny <= 0 || nz <= 0) return DECO_ERR_ARG;
  try {
    Spec sp = spec_for(cfg);
    sp.nx = nx; sp.ny = ny;
    if (sp.n_layers == sp.nz) sp.n_layers = nz;   // one layer per z-slice stays one per slice
    sp.nz = nz;
    if (seed != 0) sp.seed = seed;
    auto* g = new deco_gen();
    build_system(sp, g);
    *out = g;
    return DECO_OK;
  } catch (const std::exception& e) {
    g_gen_err = e.what();
    return DECO_ERR_ARG;
  }
}

int deco_gen_info(const deco_gen* g, int* n, int* nnz, int* bs, int* ncells) {
  if (!g) return DECO_ERR_ARG;
  if (n) *n = g->n;
  if (nnz) *nnz = g->nnz;
  if (bs) *bs = g->bs;
  if (ncells) *ncells = g->ncells;
  return DECO_OK;
}

int deco_gen_arrays(const deco_gen* g, const int** r, const int** c, const double** v,
                    const double** b, const double** xstar) {
  if (!g) return DECO_ERR_ARG;
  if (r) *r = g->r.data();
  if (c) *c = g->c.data();
  if (v) *v = g->v.data();
  if (b) *b = g->b.data();
  if (xstar) *xstar = g->xs.data();
  return DECO_OK;
}

int deco_gen_destroy(deco_gen* g) {
  delete g;
  return DECO_OK;
}

}  // extern "C"
