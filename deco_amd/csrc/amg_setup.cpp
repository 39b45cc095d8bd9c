// Host part of the AMG setup (the hypre BoomerAMG substitute): strength of connection,
// C/F splitting and extended+i interpolation with truncation. Deterministic for any thread
// count: every row is computed independently and coarsening uses fixed tie-breaking or a seeded
// hash for its random measure.
#include "amg_setup.hpp"
#include <algorithm>
#include <cmath>
#include <cstring>
#include <omp.h>

namespace deco {
namespace amg {

static inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

void prefault(void* p, size_t bytes) {
  char* b = static_cast<char*>(p);
  const int64_t pages = (int64_t)((bytes + 4095) / 4096);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < pages; ++k) b[k * 4096] = 0;
}

void strength(const HostCSR& A, double theta, double max_row_sum, std::vector<char>& strong) {
  strong.assign(A.c.size(), 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < A.n; ++i) {
    const int b = A.r[i], e = A.r[i + 1];
    double diag = 0.0;
    for (int k = b; k < e; ++k)
      if (A.c[k] == i) { diag = A.v[k]; break; }
    double row_scale = 0.0, row_sum = diag;
    for (int k = b; k < e; ++k) {
      if (A.c[k] == i) continue;
      const double a = A.v[k];
      if (diag < 0.0) row_scale = std::max(row_scale, a);
      else row_scale = std::min(row_scale, a);
      row_sum += a;
    }
    row_sum = std::fabs(row_sum / diag);
    if (row_sum > max_row_sum && max_row_sum < 1.0) continue;  // all dependencies weak
    const double thr = theta * row_scale;
    for (int k = b; k < e; ++k) {
      if (A.c[k] == i) continue;
      const double a = A.v[k];
      strong[k] = (diag < 0.0) ? (a > thr) : (a < thr);
    }
  }
}

// strong dependencies S (rows) and strong influences S^T as CSR adjacency; the S^T lists are
// filled in parallel and then sorted, which gives them the ascending order of a serial fill
static void adjacency(const HostCSR& A, const std::vector<char>& strong, hvec<int>& sr, hvec<int>& sc, hvec<int>& tr,
                      hvec<int>& tc) {
  const int n = A.n;
  sr.assign(n + 1, 0);
  tr.assign(n + 1, 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int cnt = 0;
    for (int k = A.r[i]; k < A.r[i + 1]; ++k)
      if (strong[k]) {
        ++cnt;
#pragma omp atomic
        tr[A.c[k] + 1]++;
      }
    sr[i + 1] = cnt;
  }
  for (int i = 0; i < n; ++i) { sr[i + 1] += sr[i]; tr[i + 1] += tr[i]; }
  sc.resize(sr[n]);
  tc.resize(tr[n]);
  std::vector<int> cur(tr.begin(), tr.end() - 1);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    int o = sr[i];
    for (int k = A.r[i]; k < A.r[i + 1]; ++k)
      if (strong[k]) {
        sc[o++] = A.c[k];
        int slot;
#pragma omp atomic capture
        slot = cur[A.c[k]]++;
        tc[slot] = i;
      }
  }
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; ++j) std::sort(tc.begin() + tr[j], tc.begin() + tr[j + 1]);
}

// Ruge-Stueben first pass with bucket lists (max measure first, LIFO inside a bucket). The
// measure, the list links and the C/F state of a point share one 16-byte record: the pass is
// bound by random accesses to the neighbours of each new C point, one cache line each.
static void coarsen_rs(int n, const hvec<int>& sr, const hvec<int>& sc, const hvec<int>& tr, const hvec<int>& tc,
                       std::vector<int>& cf_out) {
  struct Pt {
    int lam, nxt, prv, cf;
  };
  std::vector<Pt> P(n);
  int maxdeg = 0;
  for (int i = 0; i < n; ++i) {
    P[i] = Pt{tr[i + 1] - tr[i], -1, -1, 0};
    maxdeg = std::max(maxdeg, P[i].lam);
  }
  const int nb = 2 * maxdeg + 2;
  std::vector<int> head(nb, -1);
  auto insert = [&](int i) {
    Pt& p = P[i];
    p.nxt = head[p.lam];
    p.prv = -1;
    if (p.nxt >= 0) P[p.nxt].prv = i;
    head[p.lam] = i;
  };
  auto remove = [&](int i) {
    Pt& p = P[i];
    if (p.prv >= 0) P[p.prv].nxt = p.nxt; else head[p.lam] = p.nxt;
    if (p.nxt >= 0) P[p.nxt].prv = p.prv;
    p.nxt = p.prv = -1;
  };
  for (int i = 0; i < n; ++i) {
    if (P[i].lam == 0) P[i].cf = -1;  // influences nobody (or isolated): F
    else insert(i);
  }
  int top = nb - 1;
  while (true) {
    while (top > 0 && head[top] < 0) --top;
    if (top <= 0) break;
    const int i = head[top];
    remove(i);
    P[i].cf = 1;
    for (int k = tr[i]; k < tr[i + 1]; ++k) {  // points depending on i become F
      const int j = tc[k];
      if (P[j].cf != 0) continue;
      remove(j);
      P[j].cf = -1;
      for (int q = sr[j]; q < sr[j + 1]; ++q) {  // their influencers gain weight
        const int m = sc[q];
        if (P[m].cf != 0) continue;
        remove(m);
        P[m].lam = std::min(P[m].lam + 1, nb - 1);
        insert(m);
        if (P[m].lam > top) top = P[m].lam;
      }
    }
    for (int q = sr[i]; q < sr[i + 1]; ++q) {  // points i depends on lose weight
      const int m = sc[q];
      if (P[m].cf != 0) continue;
      remove(m);
      P[m].lam = std::max(P[m].lam - 1, 0);
      insert(m);
    }
  }
  cf_out.resize(n);
  for (int i = 0; i < n; ++i) cf_out[i] = P[i].cf == 0 ? -1 : P[i].cf;
}

// PMIS: measure |S^T_i| + U[0,1) from a seeded hash; independent sets of local maxima.
static void coarsen_pmis(int n, const hvec<int>& sr, const hvec<int>& sc, const hvec<int>& tr, const hvec<int>& tc,
                         uint64_t seed, std::vector<int>& cf) {
  cf.assign(n, 0);
  std::vector<double> meas(n);
  for (int i = 0; i < n; ++i) {
    const double u = (double)(mix64(seed ^ mix64((uint64_t)i)) >> 11) * (1.0 / 9007199254740992.0);
    meas[i] = (double)(tr[i + 1] - tr[i]) + u;
    if (meas[i] < 1.0) cf[i] = -1;
  }
  std::vector<char> newc(n, 0);
  while (true) {
    int undecided = 0;
#pragma omp parallel for schedule(static) reduction(+ : undecided)
    for (int i = 0; i < n; ++i) {
      newc[i] = 0;
      if (cf[i] != 0) continue;
      ++undecided;
      bool mx = true;
      for (int q = sr[i]; q < sr[i + 1] && mx; ++q) {
        const int j = sc[q];
        if (cf[j] == 0 && meas[j] >= meas[i] && j != i) mx = false;
      }
      for (int q = tr[i]; q < tr[i + 1] && mx; ++q) {
        const int j = tc[q];
        if (cf[j] == 0 && meas[j] >= meas[i] && j != i) mx = false;
      }
      newc[i] = mx;
    }
    if (undecided == 0) break;
    for (int i = 0; i < n; ++i)
      if (newc[i]) cf[i] = 1;
    for (int i = 0; i < n; ++i)
      if (newc[i])
        for (int q = tr[i]; q < tr[i + 1]; ++q)
          if (cf[tc[q]] == 0) cf[tc[q]] = -1;
  }
}

void coarsen(const HostCSR& A, const std::vector<char>& strong, int coarse_type, uint64_t seed, std::vector<int>& cf) {
  hvec<int> sr, sc, tr, tc;
  adjacency(A, strong, sr, sc, tr, tc);
  if (coarse_type == 8) coarsen_pmis(A.n, sr, sc, tr, tc, seed, cf);
  else coarsen_rs(A.n, sr, sc, tr, tc, cf);
}

HostCSR interp_extpi(const HostCSR& A, const std::vector<char>& strong, const std::vector<int>& cf, double trunc,
                     int p_max, int& nc) {
  const int n = A.n;
  std::vector<int> cidx(n, -1);
  nc = 0;
  for (int i = 0; i < n; ++i)
    if (cf[i] == 1) cidx[i] = nc++;
  // Rows are appended to per-thread buffers (no allocation per row); row i lives in the buffer
  // of thread owner[i] at rstart[i], rcnt[i] entries, and is copied into P afterwards.
  const int nthreads = omp_get_max_threads();
  std::vector<std::vector<std::pair<int, double>>> tbuf(nthreads);
  std::vector<int> owner(n, 0), rcnt(n, 0);
  std::vector<size_t> rstart(n, 0);
#pragma omp parallel
  {
    auto& buf = tbuf[omp_get_thread_num()];
    const int me = omp_get_thread_num();
    std::vector<int> pos(n, -1);
    std::vector<int> list;
    std::vector<double> coef;
    std::vector<std::pair<int, double>> ent, kept;
#pragma omp for schedule(dynamic, 512)
    for (int i = 0; i < n; ++i) {
      owner[i] = me;
      rstart[i] = buf.size();
      if (cf[i] == 1) { buf.push_back({cidx[i], 1.0}); rcnt[i] = 1; continue; }
      list.clear();
      coef.clear();
      auto add = [&](int j) {
        if (pos[j] < 0) { pos[j] = (int)list.size(); list.push_back(j); coef.push_back(0.0); }
      };
      const int b = A.r[i], e = A.r[i + 1];
      for (int k = b; k < e; ++k) {
        const int j = A.c[k];
        if (j == i || !strong[k]) continue;
        if (cf[j] == 1) add(j);
        else
          for (int kk = A.r[j]; kk < A.r[j + 1]; ++kk)
            if (strong[kk] && cf[A.c[kk]] == 1) add(A.c[kk]);
      }
      double diagonal = 0.0;
      for (int k = b; k < e; ++k)
        if (A.c[k] == i) diagonal += A.v[k];
      for (int k = b; k < e; ++k) {
        const int j = A.c[k];
        if (j == i) continue;
        const double a = A.v[k];
        if (pos[j] >= 0) {
          coef[pos[j]] += a;
        } else if (strong[k] && cf[j] != 1) {
          double ajj = 0.0;
          for (int kk = A.r[j]; kk < A.r[j + 1]; ++kk)
            if (A.c[kk] == j) { ajj = A.v[kk]; break; }
          const double sgn = (ajj < 0.0) ? -1.0 : 1.0;
          double sum = 0.0;
          for (int kk = A.r[j]; kk < A.r[j + 1]; ++kk) {
            const int jj = A.c[kk];
            if ((pos[jj] >= 0 || jj == i) && sgn * A.v[kk] < 0.0) sum += A.v[kk];
          }
          if (sum != 0.0) {
            const double dist = a / sum;
            for (int kk = A.r[j]; kk < A.r[j + 1]; ++kk) {
              const int jj = A.c[kk];
              if (sgn * A.v[kk] >= 0.0) continue;
              if (pos[jj] >= 0) coef[pos[jj]] += dist * A.v[kk];
              else if (jj == i) diagonal += dist * A.v[kk];
            }
          } else {
            diagonal += a;
          }
        } else {
          diagonal += a;
        }
      }
      for (int j : list) pos[j] = -1;
      if (list.empty()) continue;
      for (size_t q = 0; q < coef.size(); ++q) coef[q] = (diagonal != 0.0) ? -coef[q] / diagonal : 0.0;
      ent.clear();
      for (size_t q = 0; q < list.size(); ++q) ent.push_back({list[q], coef[q]});
      // truncation relative to the row maximum, keeping the row sum
      if (trunc > 0.0) {
        double mx = 0.0, rs = 0.0;
        for (auto& p : ent) { mx = std::max(mx, std::fabs(p.second)); rs += p.second; }
        const double thr = trunc * mx;
        kept.clear();
        double ks = 0.0;
        for (auto& p : ent)
          if (std::fabs(p.second) >= thr) { kept.push_back(p); ks += p.second; }
        if (kept.size() < ent.size() && ks != 0.0) {
          const double sc = rs / ks;
          for (auto& p : kept) p.second *= sc;
        }
        ent.swap(kept);
      }
      if (p_max > 0 && (int)ent.size() > p_max) {
        double rs = 0.0;
        for (auto& p : ent) rs += p.second;
        std::sort(ent.begin(), ent.end(), [](const std::pair<int, double>& a, const std::pair<int, double>& b) {
          const double fa = std::fabs(a.second), fb = std::fabs(b.second);
          if (fa != fb) return fa > fb;
          return a.first < b.first;
        });
        ent.resize(p_max);
        double ks = 0.0;
        for (auto& p : ent) ks += p.second;
        if (ks != 0.0) {
          const double sc = rs / ks;
          for (auto& p : ent) p.second *= sc;
        }
      }
      const size_t s0 = buf.size();
      for (auto& p : ent) buf.push_back({cidx[p.first], p.second});
      std::sort(buf.begin() + s0, buf.end());
      rcnt[i] = (int)(buf.size() - s0);
    }
  }
  HostCSR P;
  P.n = n;
  P.m = nc;
  P.r.assign(n + 1, 0);
  for (int i = 0; i < n; ++i) P.r[i + 1] = P.r[i] + rcnt[i];
  P.c.resize(P.r[n]);
  P.v.resize(P.r[n]);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const auto* src = tbuf[owner[i]].data() + rstart[i];
    for (int q = 0; q < rcnt[i]; ++q) {
      P.c[P.r[i] + q] = src[q].first;
      P.v[P.r[i] + q] = src[q].second;
    }
  }
  return P;
}

bool dense_inverse(const HostCSR& A, std::vector<double>& inv) {
  const int n = A.n;
  std::vector<double> a((size_t)n * n, 0.0);  // column-major
  for (int i = 0; i < n; ++i)
    for (int k = A.r[i]; k < A.r[i + 1]; ++k) a[i + (size_t)A.c[k] * n] = A.v[k];
  std::vector<int> piv(n);
  for (int j = 0; j < n; ++j) {
    int p = j;
    double mx = std::fabs(a[j + (size_t)j * n]);
    for (int i = j + 1; i < n; ++i)
      if (std::fabs(a[i + (size_t)j * n]) > mx) { mx = std::fabs(a[i + (size_t)j * n]); p = i; }
    piv[j] = p;
    if (a[p + (size_t)j * n] == 0.0) return false;
    if (p != j)
      for (int q = 0; q < n; ++q) std::swap(a[j + (size_t)q * n], a[p + (size_t)q * n]);
    const double rp = 1.0 / a[j + (size_t)j * n];
    for (int i = j + 1; i < n; ++i) a[i + (size_t)j * n] *= rp;
    for (int q = j + 1; q < n; ++q) {
      const double t = a[j + (size_t)q * n];
      if (t != 0.0)
        for (int i = j + 1; i < n; ++i) a[i + (size_t)q * n] -= a[i + (size_t)j * n] * t;
    }
  }
  inv.assign((size_t)n * n, 0.0);
  std::vector<double> col(n);
  for (int k = 0; k < n; ++k) {
    std::fill(col.begin(), col.end(), 0.0);
    col[k] = 1.0;
    for (int j = 0; j < n; ++j)
      if (piv[j] != j) std::swap(col[j], col[piv[j]]);
    for (int j = 0; j < n; ++j) {
      const double t = col[j];
      if (t != 0.0)
        for (int i = j + 1; i < n; ++i) col[i] -= t * a[i + (size_t)j * n];
    }
    for (int j = n - 1; j >= 0; --j) {
      col[j] /= a[j + (size_t)j * n];
      const double t = col[j];
      for (int i = 0; i < j; ++i) col[i] -= t * a[i + (size_t)j * n];
    }
    for (int i = 0; i < n; ++i) inv[(size_t)i * n + k] = col[i];
  }
  return true;
}

HostCSR host_spgemm(const HostCSR& A, const HostCSR& B) {
  HostCSR C;
  C.n = A.n;
  C.m = B.m;
  C.r.assign(A.n + 1, 0);
  std::vector<double> acc(B.m, 0.0);
  std::vector<char> on(B.m, 0);
  std::vector<int> cols;
  for (int i = 0; i < A.n; ++i) {
    cols.clear();
    for (int e = A.r[i]; e < A.r[i + 1]; ++e) {
      const int k = A.c[e];
      const double a = A.v[e];
      for (int f = B.r[k]; f < B.r[k + 1]; ++f) {
        const int j = B.c[f];
        if (!on[j]) { on[j] = 1; cols.push_back(j); acc[j] = 0.0; }
        acc[j] += a * B.v[f];
      }
    }
    std::sort(cols.begin(), cols.end());
    for (int j : cols) {
      C.c.push_back(j);
      C.v.push_back(acc[j]);
      on[j] = 0;
    }
    C.r[i + 1] = (int)C.c.size();
  }
  return C;
}

}  // namespace amg
}  // namespace deco
