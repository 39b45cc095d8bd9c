// Host part of the AMG setup that replaces hypre BoomerAMG's coarsening and interpolation
// (precond/hypre_AMG/hypre_AMG.tpp:44-67 names the parameters). The Galerkin products run on
// the device (amg.hip); this part sees one level at a time as host CSR arrays.
#pragma once
#include <cstdint>
#include <memory>
#include <utility>
#include <vector>

namespace deco {
namespace amg {

// std::allocator whose value-less construct leaves the element uninitialised: the host level
// arrays are always overwritten after a resize (by a download or a fill), so resize() need not
// zero hundreds of MB first.
template <class T>
struct NoInitAlloc : std::allocator<T> {
  using std::allocator<T>::allocator;
  template <class U>
  struct rebind { using other = NoInitAlloc<U>; };
  template <class U>
  void construct(U* p) noexcept { ::new (static_cast<void*>(p)) U; }
  template <class U, class... Args>
  void construct(U* p, Args&&... args) { ::new (static_cast<void*>(p)) U(std::forward<Args>(args)...); }
};
template <class T>
using hvec = std::vector<T, NoInitAlloc<T>>;

struct HostCSR {
  int n = 0, m = 0;
  hvec<int> r, c;
  hvec<double> v;
};

// Touches every page of [p, p + bytes) from all threads, so a copy into fresh (never written)
// host memory does not take its page faults one at a time.
void prefault(void* p, size_t bytes);

// Classical strength of connection with hypre's conventions (the sign of the diagonal decides
// which off-diagonals can be strong; rows whose |row sum / diagonal| exceeds max_row_sum have
// only weak connections). strong[k] is 1 for a strong off-diagonal entry k of A.
void strength(const HostCSR& A, double theta, double max_row_sum, std::vector<char>& strong);

// C/F splitting: cf[i] = 1 (C) or -1 (F). coarse_type 8 = PMIS (seeded random measure),
// 10 (HMIS on one process) and 0/1/3/6 = Ruge-Stueben first pass.
void coarsen(const HostCSR& A, const std::vector<char>& strong, int coarse_type, uint64_t seed, std::vector<int>& cf);

// Extended+i interpolation (hypre interp_type 6) followed by truncation
// (trunc_factor relative to the row maximum, at most p_max entries per row; rows rescaled to
// keep their sums). Returns P (n x nc) with sorted columns.
HostCSR interp_extpi(const HostCSR& A, const std::vector<char>& strong, const std::vector<int>& cf, double trunc,
                     int p_max, int& nc);

// Dense inverse of a small matrix by LU with partial pivoting; returns false if singular.
bool dense_inverse(const HostCSR& A, std::vector<double>& inv_rowmajor);

// C = A B on host CSR arrays (columns sorted); small operands (dense accumulator per row)
HostCSR host_spgemm(const HostCSR& A, const HostCSR& B);

}  // namespace amg
}  // namespace deco
