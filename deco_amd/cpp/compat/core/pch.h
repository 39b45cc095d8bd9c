// The header a reference-style driver includes first (the role of core/pch.h), for builds on
// libdeco_hip.so: the class templates of the path (deco.hpp), the binary CSR / rhs format
// (io.hpp) and the one-rank MPI and set-up calls (mpi1.hpp). Add deco_amd/cpp/compat and include/
// to the include path and link libdeco_hip.so.
#pragma once
#include "../../deco.hpp"
#include "../../io.hpp"
#include "../../mpi1.hpp"
#if !defined(DECO_USE_BOOST_PTREE)
namespace boost {
namespace property_tree {
using ptree = deco::ptree;
}  // namespace property_tree
}  // namespace boost
#endif
