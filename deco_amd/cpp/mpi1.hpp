// One-rank versions of the MPI and set-up calls the reference's driver makes around the solve
// (cpr.cpp:12-13 init_MPI / MPI_Comm_set_errhandler, :17 IO::read_ptree, :47-53
// IO::distribute_csr_matrix / memres / MPI_Scatterv / get_mpi_type, :84 profiler.post_process,
// :88 MPI_Finalize; the helpers are core/header.h:68-89,141-153 and IO/cpp_read.hpp:471-577).
// The device path runs on one rank (one GPU per process), so these keep the reference's names and
// arguments with their one-rank meaning: rank 0 of 1, the whole matrix as rank 0's block, a
// scatter that copies. With a real mpi.h included first (MPI_VERSION defined) the MPI names are
// MPI's own and init_MPI rejects more than one rank.
#pragma once
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <type_traits>
#include <stdexcept>
#include <string>
#include "deco.hpp"

#ifndef MPI_VERSION
typedef int MPI_Comm;
typedef int MPI_Datatype;
typedef int MPI_Errhandler;
#define MPI_COMM_WORLD 0
#define MPI_ERRORS_RETURN 0
#define MPI_SUCCESS 0
#define MPI_INT 4
#define MPI_FLOAT 104
#define MPI_DOUBLE 108
inline int MPI_Init(int*, char***) { return MPI_SUCCESS; }
inline int MPI_Initialized(int* flag) { *flag = 1; return MPI_SUCCESS; }
inline int MPI_Finalize() { return MPI_SUCCESS; }
inline int MPI_Comm_rank(MPI_Comm, int* rank) { *rank = 0; return MPI_SUCCESS; }
inline int MPI_Comm_size(MPI_Comm, int* size) { *size = 1; return MPI_SUCCESS; }
inline int MPI_Comm_set_errhandler(MPI_Comm, MPI_Errhandler) { return MPI_SUCCESS; }
inline int MPI_Barrier(MPI_Comm) { return MPI_SUCCESS; }
// element size of the stand-in datatypes (their value modulo 100)
inline size_t deco_mpi_type_size(MPI_Datatype t) { return (size_t)(t % 100); }
// one rank: rank 0's block [displs[0], displs[0] + recvcount) of sendbuf goes to recvbuf
inline int MPI_Scatterv(const void* sendbuf, const int* sendcounts, const int* displs, MPI_Datatype sendtype,
                        void* recvbuf, int recvcount, MPI_Datatype recvtype, int root, MPI_Comm) {
  if (root != 0) throw std::invalid_argument("MPI_Scatterv: one rank, root must be 0");
  if (sendcounts[0] != recvcount || deco_mpi_type_size(sendtype) != deco_mpi_type_size(recvtype))
    throw std::invalid_argument("MPI_Scatterv: send and receive blocks differ");
  if (recvcount > 0)
    std::memcpy(recvbuf, static_cast<const char*>(sendbuf) + (size_t)displs[0] * deco_mpi_type_size(sendtype),
                (size_t)recvcount * deco_mpi_type_size(recvtype));
  return MPI_SUCCESS;
}
#endif

// the communicator of the path (core/pch.h:23)
inline MPI_Comm deco_comm = MPI_COMM_WORLD;

// init_MPI (core/header.h:68-89): rank and size of deco_comm
inline void init_MPI(int* id, int* np, int* argc = nullptr, char*** argv = nullptr, int np_ = 1) {
  (void)np_;
  int initialized = 0;
  MPI_Initialized(&initialized);
  if (!initialized) MPI_Init(argc, argv);
  MPI_Comm_rank(deco_comm, id);
  MPI_Comm_size(deco_comm, np);
  if (*np != 1) throw std::invalid_argument("the device solve path runs on one rank (one GPU per process)");
}

// get_mpi_type<T>() (core/header.h): the MPI datatype of T
template <typename T>
MPI_Datatype get_mpi_type() {
  if (std::is_same<T, double>::value) return MPI_DOUBLE;
  if (std::is_same<T, float>::value) return MPI_FLOAT;
  if (std::is_same<T, int>::value) return MPI_INT;
  throw std::invalid_argument("get_mpi_type: unsupported type");
}

// memres<T>(n) (core/header.h:141-153): malloc'd storage (what CSR and SOL free), exit on failure
template <typename T>
T* memres(size_t n, const std::string& type = "C") {
  T* p = type == "Cpp" ? new T[n] : static_cast<T*>(std::malloc(std::max<size_t>(n, 1) * sizeof(T)));
  if (!p) {
    std::fprintf(stderr, "Error: Null pointer detected in memres\n");
    std::exit(EXIT_FAILURE);
  }
  return p;
}

namespace IO {

// IO::read_ptree(argc, argv) (IO/ptree.hpp:86-119): the parameter file named by argv[1]
inline deco::ptree read_ptree(int argc, char* argv[]) {
  if (argc < 2) throw std::invalid_argument(std::string("Usage: ") + argv[0] + " param.json");
  return deco::read_ptree(argv[1]);
}

// IO::distribute_csr_matrix (IO/cpp_read.hpp:471-577) for one rank: the whole matrix is rank 0's
// block; counts_nr / displs_nr describe that block for the scatter of b
template <template <typename, typename, typename> class TCSR, typename R, typename C, typename V>
std::shared_ptr<TCSR<R, C, V>> distribute_csr_matrix(std::shared_ptr<TCSR<R, C, V>> A, int mid, int np, int bs) {
  if (np != 1 || mid != 0) throw std::invalid_argument("distribute_csr_matrix: the device path runs on one rank");
  if (bs <= 0 || A->n % bs != 0) throw std::invalid_argument("distribute_csr_matrix: n is not a multiple of bs");
  A->np = 1;
  A->mid = 0;
  A->n_glo = A->n;
  A->shift_nr = 0;
  std::free(A->counts_nr);
  std::free(A->displs_nr);
  A->counts_nr = memres<C>(1);
  A->displs_nr = memres<C>(2);
  A->counts_nr[0] = A->n;
  A->displs_nr[0] = 0;
  A->displs_nr[1] = A->n;
  return A;
}

// the ASCII reader (IO/cpp_read.hpp) is tooling outside the solve path: the binary format
// (io.hpp) is the one this library reads
template <typename R, typename C, typename V>
std::shared_ptr<CSR<R, C, V>> cpp_read_ascii_matrix(const std::string&) {
  throw std::invalid_argument("IO.format \"ascii\" is not supported: convert to the binary format");
}
template <typename R, typename C, typename V>
V* cpp_read_ascii_rhs(const std::string&, std::shared_ptr<CSR<R, C, V>>) {
  throw std::invalid_argument("IO.format \"ascii\" is not supported: convert to the binary format");
}

}  // namespace IO
