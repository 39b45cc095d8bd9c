// A reference-style driver (the flow of cpr.cpp:8-91) on libdeco_hip.so through the compat header
// (compat/core/pch.h): MPI set-up on one rank, read the system (binary files, or a seeded
// synthetic system), in-cell reorder, row distribution and scatter of b, CPR setup, SOL, profile
// report. Usage:
//   cpr_driver param.json                (IO.format "bin" with IO.file_A / IO.file_b)
//   cpr_driver param.json synthetic:<cfg> (seeded synthetic system, cfg 1, 11, 2, 3, 4, 5)
#include <cstdio>
#include <cstring>
#include <string>
#include "core/pch.h"

using R = int;
using C = int;
using V = double;

// a seeded synthetic system as host CSR arrays and right-hand side (the generator of the tests)
static std::shared_ptr<CSR<R, C, V>> synthetic(int cfg, V*& b) {
  deco_gen* g = nullptr;
  deco_facade::check(deco_gen_create(cfg, 0, &g), "deco_gen_create");
  int n, nnz;
  deco_gen_info(g, &n, &nnz, nullptr, nullptr);
  const int *r, *c;
  const double *v, *bb;
  deco_gen_arrays(g, &r, &c, &v, &bb, nullptr);
  auto A = std::make_shared<CSR<R, C, V>>(n, n, nnz);
  std::memcpy(A->r, r, sizeof(R) * (n + 1));
  std::memcpy(A->c, c, sizeof(C) * nnz);
  std::memcpy(A->v, v, sizeof(V) * nnz);
  b = memres<V>(n);
  std::memcpy(b, bb, sizeof(V) * n);
  deco_gen_destroy(g);
  return A;
}

static int run(int argc, char* argv[]) {
  int mid = 0, nprocs = 1;
  init_MPI(&mid, &nprocs, &argc, &argv);
  MPI_Comm_set_errhandler(deco_comm, MPI_ERRORS_RETURN);
  deco_facade::check(deco_init(0), "deco_init");
  int flag = 1;
  {
    boost::property_tree::ptree param = IO::read_ptree(argc, argv);
    auto A = std::make_shared<CSR<R, C, V>>();
    V* b = nullptr;
    V* x = nullptr;
    const std::string src = argc > 2 ? argv[2] : "";
    const auto glo = param.get<std::string>("prec.prec_glo.type");
    if (mid == 0) {
      profiler.start("IO");
      if (src.rfind("synthetic:", 0) == 0) {
        A = synthetic(std::atoi(src.c_str() + 10), b);
      } else {
        const auto format = param.get<std::string>("IO.format");
        if (format == "ascii") {
          A = IO::cpp_read_ascii_matrix<R, C, V>(param.get<std::string>("IO.file_A"));
          b = IO::cpp_read_ascii_rhs<R, C, V>(param.get<std::string>("IO.file_b"), A);
        } else if (format == "bin" || format == "binary") {
          A = IO::cpp_read_binary_matrix<R, C, V>(param.get<std::string>("IO.file_A"));
          b = IO::cpp_read_binary_rhs<R, C, V>(param.get<std::string>("IO.file_b"), A);
        } else {
          throw std::invalid_argument("Unknown global IO::format: " + format);
        }
      }
      profiler.stop();
      AD::block_reorder(A, b, param);
    }
    MPI_Barrier(deco_comm);
    auto Aloc = IO::distribute_csr_matrix(A, mid, nprocs, param.get<int>("prec.block_size"));
    V* b_loc = memres<V>(Aloc->n);
    MPI_Scatterv(b, Aloc->counts_nr, Aloc->displs_nr, get_mpi_type<V>(), b_loc, Aloc->n, get_mpi_type<V>(), 0,
                 deco_comm);
    std::free(b);
    std::printf("matrix size:\tn = %d, m = %d, nz = %d\n", Aloc->n, Aloc->m, Aloc->nz);
    MPI_Barrier(deco_comm);
    profiler.start("setup & solve");
    profiler.start("Prec");
    Aloc->setup_spmv_v2();
    std::shared_ptr<CPRBase<V>> P;
    if (glo == "Jacobi") P = std::make_shared<CPR<R, C, V, Jacobi, hypre_AMG>>(param.get_child("prec"), Aloc, b_loc);
    else if (glo == "FSAI") P = std::make_shared<CPR<R, C, V, FSAI, hypre_AMG>>(param.get_child("prec"), Aloc, b_loc);
    else throw std::invalid_argument("Unknown global prec.prec_glo.type: " + glo);
    MPI_Barrier(deco_comm);
    profiler.stop();
    profiler.start("Solve");
    {
      SOL<R, C, V> S(param.get_child("sol"), Aloc, P->apply, b_loc, x);   // SOL frees b_loc and x
      flag = S.flag;
      std::printf("iteration count =  %d\nresidual = %.2e\nflag = %d\nsolve (device) %.3f ms\n", S.iter, S.res_rel,
                  S.flag, S.info.time_ms);
    }
    MPI_Barrier(deco_comm);
    profiler.stop();
    profiler.stop();
    profiler.post_process();
    if (mid == 0) profiler.report();
  }
  MPI_Finalize();
  return flag;
}

// Errors the reference's driver would end on (unreadable files, bad parameters, numerical
// failures reported by the library) print their message and return 2 here.
int main(int argc, char* argv[]) {
  if (argc < 2) {
    std::fprintf(stderr, "Usage: %s param.json [synthetic:<cfg>]\n", argv[0]);
    return 1;
  }
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "cpr_driver: %s\n", e.what());
    return 2;
  }
}
