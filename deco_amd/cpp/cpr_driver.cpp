// cpr.cpp-style driver (cpr.cpp:8-91) on the device: read parameters, read or generate the
// system, CPR setup, SOL, report. Usage:
//   cpr_driver param.json                (IO.format "bin" with IO.file_A / IO.file_b)
//   cpr_driver param.json synthetic:<cfg> (seeded synthetic system, cfg 1, 11, 2, 3, 4, 5)
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include "deco.hpp"
#include "io.hpp"

using Clock = std::chrono::steady_clock;
static double secs(Clock::time_point a, Clock::time_point b) { return std::chrono::duration<double>(b - a).count(); }

static int run(int argc, char* argv[]) {
  deco_facade::check(deco_init(0), "deco_init");
  deco_ptree param = deco::read_ptree(argv[1]);
  std::shared_ptr<CSR<int, int, double>> A;
  double* b = nullptr;
  std::string src = argc > 2 ? argv[2] : "";
  auto t0 = Clock::now();
  profiler.start("IO");                                                       // cpr.cpp:28-39
  if (src.rfind("synthetic:", 0) == 0) {
    deco_gen* g = nullptr;
    deco_facade::check(deco_gen_create(std::atoi(src.c_str() + 10), 0, &g), "deco_gen_create");
    int n, nnz;
    deco_gen_info(g, &n, &nnz, nullptr, nullptr);
    const int *r, *c;
    const double *v, *bb;
    deco_gen_arrays(g, &r, &c, &v, &bb, nullptr);
    A = std::make_shared<CSR<int, int, double>>(n, n, nnz);
    std::memcpy(A->r, r, sizeof(int) * (n + 1));
    std::memcpy(A->c, c, sizeof(int) * nnz);
    std::memcpy(A->v, v, sizeof(double) * nnz);
    b = (double*)std::malloc(sizeof(double) * n);
    std::memcpy(b, bb, sizeof(double) * n);
    deco_gen_destroy(g);
  } else {
    const std::string fmt = param.get<std::string>("IO.format");
    if (fmt != "bin" && fmt != "binary") throw std::invalid_argument("Unknown global IO::format: " + fmt);
    A = IO::cpp_read_binary_matrix<int, int, double>(param.get<std::string>("IO.file_A"));
    b = IO::cpp_read_binary_rhs<int, int, double>(param.get<std::string>("IO.file_b"), A);
  }
  profiler.stop();
  AD::block_reorder(A, b, param);   // cpr.cpp:41
  auto t1 = Clock::now();
  std::printf("matrix size:\tn = %d, m = %d, nz = %d\n", A->n, A->m, A->nz);
  const std::string glo = param.get<std::string>("prec.prec_glo.type");
  profiler.start("setup & solve");                                            // cpr.cpp:57-79
  profiler.start("Prec");
  std::shared_ptr<CPRBase<double>> P;
  if (glo == "Jacobi") P = std::make_shared<CPR<int, int, double, Jacobi, hypre_AMG>>(param.get_child("prec"), A, b);
  else if (glo == "FSAI") P = std::make_shared<CPR<int, int, double, FSAI, hypre_AMG>>(param.get_child("prec"), A, b);
  else throw std::invalid_argument("Unknown global prec.prec_glo.type: " + glo);
  profiler.stop();
  auto t2 = Clock::now();
  int iter, flag;
  double res;
  profiler.start("Solve");
  {
    SOL<int, int, double> S(param.get_child("sol"), A, P->apply, b, nullptr);
    iter = S.iter;
    flag = S.flag;
    res = S.res_rel;
    std::printf("solve (device) %.3f ms\n", S.info.time_ms);
  }   // SOL frees b and x, as in the reference
  profiler.stop();
  profiler.stop();
  auto t3 = Clock::now();
  std::printf("\n--------------------------------\nProfile of Execution Times (seconds):\n");
  std::printf("└── IO                  %.4fs\n└── Prec                %.4fs\n└── Solve               %.4fs\n",
              secs(t0, t1), secs(t1, t2), secs(t2, t3));
  std::printf("iteration count =  %d\nresidual = %.2e\nflag = %d\n", iter, res, flag);
  profiler.report();
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
