"""The C++ drop-in facade: the ptree shim (host, no GPU) and the cpr.cpp-style driver (GPU)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CPP = os.path.join(ROOT, "deco_amd", "cpp")
PARAM = os.path.join(ROOT, "tests", "golden", "param_synthetic.json")


def test_ptree_shim_reads_commented_json(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <cstdio>
#include "ptree.hpp"
int main(int argc, char** argv) {
  deco::ptree p = deco::read_ptree(argv[1]);
  const deco::ptree& prec = p.get_child("prec");
  std::printf("%d %s %g %d %s\n", prec.get<int>("block_size"), prec.get<std::string>("weights").c_str(),
              prec.get<double>("drs.e_dd"), p.get<int>("sol.restart"),
              prec.get_child("prec_glo").get<std::string>("type").c_str());
  try { p.get<int>("prec.nope"); } catch (const std::exception&) { std::printf("missing-ok\n"); }
  return 0;
}''')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-I", CPP, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), PARAM], capture_output=True, text=True, check=True).stdout.split()
    assert out == ["2", "quasiimpes_GE", "0.2", "30", "FSAI", "missing-ok"]


def test_prec_layout_key(tmp_path):
    """The "layout" key of the prec tree in the C++ facade and the Python binding: block (the
    default) or csr; anything else is an invalid argument, like an unknown weights type."""
    import deco_amd
    from conftest import small_params
    assert deco_amd.cpr_params(small_params(2)).layout == deco_amd.LAYOUT_BLOCK
    for k, v in (("csr", deco_amd.LAYOUT_CSR), ("block", deco_amd.LAYOUT_BLOCK)):
        assert deco_amd.cpr_params(small_params(2, layout=k)).layout == v
    with pytest.raises(ValueError):
        deco_amd.cpr_params(small_params(2, layout="ell"))
    src = tmp_path / "t.cpp"
    src.write_text(r'''
#include <cstdio>
#include "deco.hpp"
int main() {
  for (const char* k : {"", "csr", "block", "ell"}) {
    deco::ptree t;
    t.put("block_size", "2");
    if (*k) t.put("layout", k);
    try { std::printf("%d\n", deco_facade::params_from(t).layout); }
    catch (const std::invalid_argument&) { std::printf("invalid\n"); }
  }
  return 0;
}''')
    exe = tmp_path / "t"
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CPP, str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()
    assert out == [str(deco_amd.LAYOUT_BLOCK), str(deco_amd.LAYOUT_CSR), str(deco_amd.LAYOUT_BLOCK), "invalid"]


def test_facade_compiles_against_header():
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), "-I",
                    os.path.join(CPP, "compat"), os.path.join(CPP, "cpr_driver.cpp")], check=True)


ONE_RANK_PROBE = r"""
#include "core/pch.h"
#include <cassert>
int main(int argc, char* argv[]) {
  int mid = -1, np = -1;
  init_MPI(&mid, &np, &argc, &argv);                       // cpr.cpp:12
  assert(mid == 0 && np == 1);
  assert(MPI_Comm_set_errhandler(deco_comm, MPI_ERRORS_RETURN) == MPI_SUCCESS);
  boost::property_tree::ptree param = IO::read_ptree(argc, argv);
  assert(param.get<int>("prec.block_size") == 2);
  auto A = std::make_shared<CSR<int, int, double>>(4, 4, 4);
  for (int i = 0; i <= 4; ++i) A->r[i] = i;
  for (int i = 0; i < 4; ++i) { A->c[i] = i; A->v[i] = 1.0 + i; }
  double* b = memres<double>(4);
  for (int i = 0; i < 4; ++i) b[i] = 10.0 * i;
  auto Aloc = IO::distribute_csr_matrix(A, mid, np, 2);   // cpr.cpp:47
  assert(Aloc.get() == A.get() && Aloc->counts_nr[0] == 4 && Aloc->displs_nr[0] == 0);
  double* b_loc = memres<double>(Aloc->n);
  assert(MPI_Scatterv(b, Aloc->counts_nr, Aloc->displs_nr, get_mpi_type<double>(), b_loc, Aloc->n,
                      get_mpi_type<double>(), 0, deco_comm) == MPI_SUCCESS);
  for (int i = 0; i < 4; ++i) assert(b_loc[i] == b[i]);
  assert(MPI_Barrier(deco_comm) == MPI_SUCCESS);
  bool threw = false;
  try { IO::distribute_csr_matrix(A, 0, 2, 2); } catch (const std::invalid_argument&) { threw = true; }
  assert(threw);
  profiler.post_process();
  std::free(b);
  std::free(b_loc);
  assert(MPI_Finalize() == MPI_SUCCESS);
  std::puts("one-rank calls ok");
  return 0;
}
"""


def test_one_rank_calls_of_the_reference_driver(tmp_path):
    """The set-up calls the reference's driver makes around the solve (cpr.cpp:12-13,17,47-53,84,88:
    init_MPI, MPI_Comm_set_errhandler, IO::read_ptree, IO::distribute_csr_matrix, memres,
    MPI_Scatterv with get_mpi_type, MPI_Barrier, profiler.post_process, MPI_Finalize) with the
    reference's arguments, through compat/core/pch.h, in their one-rank meaning; more than one rank
    is an error. Host only (the library is linked, no device call is made)."""
    src = tmp_path / "probe.cpp"
    src.write_text(ONE_RANK_PROBE)
    exe = tmp_path / "probe"
    lib = os.path.join(ROOT, "deco_amd")
    subprocess.run(["g++", "-std=c++17", "-O0", "-I", os.path.join(ROOT, "include"), "-I", os.path.join(CPP, "compat"),
                    str(src), "-o", str(exe), "-L", lib, "-ldeco_hip", f"-Wl,-rpath,{lib}"], check=True)
    res = subprocess.run([str(exe), PARAM], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "one-rank calls ok" in res.stdout, res.stdout + res.stderr


@pytest.mark.gpu
def test_cpp_driver_solves_synthetic_system(gpu):
    exe = os.path.join(CPP, "cpr_driver")
    assert os.path.exists(exe), "build the driver with __graft_entry__.build()"
    res = subprocess.run([exe, PARAM, "synthetic:1"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "iteration count" in res.stdout
    res_line = [l for l in res.stdout.splitlines() if l.startswith("residual")][-1]
    assert float(res_line.split("=")[1]) < 1e-8


def _write_system(tmp_path, g):
    """The system in the reference's binary layout (IO/cpp_read.hpp:331-442): int32 n, m, nz;
    int32 r[n+1]; int32 c[nz]; f64 v[nz]; rhs int32 n; f64 b[n]."""
    import numpy as np
    n, nnz = g["n"], g["nnz"]
    with open(tmp_path / "A.bin", "wb") as f:
        np.array([n, n, nnz], np.int32).tofile(f)
        np.asarray(g["r"], np.int32).tofile(f)
        np.asarray(g["c"], np.int32).tofile(f)
        np.asarray(g["v"], np.float64).tofile(f)
    with open(tmp_path / "b.bin", "wb") as f:
        np.array([n], np.int32).tofile(f)
        np.asarray(g["b"], np.float64).tofile(f)


def _param_with_files(tmp_path, file_a, file_b):
    import json
    from deco_amd import remove_json_comments
    p = json.loads(remove_json_comments(open(PARAM).read()))
    p["IO"] = {"format": "bin", "file_A": str(file_a), "file_b": str(file_b)}
    out = tmp_path / "param.json"
    out.write_text(json.dumps(p))
    return out


@pytest.mark.gpu
def test_cpp_driver_reads_binary_files(gpu, tmp_path):
    """The driver on files in the reference's binary format: the generator's system written with
    numpy converges as it does from memory; rhs "ones" builds b = A 1; a truncated matrix file and
    an rhs of the wrong length are errors."""
    import deco_amd
    exe = os.path.join(CPP, "cpr_driver")
    assert os.path.exists(exe), "build the driver with __graft_entry__.build()"
    g = deco_amd.generate(1)
    _write_system(tmp_path, g)
    for rhs in (tmp_path / "b.bin", "ones"):
        res = subprocess.run([exe, str(_param_with_files(tmp_path, tmp_path / "A.bin", rhs))],
                             capture_output=True, text=True, timeout=120)
        assert res.returncode == 0, res.stdout + res.stderr
        res_line = [l for l in res.stdout.splitlines() if l.startswith("residual")][-1]
        assert float(res_line.split("=")[1]) < 1e-8
    data = (tmp_path / "A.bin").read_bytes()
    (tmp_path / "T.bin").write_bytes(data[: len(data) // 2])
    res = subprocess.run([exe, str(_param_with_files(tmp_path, tmp_path / "T.bin", "ones"))],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode != 0 and "truncated" in (res.stdout + res.stderr)
    import numpy as np
    with open(tmp_path / "s.bin", "wb") as f:
        np.array([7], np.int32).tofile(f)
        np.zeros(7).tofile(f)
    res = subprocess.run([exe, str(_param_with_files(tmp_path, tmp_path / "A.bin", tmp_path / "s.bin"))],
                         capture_output=True, text=True, timeout=120)
    assert res.returncode != 0 and "does not match" in (res.stdout + res.stderr)


@pytest.mark.gpu
def test_binary_writer_reader_round_trip(gpu, tmp_path):
    """IO::cpp_write_binary_matrix / cpp_write_binary_array then the readers: bit-identical."""
    src = tmp_path / "rt.cpp"
    src.write_text(r'''
#include <cstdio>
#include <cstring>
#include "io.hpp"
int main(int argc, char** argv) {
  auto A = std::make_shared<CSR<int, int, double>>(3, 4, 5);
  const int r[] = {0, 2, 3, 5}, c[] = {0, 3, 1, 0, 2};
  const double v[] = {1.5, -2.25, 3e-300, 1e300, -0.0};
  std::memcpy(A->r, r, sizeof r); std::memcpy(A->c, c, sizeof c); std::memcpy(A->v, v, sizeof v);
  const std::string fa = std::string(argv[1]) + "/m.bin", fb = std::string(argv[1]) + "/v.bin";
  IO::cpp_write_binary_matrix(fa, A);
  const double b[] = {0.1, -7.0, 2.5};
  IO::cpp_write_binary_array(fb, 3, b);
  auto B = IO::cpp_read_binary_matrix<int, int, double>(fa);
  double* bb = IO::cpp_read_binary_rhs<int, int, double>(fb, B);
  int ok = B->n == 3 && B->m == 4 && B->nz == 5 && !std::memcmp(B->r, r, sizeof r) &&
           !std::memcmp(B->c, c, sizeof c) && !std::memcmp(B->v, v, sizeof v) && !std::memcmp(bb, b, sizeof b);
  std::free(bb);
  std::printf("%s\n", ok ? "round-trip-ok" : "MISMATCH");
  return ok ? 0 : 1;
}''')
    exe = tmp_path / "rt"
    lib_dir = os.path.join(ROOT, "deco_amd")
    subprocess.run(["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-I", CPP, str(src), "-o", str(exe),
                    "-L", lib_dir, "-ldeco_hip", "-Wl,-rpath," + lib_dir], check=True)
    res = subprocess.run([str(exe), str(tmp_path)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0 and "round-trip-ok" in res.stdout, res.stdout + res.stderr


@pytest.mark.gpu
def test_profiler_records_the_reference_phases_of_a_device_solve(gpu):
    """The phase tree of a CPR setup and solve on the device carries the reference's labels
    (precond/CPR.hpp:47-170, precond/FSAI/FSAI.hpp:42-45, precond/hypre_AMG/hypre_AMG.hpp:27,
    CPR.hpp:240-291, and the driver's "setup & solve" / "Prec" / "Solve", cpr.cpp:57-75) nested as
    the reference nests them, with times consistent with the solve's device-timed duration."""
    import deco_amd
    from conftest import gen_system, small_params
    g = gen_system(1)
    n = g["n"]
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    sol = {"type": "rgmres", "reltol": "1e-8", "restart": "30", "itmax": "300", "verbosity": "0"}
    deco_amd.profile_reset()
    deco_amd.profile_enable(True)
    try:
        deco_amd.profile_start("setup & solve")
        deco_amd.profile_start("Prec")
        M = deco_amd.CPR(small_params(2), A, b)
        deco_amd.profile_stop()
        deco_amd.profile_start("Solve")
        S = deco_amd.SOL(sol, A, M, b)
        deco_amd.profile_stop()
        deco_amd.profile_stop()
    finally:
        deco_amd.profile_enable(False)
    assert S.flag == 0
    rep = deco_amd.profile_report()
    names = [l.split("└── ", 1)[1].rsplit("%", 1)[0].rsplit(None, 1)[0] for l in rep.splitlines()
             if "└" in l]
    depth = {nm: (l.index("└") // 4) for nm, l in zip(names, [l for l in rep.splitlines() if "└" in l])}
    for label in ("setup & solve", "Prec", "decoupling", "quasiimpes weights", "CPD decoup", "Global precond",
                  "Local  precond", "Solve", "FSAI apply", "AMG apply", "CPR_rest"):
        assert label in names, (label, rep)
    assert depth["Prec"] == depth["Solve"] == depth["setup & solve"] + 1
    assert depth["Global precond"] == depth["Local  precond"] == depth["Prec"] + 1
    assert depth["FSAI apply"] > depth["Solve"] and depth["AMG apply"] > depth["Solve"]
    t_solve = deco_amd.profile_time("Solve")
    parts = deco_amd.profile_time("FSAI apply") + deco_amd.profile_time("AMG apply")
    # the host clock around a solve that waits for the device at every phase boundary
    assert t_solve * 1e3 >= 0.9 * S.info.time_ms
    assert 0.0 < parts <= t_solve
    assert deco_amd.profile_time("setup & solve") >= deco_amd.profile_time("Prec") + t_solve * 0.999
    deco_amd.profile_reset()
