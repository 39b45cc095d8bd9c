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


def test_facade_compiles_against_header():
    subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                    os.path.join(CPP, "cpr_driver.cpp")], check=True)


@pytest.mark.gpu
def test_cpp_driver_solves_synthetic_system(gpu):
    exe = os.path.join(CPP, "cpr_driver")
    assert os.path.exists(exe), "build the driver with __graft_entry__.build()"
    res = subprocess.run([exe, PARAM, "synthetic:1"], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "iteration count" in res.stdout
    res_line = [l for l in res.stdout.splitlines() if l.startswith("residual")][-1]
    assert float(res_line.split("=")[1]) < 1e-8
