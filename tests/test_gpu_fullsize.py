"""GPU checks at the full storage-aquifer size (168x100x120 cells, 2 unknowns per cell: n =
4,032,000, nnz = 56,056,320) through properties that do not depend on the size: the generator's
known answer A x* = b, row sums, the in-cell reorder as an involution, convergence of the CPR
solves on the decoupled system and of their solutions on the original one, FP32/FP64 agreement
and bitwise reproducibility."""
import numpy as np
import pytest

import deco_amd

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# the bench's CPR parameters (the reference's param_sleipner.json "prec" set)
from bench import PREC_TREE as PREC  # noqa: E402


@pytest.fixture(scope="module")
def big(gpu):
    g = deco_amd.generate(3)
    assert g["n"] == 4_032_000 and g["nnz"] == 56_056_320
    return g


def test_fullsize_products(big):
    g = big
    n, bs = g["n"], g["bs"]
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    assert rel(A.apply(g["xstar"]).numpy(), g["b"]) <= 1e-14
    assert rel(A.apply(g["xstar"], block=(bs, bs)).numpy(), g["b"]) <= 1e-14
    rows = np.add.reduceat(g["v"], g["r"][:-1])          # every row has its diagonal: no empty rows
    assert rel(A.apply(np.ones(n)).numpy(), rows) <= 1e-14
    assert rel(A.apply(np.ones(n), block=(bs, bs)).numpy(), rows) <= 1e-14
    y32 = A.apply_f32(g["xstar"].astype(np.float32), block=(bs, bs)).numpy().astype(np.float64)
    assert rel(y32, g["b"]) <= 1e-5


def test_fullsize_reorder_involution(big):
    g = big
    n = g["n"]
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    A.block_reorder(2, "21", "21", b)
    A.block_reorder(2, "21", "21", b)
    r, c, v = A.download()
    assert np.array_equal(r, g["r"]) and np.array_equal(c, g["c"]) and np.array_equal(v, g["v"])
    assert np.array_equal(b.numpy(), g["b"])


def test_fullsize_cpr_solves(big):
    g = big
    n = g["n"]
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(deco_amd.PTree(PREC), A, b)          # decouples A and b in place
    sol = {"type": "rgmres", "reltol": "1e-8", "restart": "50", "itmax": "300", "verbosity": "0"}
    S64 = deco_amd.SOL(sol, A, M, b)
    x64 = S64.x.numpy()
    M.set_precision(deco_amd.PREC_FP32)
    S32 = deco_amd.SOL(sol, A, M, b)
    x32 = S32.x.numpy()
    S32b = deco_amd.SOL(sol, A, M, b)
    M.set_precision(deco_amd.PREC_FP64)
    assert S64.flag == 0 and S32.flag == 0
    assert S64.res_rel <= 1e-8 and S32.res_rel <= 1e-8
    assert abs(S64.iter - S32.iter) <= 2
    assert rel(x32, x64) <= 1e-6
    assert np.array_equal(S32b.x.numpy(), x32)            # deterministic reductions
    # the left transforms keep the solution: x solves the original system too
    A0 = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    for x in (x64, x32):
        assert rel(A0.apply(x).numpy(), g["b"]) <= 1e-7
        assert rel(x, g["xstar"]) <= 1e-2
