"""The strongly anisotropic case (lognormal permeability sigma 3, thin layers with kz/kx 1e-3, two
unknowns per cell) with the FSAI pattern taken from A^2 (power 2, device SpGEMM) and the device
Galerkin products.

Reduced size (the same physics on a 16 x 15 x 33 grid): the device against the oracle -- decoupled
system bit for bit, FSAI power-2 pattern bit for bit and factors to 1e-13, the AMG hierarchy's
setup products against the oracle's own, and CPR-GMRES iterations within 2 and solutions within
1e-6 at the same tolerance.

Full size (64 x 118 x 263 cells, n = 3,972,352, nnz = 55,169,584, the structure of the reference's
Sleipner system): properties that do not depend on the size -- the generator's known answer
A x* = b, convergence of the FP64 and FP32-preconditioned solves on the decoupled system and of
their solutions on the original one, their agreement, and the adjoint: the transposed solve's
lambda = W^T mu satisfies A^T lambda = w, and dL/dA is -lambda_i x_j on A's pattern."""
import numpy as np
import pytest

import deco_amd
import oracle
from conftest import to_mat, oracle_amg
from bench import PREC_TREE

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def tree5():
    t = dict(PREC_TREE, block_size="2")
    t["prec_glo"] = dict(t["prec_glo"], power="2")
    return deco_amd.PTree(t)


SOL = {"type": "rgmres", "reltol": "1e-8", "restart": "50", "itmax": "500", "verbosity": "0"}


def test_scaled_anisotropic_against_oracle(gpu):
    g = deco_amd.generate(5, grid=(16, 15, 33))
    n = g["n"]
    tree = tree5()
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(tree, A, b)
    Ao = to_mat(g)
    bo = g["b"].copy()
    Mo = oracle.CPR(deco_amd.cpr_params(tree), Ao, bo)
    r, c, v = A.download()
    assert np.array_equal(r, Ao.r) and np.array_equal(c, Ao.c) and np.array_equal(v, Ao.v)
    assert np.array_equal(b.numpy(), bo)
    # FSAI on the A^2 pattern
    FL, FUt = M.fsai().factors()
    FLo, FUto = oracle.fsai_factors(Mo)
    for F, Fo in ((FL, FLo), (FUt, FUto)):
        fr, fc, fv = F.download()
        assert np.array_equal(fr, Fo.r) and np.array_equal(fc, Fo.c[:Fo.nnz])
        assert rel(fv, Fo.v[:Fo.nnz]) <= 1e-13
    assert FL.info()[2] > 2 * (M.pressure_matrix().info()[2])      # the power-2 pattern is wider
    # AMG hierarchy against the oracle's own setup products, then the solves
    Go, levels, _ = oracle_amg(M.amg(), Mo.pressure_matrix(), tree)
    assert len(levels) >= 2
    Mo.set_local(Go)
    xo, info = oracle.solve(deco_amd.sol_params(SOL), Ao, Mo, bo)
    S = deco_amd.SOL(SOL, A, M, b)
    assert S.flag == 0 and info.flag == 0 and abs(S.iter - info.iter) <= 2
    assert rel(S.x.numpy(), xo) <= 1e-6
    M.set_precision(deco_amd.PREC_FP32)
    S32 = deco_amd.SOL(SOL, A, M, b)
    M.set_precision(deco_amd.PREC_FP64)
    assert S32.flag == 0 and abs(S32.iter - info.iter) <= 2
    print(f"\nanisotropic 16x15x33: iterations device {S.iter} / FP32 {S32.iter} / oracle {info.iter}, "
          f"x vs oracle {rel(S.x.numpy(), xo):.2e}, FP32 vs FP64 {rel(S32.x.numpy(), S.x.numpy()):.2e}")


@pytest.fixture(scope="module")
def big5(gpu):
    g = deco_amd.generate(5)
    assert g["n"] == 3_972_352 and g["nnz"] == 55_169_584
    return g


def test_fullsize_anisotropic_solves_and_adjoint(big5):
    g = big5
    n = g["n"]
    A = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])
    A0 = deco_amd.CSR(n, n, g["r"], g["c"], g["v"])      # the system before decoupling
    assert rel(A0.apply(g["xstar"]).numpy(), g["b"]) <= 1e-14
    assert rel(A0.apply(g["xstar"], block=(2, 2)).numpy(), g["b"]) <= 1e-14
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(tree5(), A, b)
    out = {}
    for prec, key in ((deco_amd.PREC_FP64, "fp64"), (deco_amd.PREC_FP32, "fp32")):
        M.set_precision(prec)
        for tol in ("1e-8", "1e-11"):
            S = deco_amd.SOL(dict(SOL, reltol=tol), A, M, b)
            assert S.flag == 0 and S.res_rel <= float(tol), (key, tol)
            out[key, tol] = (S.iter, S.x.numpy())
    M.set_precision(deco_amd.PREC_FP64)
    for key in ("fp64", "fp32"):
        for tol in ("1e-8", "1e-11"):
            x = out[key, tol][1]
            # W (A x - b) is small on the decoupled system; on the original one up to W's condition
            assert rel(A0.apply(x).numpy(), g["b"]) <= 1e3 * float(tol), (key, tol)
    assert abs(out["fp64", "1e-8"][0] - out["fp32", "1e-8"][0]) <= 2
    d8 = rel(out["fp32", "1e-8"][1], out["fp64", "1e-8"][1])
    d11 = rel(out["fp32", "1e-11"][1], out["fp64", "1e-11"][1])
    e8 = rel(out["fp64", "1e-8"][1], out["fp64", "1e-11"][1])
    print(f"\nanisotropic full size: iterations FP64 {out['fp64', '1e-8'][0]} / FP32 {out['fp32', '1e-8'][0]} "
          f"(rtol 1e-8), {out['fp64', '1e-11'][0]} / {out['fp32', '1e-11'][0]} (1e-11); FP32 vs FP64 "
          f"{d8:.2e} at 1e-8, {d11:.2e} at 1e-11; FP64 at 1e-8 vs 1e-11 {e8:.2e}")
    # at rtol 1e-8 the two precisions differ by less than either differs from the tighter solve;
    # at 1e-11 they agree to 1e-6
    assert d8 <= e8
    assert d11 <= 1e-6
    # adjoint: (W A)^T mu = w, lambda = W^T mu solves A^T lambda = w; dL/dA = -lambda_i x_j
    w = np.random.default_rng(1005).standard_normal(n)
    St = deco_amd.SOL(SOL, A, M, w, transpose=True)
    assert St.flag == 0 and St.res_rel <= 1e-8
    lam = M.transform_transpose(St.x)
    assert rel(A0.apply_transpose(lam).numpy(), w) <= 1e-7
    x = deco_amd.DeviceArray.from_numpy(out["fp64", "1e-8"][1])
    dA = deco_amd.grad_A_pattern(A0, lam, x).numpy()
    k = np.random.default_rng(7).choice(g["nnz"], 4096, replace=False)
    rows = np.searchsorted(g["r"], k, side="right") - 1
    ln, xn = lam.numpy(), x.numpy()
    assert np.array_equal(dA[k], -ln[rows] * xn[g["c"][k]])
