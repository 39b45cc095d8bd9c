"""torch.autograd adjoint of the CPR solve: central finite differences (h = 1e-6) on 64 random
entries of the 16x16x8 two-unknown system, and the full gradient against a direct transposed
solve (scipy sparse LU)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import gen_system, small_params

pytestmark = pytest.mark.gpu


def test_adjoint_gradients_fd_and_direct(gpu):
    import torch
    import deco_amd
    from deco_amd.torch_adjoint import SparsePattern, CPRSolver, cpr_solve

    g = gen_system(4)
    n = g["n"]
    pat = SparsePattern(n, g["r"], g["c"])
    sol = {"type": "rgmres", "reltol": "1e-13", "restart": "60", "itmax": "2000", "verbosity": "0"}
    solver = CPRSolver(pat, small_params(2), sol)
    vals = torch.tensor(g["v"], dtype=torch.float64, device="cuda", requires_grad=True)
    b = torch.tensor(g["b"], dtype=torch.float64, device="cuda", requires_grad=True)
    w = torch.tensor(np.random.default_rng(1004).standard_normal(n), dtype=torch.float64, device="cuda")
    x = cpr_solve(vals, b, solver)
    L = (w * x).sum()
    L.backward()
    dA = vals.grad.cpu().numpy()
    db = b.grad.cpu().numpy()

    S = sp.csr_matrix((g["v"], g["c"], g["r"]), shape=(n, n)).tocsc()
    xs = spla.spsolve(S, g["b"])
    assert np.linalg.norm(x.detach().cpu().numpy() - xs) / np.linalg.norm(xs) < 1e-9
    wn = w.cpu().numpy()
    lam = spla.spsolve(S.T.tocsc(), wn)
    rows = np.repeat(np.arange(n), np.diff(g["r"]))
    full = -lam[rows] * xs[g["c"]]
    assert np.linalg.norm(dA - full) / np.linalg.norm(full) < 1e-8
    assert np.linalg.norm(db - lam) / np.linalg.norm(lam) < 1e-8

    # central finite differences on 64 random entries, exact solves of the perturbed systems
    h = 1e-6
    rng = np.random.default_rng(64)
    picks = rng.choice(g["nnz"], 64, replace=False)
    Scsr = sp.csr_matrix((g["v"], g["c"], g["r"]), shape=(n, n))
    for k in picks:
        vp = g["v"].copy()
        vm = g["v"].copy()
        vp[k] += h
        vm[k] -= h
        Lp = wn @ spla.spsolve(sp.csr_matrix((vp, g["c"], g["r"]), shape=(n, n)).tocsc(), g["b"])
        Lm = wn @ spla.spsolve(sp.csr_matrix((vm, g["c"], g["r"]), shape=(n, n)).tocsc(), g["b"])
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - dA[k]) <= 1e-5 * max(np.abs(full).max(), 1e-300), (k, fd, dA[k])
    del Scsr


def test_adjoint_rejects_mistyped_tensors(gpu):
    """The binding reads values and b as float64 device arrays of the pattern's sizes: other
    dtypes, host tensors and wrong lengths are errors before any library call."""
    import torch
    from deco_amd.torch_adjoint import SparsePattern, CPRSolver, cpr_solve

    g = gen_system(4)
    pat = SparsePattern(g["n"], g["r"], g["c"])
    solver = CPRSolver(pat, small_params(2), {"type": "rgmres", "reltol": "1e-8", "restart": "30",
                                              "itmax": "100", "verbosity": "0"})
    vals = torch.tensor(g["v"], dtype=torch.float64, device="cuda")
    b = torch.tensor(g["b"], dtype=torch.float64, device="cuda")
    with pytest.raises(TypeError):
        cpr_solve(vals.float(), b, solver)
    with pytest.raises(TypeError):
        cpr_solve(vals, b.cpu(), solver)
    with pytest.raises(ValueError):
        cpr_solve(vals[:-1], b, solver)
    x = cpr_solve(vals, b, solver)
    assert x.shape == b.shape and solver.last.flag == 0


def test_adjoint_raises_when_a_solve_does_not_converge(gpu):
    """A forward solve that stops at itmax raises instead of returning x; so does a transposed
    solve in the backward pass that stops at itmax (the forward one converged)."""
    import torch
    from deco_amd.torch_adjoint import SparsePattern, CPRSolver, cpr_solve, NotConverged

    g = gen_system(4)
    pat = SparsePattern(g["n"], g["r"], g["c"])
    vals = torch.tensor(g["v"], dtype=torch.float64, device="cuda", requires_grad=True)
    b = torch.tensor(g["b"], dtype=torch.float64, device="cuda")
    tight = {"type": "rgmres", "reltol": "1e-12", "restart": "30", "itmax": "1", "verbosity": "0"}
    solver = CPRSolver(pat, small_params(2), tight)
    with pytest.raises(NotConverged) as e:
        cpr_solve(vals, b, solver)
    assert e.value.info.flag == 1 and e.value.info.iter == 1
    solver = CPRSolver(pat, small_params(2), dict(tight, itmax="300"))
    x = cpr_solve(vals, b, solver)
    assert solver.last.flag == 0
    solver.sol["itmax"] = "1"                       # the backward's transposed solve stops early
    with pytest.raises(NotConverged) as e:
        x.sum().backward()
    assert e.value.info.flag == 1 and solver.last_adjoint.flag == 1
