"""Pins the CPU oracle's CPR decoupling, FSAI, Jacobi, V-cycle and Krylov solvers with
invariants, closed forms and independent library results (numpy / scipy). No GPU."""
import numpy as np
import pytest
import scipy.sparse as sp

import deco_amd
import oracle
from conftest import gen_system, to_mat, small_params


def blocks(S, bs, s):
    return S[s * bs:(s + 1) * bs, s * bs:(s + 1) * bs]


def test_quasi_impes_weights_satisfy_their_defining_equations():
    for cfg, custom in [(4, None), (0, (4, 3, 3, 3))]:
        g = gen_system(cfg, custom=custom)
        A = to_mat(g)
        bs = g["bs"]
        D = A.to_scipy().toarray()
        w = oracle.quasi_impes_ge(A, bs)
        wq = oracle.quasi_impes(A, bs)
        for s in range(g["ncells"]):
            B = blocks(D, bs, s)
            ws = w[s * bs:(s + 1) * bs]
            assert ws[0] == 1.0
            # sum_k w_k B[k, m] = 0 for m >= 1 (the GE weights annihilate the saturation columns)
            assert np.allclose(ws @ B[:, 1:], 0.0, atol=1e-12 * np.abs(B).max())
            # quasi-IMPES: B^T w = e_0
            assert np.allclose(B.T @ wq[s * bs:(s + 1) * bs], np.eye(bs)[0], atol=1e-12)


def test_drs_weights_small_closed_case():
    # one cell, bs = 2: row 1's pressure-column diagonal (signed) vs e_dd * off-diagonal sum
    # cells 0,1 coupled; row 1 of cell 0 has a negative pressure-column entry -> fails the test
    r = np.array([0, 4, 8, 12, 16], np.int32)
    c = np.array([0, 1, 2, 3] * 4, np.int32)
    v = np.array([4, 0.1, -1, 0.1,
                  -2, 3, -1, 0.2,
                  -1, 0.1, 4, 0.1,
                  5, 0.3, -1, 3], float)
    A = oracle.Mat(4, 4, r, c, v)
    w = oracle.drs_weights(A, 2, 0.2, 0.0)
    # row 0 and row 2 are pressure rows (forced true); row 1: diag(-2) < 0.2*(|-2|+|-1|-|-2|) -> false
    # row 3: diag(-1 at col 2) ... pressure column of cell 1 is col 2: diag=-1 < 0.2*(5+1-1)=1 -> false
    assert list(w) == [True, False, True, False]
    w2 = oracle.drs_weights(A, 2, -10.0, 0.0)
    assert list(w2) == [True, True, True, True]
    # the e_ps test (CPR.hpp:551-577): a cell's pressure row, summed over the columns of each
    # unknown class without its diagonal, against e_ps * |a_ii|; rows 0 and 2 hold 0.2 in the
    # class of unknown 1 against e_ps * 4, which clears the flags of rows 1 and 3 for e_ps 0.1
    assert list(oracle.drs_weights(A, 2, -10.0, 0.1)) == [True, False, True, False]
    assert list(oracle.drs_weights(A, 2, -10.0, 0.01)) == [True, True, True, True]


@pytest.mark.parametrize("custom", [(4, 4, 3, 2), (4, 3, 3, 3)])
def test_cpr_decoupling_invariants_and_left_transform(custom):
    g = gen_system(0, custom=custom)
    bs = g["bs"]
    A = to_mat(g)
    S0 = A.to_scipy().toarray()
    b0 = g["b"].copy()
    b = g["b"].copy()
    params = deco_amd.cpr_params(small_params(bs, use_drs="1"))
    M = oracle.CPR(params, A, b)
    S1 = A.to_scipy().toarray()
    W = M.left_transform(bs)
    nc = g["ncells"]
    for s in range(nc):
        Ws = W[s * bs * bs:(s + 1) * bs * bs].reshape(bs, bs)
        rows = slice(s * bs, (s + 1) * bs)
        assert np.allclose(Ws @ S0[rows], S1[rows], rtol=1e-12, atol=1e-12 * np.abs(S1[rows]).max())
        assert np.allclose(Ws @ b0[rows], b[rows], rtol=1e-12, atol=1e-12 * np.abs(b[rows]).max())
        # CPD: the diagonal block becomes diagonal (CPR.hpp:603-726)
        B = blocks(S1, bs, s)
        off = B - np.diag(np.diag(B))
        assert np.abs(off).max() <= 1e-12 * np.abs(np.diag(B)).max()
    # the solution is unchanged: (W A) x* = W b
    assert np.allclose(S1 @ g["xstar"], b, rtol=1e-10, atol=1e-10 * np.abs(b).max())
    # pressure matrix = decoupled pressure rows/columns
    Ap = M.pressure_matrix().to_scipy().toarray()
    assert Ap.shape == (nc, nc)


def test_fsai_invariants_and_exact_inverse_on_full_pattern():
    g = gen_system(0, custom=(3, 3, 2, 2))
    A = to_mat(g)
    params = deco_amd.cpr_params(small_params(2, **{"prec_glo.tol_rel": "0", "prec_glo.tol_abs": "0"}))
    F = oracle.fsai(params, A)
    FL, FUt = oracle.fsai_factors(F)
    S = A.to_scipy().toarray()
    L = FL.to_scipy().toarray()
    U = FUt.to_scipy().toarray().T   # FU (lower)
    Dg = L @ S @ U.T
    assert np.allclose(np.diag(Dg), 1.0, atol=1e-12)
    for i in range(A.n):
        P = [j for j in FL.c[FL.r[i]:FL.r[i + 1]] if j != i]
        if P:
            assert np.allclose((L @ S)[i, P], 0.0, atol=1e-12)          # (FL A)[i,P] = 0
            assert np.allclose((S @ U.T)[P, i], 0.0, atol=1e-12)        # (A FU^T)[P,i] = 0
    # a dense matrix with the full lower pattern: FU^T FL = A^-1
    rng = np.random.default_rng(0)
    D = rng.standard_normal((7, 7)) + 7 * np.eye(7)
    Ad = oracle.Mat.from_scipy(sp.csr_matrix(D))
    Fd = oracle.fsai(params, Ad)
    FLd, FUtd = oracle.fsai_factors(Fd)
    inv = FUtd.to_scipy().toarray() @ FLd.to_scipy().toarray()
    assert np.allclose(inv, np.linalg.inv(D), rtol=1e-10, atol=1e-12)
    x = rng.standard_normal(7)
    assert np.allclose(Fd.apply(x), np.linalg.solve(D, x), rtol=1e-10)
    assert np.allclose(Fd.apply_t(x), np.linalg.solve(D.T, x), rtol=1e-10)


def test_jacobi_block_inverse_matches_numpy():
    g = gen_system(0, custom=(3, 2, 2, 3))
    A = to_mat(g)
    J = oracle.jacobi(3, A)
    S = A.to_scipy().toarray()
    x = np.random.default_rng(1).standard_normal(A.n)
    want = np.concatenate([np.linalg.solve(blocks(S, 3, s), x[3 * s:3 * s + 3]) for s in range(A.n // 3)])
    assert np.allclose(J.apply(x), want, rtol=1e-12)
    J1 = oracle.jacobi(1, A)
    assert np.allclose(J1.apply(x), x / np.diag(S), rtol=1e-15)


def laplacian_1d_blocks(n):
    main = 2.0 * np.ones(n)
    off = -np.ones(n - 1)
    return sp.diags([off, main, off], [-1, 0, 1]).tocsr()


def test_host_interpolation_structure():
    n = 30
    S = laplacian_1d_blocks(n)
    S = S.tolil()
    S[0, 0] = 1.0
    S[n - 1, n - 1] = 1.0      # zero row sums everywhere
    S = S.tocsr()
    cf, nc, (Pr, Pc, Pv) = deco_amd.amg_interp_host({"amg_SoC_alpha": "0.25", "amg_max_row_sum": "1.0",
                                                      "amg_prol_trunc_fac": "0", "amg_P_max_elmts": "0"},
                                                     n, S.indptr, S.indices, S.data)
    P = sp.csr_matrix((Pv, Pc, Pr), shape=(n, nc)).toarray()
    assert nc == int((cf == 1).sum()) and 0 < nc < n
    ci = np.cumsum(cf == 1) - 1
    for i in range(n):
        if cf[i] == 1:
            assert np.array_equal(P[i], np.eye(nc)[ci[i]])      # C rows are identity rows
    # zero row sums: interpolation preserves constants
    assert np.allclose(P.sum(axis=1), 1.0, atol=1e-14)
    for i in range(n):
        assert np.all(np.diff(Pc[Pr[i]:Pr[i + 1]]) > 0)


def _solve(typ, A, M, b, **kw):
    sp_ = deco_amd.sol_params(dict(type=typ, reltol=str(kw.get("tol", 1e-10)), restart=str(kw.get("m", 30)),
                                   itmax=str(kw.get("itmax", 500)), verbosity="0"))
    return oracle.solve(sp_, A, M, b, resvec=True)


def test_gmres_exact_termination_and_solution():
    rng = np.random.default_rng(3)
    n = 25
    D = rng.standard_normal((n, n)) + 8 * np.eye(n)
    A = oracle.Mat.from_scipy(sp.csr_matrix(D))
    b = rng.standard_normal(n)
    for typ in ("rgmres", "lgmres"):
        x, info, res = _solve(typ, A, oracle.jacobi(1, A), b, tol=1e-12, m=n + 5)
        assert info.flag == 0 and info.iter <= n
        assert np.allclose(x, np.linalg.solve(D, b), rtol=1e-9)
        # GMRES residual estimates are non-increasing within a cycle
        h = res[:info.iter]
        assert np.all(np.diff(h) <= 1e-12)


def test_bicgstab_pcg_match_direct_solutions():
    rng = np.random.default_rng(4)
    n = 60
    S = laplacian_1d_blocks(n) + sp.eye(n) * 0.5
    A = oracle.Mat.from_scipy(S)
    b = rng.standard_normal(n)
    ref = np.linalg.solve(S.toarray(), b)
    x, info, _ = _solve("bicgstab", A, oracle.jacobi(1, A), b, tol=1e-12)
    assert info.flag == 0 and np.allclose(x, ref, rtol=1e-9)
    x, info, _ = _solve("pcg", A, oracle.jacobi(1, A), b, tol=1e-12)
    assert info.flag == 0 and np.allclose(x, ref, rtol=1e-9)
    # reference quirk: lbicgstab's x update uses A p / A r (lbicgstab.tpp:193,202); its residual
    # recurrence still converges, so flag == 0 while x is not the solution
    x, info, _ = _solve("lbicgstab", A, oracle.jacobi(1, A), b, tol=1e-10)
    assert info.flag == 0


def test_cpr_gmres_with_exact_local_stage_recovers_xstar():
    g = gen_system(4)
    A = to_mat(g)
    b = g["b"].copy()
    M = oracle.CPR(deco_amd.cpr_params(small_params(2)), A, b)
    M.set_local(oracle.dense(M.pressure_matrix()))
    x, info, _ = _solve("rgmres", A, M, b, tol=1e-10, m=30, itmax=200)
    assert info.flag == 0
    assert np.linalg.norm(x - g["xstar"]) / np.linalg.norm(g["xstar"]) < 1e-7
    # transposed CPR is the transpose of the forward operator: <M u, v> = <u, M^T v>
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal(A.n), rng.standard_normal(A.n)
    assert np.isclose(M.apply(u) @ v, u @ M.apply_t(v), rtol=1e-10)


@pytest.mark.parametrize("bs", [2, 3])
def test_fix_diag_zero_diagonals(bs):
    """The tiny-diagonal fix (precond/CPR.hpp:753-855): the swapped rows of a 3-unknown cell are
    part of the left transform (W A0 = A_dec and W b0 = b_dec hold for every other entry), a
    flagged diagonal is 0 again in the decoupled matrix after the setup (CPR.hpp:169,842-855),
    and only the cells that carry a zero get either treatment."""
    from conftest import zero_diag_system, diag_pos
    g, cases = zero_diag_system(bs)
    A = to_mat(g)
    S0 = A.to_scipy().toarray()
    b0 = g["b"].copy()
    b = g["b"].copy()
    M = oracle.CPR(deco_amd.cpr_params(small_params(bs, use_drs="1")), A, b)
    S1 = A.to_scipy().toarray()
    W = M.left_transform(bs)
    flagged = [i for i, what in cases.items() if what == "flag"]
    for i in flagged:
        assert S1[i, i] == 0.0
    for s in range(g["ncells"]):
        Ws = W[s * bs * bs:(s + 1) * bs * bs].reshape(bs, bs)
        rows = slice(s * bs, (s + 1) * bs)
        D = Ws @ S0[rows]
        mask = np.ones_like(D, bool)
        for i in flagged:
            if s * bs <= i < (s + 1) * bs:
                mask[i - s * bs, i] = False
        scale = np.abs(S1[rows]).max()
        assert np.allclose(D[mask], S1[rows][mask], rtol=1e-12, atol=1e-12 * scale), s
        assert np.allclose(Ws @ b0[rows], b[rows], rtol=1e-12, atol=1e-12 * np.abs(b[rows]).max()), s
    # after the swap the saturation rows of the cell trade places, so their decoupled diagonals
    # come from the nonzero pair (i+2, i+1), (i+1, i+2) (scaled by the weights): a missing swap
    # would leave the zero there, the -1e-20 branch a value of that size
    for i, what in cases.items():
        if what == "swap":
            for q in (i, i + 1):
                assert abs(S1[q, q]) > 1e-8 * np.abs(S1[q]).max(), q
