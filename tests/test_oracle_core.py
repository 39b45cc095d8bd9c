"""Pins the CPU oracle's kernels against closed forms, brute force and scipy (no GPU)."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle
from conftest import gen_system, to_mat


def rand_csr(n, m, density, seed, diag=True):
    rng = np.random.default_rng(seed)
    S = sp.random(n, m, density=density, random_state=rng, format="csr")
    S.data = rng.standard_normal(S.nnz)
    if diag:
        S = S + sp.eye(n, m) * 4.0
    S = S.tocsr()
    S.sort_indices()
    return S


def test_spmv_matches_dense_and_strided_rows():
    S = rand_csr(40, 40, 0.2, 1)
    A = oracle.Mat.from_scipy(S)
    x = np.random.default_rng(2).standard_normal(40)
    y = oracle.spmv(A, x)
    assert np.allclose(y, S.toarray() @ x, rtol=1e-14, atol=1e-13)
    y2 = oracle.spmv(A, x, bs=2)  # rows 0, 2, 4, ... (core/alg.hpp:55,62)
    assert np.allclose(y2, (S.toarray() @ x)[::2], rtol=1e-14, atol=1e-13)


def test_dot_nrm2_closed_forms():
    n = 1000
    a = np.arange(1, n + 1, dtype=np.float64)
    assert oracle.dot(a, a) == n * (n + 1) * (2 * n + 1) / 6
    assert oracle.nrm2(np.array([3.0, 4.0])) == 5.0


def brute_filter(S, tol, proper):
    D = S.toarray()
    keep = []
    for i in range(D.shape[0]):
        row = S.indices[S.indptr[i]:S.indptr[i + 1]]
        thr = abs(D[i, i] * tol) if proper else tol
        keep.append([j for j in row if abs(D[i, j]) > thr or j == i])
    return keep


@pytest.mark.parametrize("proper", [False, True])
def test_filters_match_brute_force(proper):
    S = rand_csr(60, 60, 0.15, 3)
    A = oracle.Mat.from_scipy(S)
    tol = 0.3 if proper else 0.5
    (oracle.filter_proper if proper else oracle.filter)(A, tol)
    want = brute_filter(S, tol, proper)
    for i in range(60):
        got = list(A.c[A.r[i]:A.r[i + 1]])
        assert got == want[i]
        assert np.array_equal(A.v[A.r[i]:A.r[i + 1]], S.toarray()[i, want[i]])


def test_local_matrix_brute_force():
    g = gen_system(0, custom=(4, 3, 2, 3))
    A = to_mat(g)
    bs = 3
    L = oracle.local_matrix(A, bs, 0)
    D = A.to_scipy().toarray()
    Dp = D[0::bs, 0::bs]
    # pattern: entries of pressure rows in pressure columns (structural, explicit zeros kept)
    for i in range(L.n):
        cols = [c // bs for c in A.c[A.r[bs * i]:A.r[bs * i + 1]] if c % bs == 0]
        assert list(L.c[L.r[i]:L.r[i + 1]]) == cols
        assert np.array_equal(L.v[L.r[i]:L.r[i + 1]], Dp[i, cols])


def cell_perm(n, bs, order):
    """new index of every old unknown: position i of a cell receives unknown order[i]-1."""
    reord = np.empty(bs, np.int64)
    for i, ch in enumerate(order[:bs]):
        reord[int(ch) - 1] = i
    old = np.arange(n)
    return (old // bs) * bs + reord[old % bs]


@pytest.mark.parametrize("bs,ro,co", [(3, "213", "213"), (3, "312", "132"), (2, "21", "12"), (3, "123", "123")])
def test_block_reorder_is_the_cell_permutation(bs, ro, co):
    """AD::block_reorder against permutation matrices: A' = P_r A P_c^T, b' = P_r b, rows sorted,
    explicit zeros kept."""
    g = gen_system(0, custom=(5, 4, 3, bs))
    n = g["n"]
    A = to_mat(g)
    Ar, br = oracle.block_reorder(A, g["b"], bs, ro, co)
    pr, pc = cell_perm(n, bs, ro), cell_perm(n, bs, co)
    S = A.to_scipy().tocoo()
    E = sp.csr_matrix((S.data, (pr[S.row], pc[S.col])), shape=(n, n))
    E.sort_indices()
    assert np.array_equal(Ar.r, E.indptr) and np.array_equal(Ar.c[:Ar.nnz], E.indices)
    assert np.array_equal(Ar.v[:Ar.nnz], E.data)
    bexp = np.empty(n)
    bexp[pr] = g["b"]
    assert np.array_equal(br, bexp)
    # the permuted system has the permuted solution: x'[pc] = x*
    xs = np.empty(n)
    xs[pc] = g["xstar"]
    assert np.allclose(E @ xs, bexp, rtol=1e-12, atol=1e-12 * np.abs(bexp).max())


def test_block_reorder_identity_and_errors():
    g = gen_system(0, custom=(4, 3, 2, 2))
    A = to_mat(g)
    Ar, br = oracle.block_reorder(A, g["b"], 2, "12", "12")          # csr2coo.cpp:40-41 early return
    assert np.array_equal(Ar.v, A.v) and np.array_equal(br, g["b"])
    Ar, br = oracle.block_reorder(A, g["b"], 1, "1", "1")            # bs 1 early return
    assert np.array_equal(Ar.c, A.c)
    with pytest.raises(RuntimeError):
        oracle.block_reorder(A, g["b"], 2, "11", "12")
    with pytest.raises(RuntimeError):
        oracle.block_reorder(A, g["b"], 2, "1", "12")


def test_transpose_matches_scipy():
    S = rand_csr(50, 37, 0.2, 4, diag=False)
    T = oracle.transpose(oracle.Mat.from_scipy(S))
    ST = S.T.tocsr()
    ST.sort_indices()
    assert np.array_equal(T.r, ST.indptr) and np.array_equal(T.c, ST.indices)
    assert np.array_equal(T.v, ST.data)


def test_spgemm_matches_scipy_and_boolean_product():
    A = rand_csr(30, 25, 0.2, 5, diag=False)
    B = rand_csr(25, 40, 0.2, 6, diag=False)
    C = oracle.spgemm(oracle.Mat.from_scipy(A), oracle.Mat.from_scipy(B))
    # structural pattern = boolean product (no numerical dropping)
    P = (abs(A).astype(bool).astype(int) @ abs(B).astype(bool).astype(int)).tocsr()
    P.sort_indices()
    assert np.array_equal(C.r, P.indptr) and np.array_equal(C.c, P.indices)
    ref = (A @ B).toarray()
    assert np.allclose(C.to_scipy().toarray(), ref, rtol=1e-14, atol=1e-14)


def test_fsai_pattern_power1_is_lower_triangle():
    S = rand_csr(40, 40, 0.2, 7)
    F = oracle.fsai_pattern(oracle.Mat.from_scipy(S), 1)
    T = sp.tril(S).tocsr()
    T.sort_indices()
    assert np.array_equal(F.r, T.indptr) and np.array_equal(F.c, T.indices)


def test_fsai_pattern_power2_brute_force_and_stencil_closed_form():
    g = gen_system(0, custom=(6, 6, 6, 2))
    A = to_mat(g)
    F = oracle.fsai_pattern(A, 2)
    B = abs(A.to_scipy()).astype(bool).astype(int)
    P = sp.tril(B @ B).tocsr()
    P.sort_indices()
    assert np.array_equal(F.r, P.indptr) and np.array_equal(F.c, P.indices)
    # interior cell of a 7-point grid: A^2 couples the 25 cells with |d|_1 <= 2, so the lower
    # part of row (cell, k) has 12 cells * bs columns plus k+1 of its own cell.
    nx = 6
    cell = 3 + nx * (3 + nx * 3)
    for k in range(2):
        i = cell * 2 + k
        assert F.r[i + 1] - F.r[i] == 12 * 2 + k + 1


def test_lu_matches_numpy_both_transposes():
    rng = np.random.default_rng(8)
    for k in (1, 2, 5, 9):
        M = rng.standard_normal((k, k)) + np.eye(k)
        b = rng.standard_normal(k)
        lu, piv, info = oracle.lu_factor(M)
        assert info == 0
        assert np.allclose(oracle.lu_solve(lu, piv, b, 0), np.linalg.solve(M, b), rtol=1e-12)
        assert np.allclose(oracle.lu_solve(lu, piv, b, 1), np.linalg.solve(M.T, b), rtol=1e-12)
    # singular: zero column reports the 1-based column index (dgetf2 info)
    M = np.array([[1.0, 0.0], [2.0, 0.0]])
    assert oracle.lu_factor(M)[2] == 2


@pytest.mark.parametrize("cfg,n,nnz", [(1, 32768, 442368), (4, 4096, 53248)])
def test_generator_sizes_and_structure(cfg, n, nnz):
    g = gen_system(cfg)
    assert g["n"] == n and g["nnz"] == nnz
    r, c, v = g["r"], g["c"], g["v"]
    bs = g["bs"]
    for i in range(0, n, max(1, n // 97)):
        row = c[r[i]:r[i + 1]]
        assert np.all(np.diff(row) > 0) and i in row
        s = i // bs
        first = c[r[s * bs]:r[s * bs + 1]]
        assert np.array_equal(row, first)          # all rows of a cell share one column set
    A = to_mat(g)
    assert np.allclose(oracle.spmv(A, g["xstar"]), g["b"], rtol=0, atol=0)  # b = A x*
