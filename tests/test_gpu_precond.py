"""GPU parity of SpGEMM, FSAI, the CPR setup and apply, and the AMG V-cycle against the CPU
oracle (through the C ABI)."""
import numpy as np
import pytest

import deco_amd
import oracle
from conftest import gen_system, to_mat, small_params, oracle_amg

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def dev_csr(g):
    return deco_amd.CSR(g["n"], g["n"], g["r"], g["c"], g["v"])


def same_pattern(gpu_csr, mat):
    r, c, v = gpu_csr.download()
    return np.array_equal(r, mat.r) and np.array_equal(c, mat.c[:mat.nnz]), v


@pytest.mark.parametrize("cfg", [1, 4])
def test_spgemm_and_power2_pattern_bitexact(gpu, cfg):
    g = gen_system(cfg)
    A = dev_csr(g)
    M = to_mat(g)
    C = A @ A
    Co = oracle.spgemm(M, M)
    ok, v = same_pattern(C, Co)
    assert ok
    assert np.array_equal(v, Co.v)          # same (j,k) accumulation order
    F = A.tril_square_pattern()
    Fo = oracle.fsai_pattern(M, 2)
    assert same_pattern(F, Fo)[0]


def test_spgemm_rectangular(gpu):
    import scipy.sparse as sp
    rng = np.random.default_rng(0)
    S1 = sp.random(300, 200, density=0.05, random_state=rng, format="csr")
    S2 = sp.random(200, 150, density=0.05, random_state=rng, format="csr")
    C = deco_amd.CSR.from_scipy(S1) @ deco_amd.CSR.from_scipy(S2)
    Co = oracle.spgemm(oracle.Mat.from_scipy(S1), oracle.Mat.from_scipy(S2))
    ok, v = same_pattern(C, Co)
    assert ok and np.array_equal(v, Co.v)


@pytest.mark.parametrize("cfg,power", [(1, 1), (2, 1), (4, 2)])
def test_fsai_factors_parity(gpu, cfg, power):
    g = gen_system(cfg)
    bs = g["bs"]
    A = dev_csr(g)
    M = to_mat(g)
    tree = small_params(bs, **{"prec_glo.power": str(power)})
    F = deco_amd.FSAI(tree, A)
    Fo = oracle.fsai(deco_amd.cpr_params(tree), M)
    FL, FUt = F.factors()
    FLo, FUto = oracle.fsai_factors(Fo)
    okL, vL = same_pattern(FL, FLo)
    okU, vU = same_pattern(FUt, FUto)
    assert okL and okU
    # same LAPACK-order LU: equal up to the last bits
    assert rel(vL, FLo.v) <= 1e-13 and rel(vU, FUto.v) <= 1e-13
    x = np.random.default_rng(1).uniform(-1, 1, g["n"])
    y, yt = F.apply(x).numpy(), F.apply_t(x).numpy()
    assert rel(y, Fo.apply(x)) <= 1e-13
    assert rel(yt, Fo.apply_t(x)) <= 1e-13
    # FP32 factors (cell blocks), FP64 vectors in and out
    F.set_precision(deco_amd.PREC_FP32)
    y32, yt32 = F.apply(x).numpy(), F.apply_t(x).numpy()
    F.set_precision(deco_amd.PREC_FP64)
    assert rel(y32, y) <= 1e-5 and rel(yt32, yt) <= 1e-5
    assert not np.array_equal(y32, y)                     # the FP32 path did run


@pytest.mark.parametrize("cfg", [4, 2])
def test_cpr_setup_bitexact_and_apply_parity(gpu, cfg):
    g = gen_system(cfg)
    bs = g["bs"]
    tree = small_params(bs)
    A = dev_csr(g)
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(tree, A, b)
    Mo_A = to_mat(g)
    bo = g["b"].copy()
    Mo = oracle.CPR(deco_amd.cpr_params(tree), Mo_A, bo)
    # decoupled system and pressure matrix (index maps) identical
    r, c, v = A.download()
    assert np.array_equal(r, Mo_A.r) and np.array_equal(c, Mo_A.c) and np.array_equal(v, Mo_A.v)
    assert np.array_equal(b.numpy(), bo)
    Ap = M.pressure_matrix()
    Apo = Mo.pressure_matrix()
    ok, vp = same_pattern(Ap, Apo)
    assert ok and np.array_equal(vp, Apo.v)
    assert np.array_equal(M.left_transform(), Mo.left_transform(bs))
    # AMG: the device hierarchy's setup products against the oracle's own (from P_l and the C/F
    # splittings only), and the V-cycle
    G = M.amg()
    Go, levels, cinv = oracle_amg(G, Apo, tree)
    assert len(levels) >= 1
    xp = np.random.default_rng(2).uniform(-1, 1, Ap.n)
    assert rel(G.apply(xp).numpy(), Go.apply(xp)) <= 1e-13
    assert rel(G.apply_t(xp).numpy(), Go.apply_t(xp)) <= 1e-12
    # CPR apply with the oracle's AMG
    Go2, _, _ = oracle_amg(G, Apo, tree, check=False)
    Mo.set_local(Go2)
    x = np.random.default_rng(3).uniform(-1, 1, g["n"])
    assert rel(M.apply(x).numpy(), Mo.apply(x)) <= 1e-12
    assert rel(M.apply_t(x).numpy(), Mo.apply_t(x)) <= 1e-11
    # FP32 preconditioner: same operator to single precision
    M.set_precision(deco_amd.PREC_FP32)
    assert rel(M.apply(x).numpy(), Mo.apply(x)) <= 2e-4
    M.set_precision(deco_amd.PREC_FP64)


@pytest.mark.parametrize("seq,relax", [(0, 18), (1, 18), (2, 18), (3, 17), (0, 17), (3, -1), (2, -1), (3, 12)])
def test_cpr_sequences(gpu, seq, relax):
    """CPR sequences 0-3 against the oracle. With a Jacobi-type or two-stage Gauss-Seidel smoother
    (18, -1, 12) the pressure stage runs fused (the residual epilogue writes the first sweep's
    b/d, the last sweep adds into x); with F-C-F half-sweeps (17) it runs as separate kernels.
    FP32 against FP64 on the device."""
    g = gen_system(4)
    tree = small_params(2, seq=str(seq), **{"prec_loc.amg_relax_type": relax})
    A = dev_csr(g)
    deco_amd.set_amg_dense_tail(0)            # the level-by-level cycle (2048 pressure rows)
    try:
        M = deco_amd.CPR(tree, A, None)
    finally:
        deco_amd.set_amg_dense_tail(4608)
    Mo_A = to_mat(g)
    Mo = oracle.CPR(deco_amd.cpr_params(tree), Mo_A, None)
    Go, _, _ = oracle_amg(M.amg(), Mo.pressure_matrix(), tree)
    Mo.set_local(Go)
    x = np.random.default_rng(seq).uniform(-1, 1, g["n"])
    y = M.apply(x).numpy()
    assert rel(y, Mo.apply(x)) <= 1e-12
    M.set_precision(deco_amd.PREC_FP32)
    y32 = M.apply(x).numpy()
    M.set_precision(deco_amd.PREC_FP64)
    assert rel(y32, y) <= 2e-4


@pytest.mark.parametrize("relax,coarse,weight", [(17, 10, 0), (18, 8, 0), (7, 10, 0), (7, 10, 0.6), (-1, 10, 0),
                                                (0, 8, 0), (11, 10, 0), (12, 10, 0), (14, 10, 0)])
def test_amg_variants(gpu, relax, coarse, weight):
    """Every smoother (F-C-F, l1-Jacobi, weighted Jacobi, the two-stage Gauss-Seidel forms of -1/0
    and 11-14) and both coarsenings: the device hierarchy against the oracle's own setup, and the
    V-cycle and its transpose against the oracle's."""
    g = gen_system(1)
    A = dev_csr(g)
    Ap = A.local_matrix(2, 0)
    tree = {"amg_SoC_alpha": "0.5", "amg_coarse_type": str(coarse), "amg_relax_type": str(relax),
            "amg_prol_trunc_fac": "1e-3", "amg_A_drop_tol": "1e-4", "amg_max_coarse": "50"}
    if weight:
        tree["amg_relax_weight"] = str(weight)
    G = deco_amd.AMG(tree, Ap)
    Go, levels, _ = oracle_amg(G, oracle.local_matrix(to_mat(g), 2, 0), tree)
    assert len(levels) >= 2
    x = np.random.default_rng(4).uniform(-1, 1, Ap.n)
    y = G.apply(x).numpy()
    # for the Jacobi-type smoothers the device's up-sweep is one product with P - D^-1 A P (and
    # P - D^-1 A^T P transposed), the oracle's the two steps of the reference form; the last
    # level's sub-cycle is one dense product on the device (the F-C-F smoother aside)
    assert rel(y, Go.apply(x)) <= 1e-12
    assert rel(G.apply_t(x).numpy(), Go.apply_t(x)) <= 1e-12
    G.set_precision(deco_amd.PREC_FP32)
    y32 = G.apply(x).numpy()
    G.set_precision(deco_amd.PREC_FP64)
    assert rel(y32, y) <= 1e-5 and not np.array_equal(y32, y)


def test_jacobi_block_parity(gpu):
    g = gen_system(2)
    A = dev_csr(g)
    J = deco_amd.Jacobi({"block_size": "3"}, A)
    Jo = oracle.jacobi(3, to_mat(g))
    x = np.random.default_rng(5).uniform(-1, 1, g["n"])
    y, yt = J.apply(x).numpy(), J.apply_t(x).numpy()
    assert rel(y, Jo.apply(x)) <= 1e-15
    assert rel(yt, Jo.apply_t(x)) <= 1e-15
    J.set_precision(deco_amd.PREC_FP32)                   # FP32 inverses, FP64 vectors
    y32, yt32 = J.apply(x).numpy(), J.apply_t(x).numpy()
    J.set_precision(deco_amd.PREC_FP64)
    assert rel(y32, y) <= 1e-6 and rel(yt32, yt) <= 1e-6
    assert not np.array_equal(y32, y)                     # the FP32 path did run


def test_spgemm_long_rows(gpu):
    import scipy.sparse as sp
    rng = np.random.default_rng(9)
    # rows whose candidate count exceeds the LDS tables (global accumulator path)
    S1 = sp.random(64, 3000, density=0.5, random_state=rng, format="csr")
    S2 = sp.random(3000, 9000, density=0.004, random_state=rng, format="csr")
    C = deco_amd.CSR.from_scipy(S1) @ deco_amd.CSR.from_scipy(S2)
    Co = oracle.spgemm(oracle.Mat.from_scipy(S1), oracle.Mat.from_scipy(S2))
    ok, v = same_pattern(C, Co)
    assert ok and np.array_equal(v, Co.v)


@pytest.mark.parametrize("cfg", [4, 2])
def test_cpr_product_layouts(gpu, cfg):
    """CPR (FSAI factors, pressure rows and decoupled A) with its products on the CSR arrays and
    on the cell-block shadows (2 x 2 and 3 x 3 cells), in both precisions and transposed, against
    the oracle's CPR on the same hierarchy; the layout given in the "prec" tree and switched later
    give the same operator."""
    g = gen_system(cfg)
    bs = g["bs"]
    A = dev_csr(g)
    M = deco_amd.CPR(small_params(bs, layout="csr"), A, None)
    Mo = oracle.CPR(deco_amd.cpr_params(small_params(bs)), to_mat(g), None)
    Go, _, _ = oracle_amg(M.amg(), Mo.pressure_matrix(), small_params(bs))
    Mo.set_local(Go)
    x = np.random.default_rng(11).uniform(-1, 1, g["n"])
    yo, yto = Mo.apply(x), Mo.apply_t(x)
    first = M.apply(x).numpy()
    for lay in (deco_amd.LAYOUT_CSR, deco_amd.LAYOUT_BLOCK):
        M.set_layout(lay)
        y, yt = M.apply(x).numpy(), M.apply_t(x).numpy()
        assert rel(y, yo) <= 1e-12 and rel(yt, yto) <= 1e-11
        if lay == deco_amd.LAYOUT_CSR:
            assert np.array_equal(y, first)
        M.set_precision(deco_amd.PREC_FP32)
        y32, yt32 = M.apply(x).numpy(), M.apply_t(x).numpy()
        M.set_precision(deco_amd.PREC_FP64)
        assert rel(y32, yo) <= 2e-4 and rel(yt32, yto) <= 2e-4


def test_fsai_product_layouts(gpu):
    """Standalone FSAI with FP64 products on its factors' 2 x 2 block shadows and on CSR."""
    g = gen_system(4)
    A = dev_csr(g)
    tree = small_params(2)
    F = deco_amd.FSAI(tree, A)
    Fo = oracle.fsai(deco_amd.cpr_params(tree), to_mat(g))
    x = np.random.default_rng(12).uniform(-1, 1, g["n"])
    for lay in (deco_amd.LAYOUT_BLOCK, deco_amd.LAYOUT_CSR):
        F.set_layout(lay)
        assert rel(F.apply(x).numpy(), Fo.apply(x)) <= 1e-13
        assert rel(F.apply_t(x).numpy(), Fo.apply_t(x)) <= 1e-13
    with pytest.raises(deco_amd._lib.DecoError):
        F.set_layout(7)


@pytest.mark.parametrize("bs", [2, 3])
def test_cpr_zero_diagonals_bitexact(gpu, bs):
    """Systems whose zero diagonals drive every branch of the tiny-diagonal fix (flagged values,
    the 3-unknown row swap, the undo after setup): decoupled A and b, pressure matrix and left
    transform bit-identical to the oracle's."""
    from conftest import zero_diag_system
    g, cases = zero_diag_system(bs)
    tree = small_params(bs)
    A = dev_csr(g)
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(tree, A, b)
    Mo_A = to_mat(g)
    bo = g["b"].copy()
    Mo = oracle.CPR(deco_amd.cpr_params(tree), Mo_A, bo)
    r, c, v = A.download()
    assert np.array_equal(r, Mo_A.r) and np.array_equal(c, Mo_A.c) and np.array_equal(v, Mo_A.v)
    assert np.array_equal(b.numpy(), bo)
    for i, what in cases.items():
        if what == "flag":
            k = r[i] + int(np.searchsorted(c[r[i]:r[i + 1]], i))
            assert v[k] == 0.0
    ok, vp = same_pattern(M.pressure_matrix(), Mo.pressure_matrix())
    assert ok and np.array_equal(vp, Mo.pressure_matrix().v)
    assert np.array_equal(M.left_transform(), Mo.left_transform(bs))


@pytest.mark.parametrize("cfg,opts", [
    (4, {"weights": "none", "use_drs": "0"}),
    (4, {"weights": "quasiimpes", "use_drs": "1", "drs.e_ps": "0.05"}),
    (4, {"CPD": "0"}),
    (4, {"tol_zero": "1e-3"}),
    (2, {"weights": "quasiimpes", "use_drs": "0"}),
    (2, {"weights": "none", "CPD": "0", "tol_zero": "1e-4"}),
])
def test_cpr_setup_options_bitexact(gpu, cfg, opts):
    """The decoupling options beyond the default set (weights none / quasiimpes, DRS with its
    e_ps test, no CPD, the tol_zero filter) in 2- and 3-unknown cells: decoupled A and b,
    pressure matrix and left transform bit-identical to the oracle's, and the CPR apply on the
    same hierarchy."""
    g = gen_system(cfg)
    bs = g["bs"]
    tree = small_params(bs, **opts)
    A = dev_csr(g)
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    M = deco_amd.CPR(tree, A, b)
    Mo_A = to_mat(g)
    bo = g["b"].copy()
    Mo = oracle.CPR(deco_amd.cpr_params(tree), Mo_A, bo)
    r, c, v = A.download()
    assert np.array_equal(r, Mo_A.r) and np.array_equal(c, Mo_A.c) and np.array_equal(v, Mo_A.v)
    assert np.array_equal(b.numpy(), bo)
    ok, vp = same_pattern(M.pressure_matrix(), Mo.pressure_matrix())
    assert ok and np.array_equal(vp, Mo.pressure_matrix().v)
    assert np.array_equal(M.left_transform(), Mo.left_transform(bs))
    Go, _, _ = oracle_amg(M.amg(), Mo.pressure_matrix(), tree)
    Mo.set_local(Go)
    x = np.random.default_rng(13).uniform(-1, 1, g["n"])
    assert rel(M.apply(x).numpy(), Mo.apply(x)) <= 1e-12


def test_amg_chunk_columns16_bit_identical(gpu):
    """The AMG levels' stream products on 16-bit chunk-relative columns give the V-cycle of the
    32-bit columns bit for bit, in both precisions."""
    g = gen_system(1)
    tree = {"amg_SoC_alpha": "0.5", "amg_coarse_type": "10", "amg_relax_type": "18",
            "amg_prol_trunc_fac": "1e-3", "amg_A_drop_tol": "1e-4", "amg_max_coarse": "50"}
    x = np.random.default_rng(14).uniform(-1, 1, g["n"] // 2)
    out = {}
    for off16 in (True, False):
        deco_amd.set_sell_offsets16(off16)
        try:
            Ap = dev_csr(g).local_matrix(2, 0)
            G = deco_amd.AMG(tree, Ap)
            y = G.apply(x).numpy()
            G.set_precision(deco_amd.PREC_FP32)
            out[off16] = (y, G.apply(x).numpy())
        finally:
            deco_amd.set_sell_offsets16(True)
    assert np.array_equal(out[True][0], out[False][0]) and np.array_equal(out[True][1], out[False][1])


def test_amg_row_piece_products_match_stream(gpu):
    """The AMG operators' row-piece products (rows cut into pieces of 1-8 entries, 64 per
    wavefront, segmented sums) against the CSR-stream kernel on the same hierarchy: the V-cycle
    agrees to 1e-13 in FP64 and to FP32 rounding in FP32, forward and transposed, and the
    oracle's V-cycle to 1e-12."""
    g = gen_system(1)
    tree = {"amg_SoC_alpha": "0.5", "amg_coarse_type": "10", "amg_relax_type": "-1",
            "amg_prol_trunc_fac": "1e-3", "amg_A_drop_tol": "1e-4", "amg_max_coarse": "50"}
    x = np.random.default_rng(16).uniform(-1, 1, g["n"] // 2)
    out = {}
    for pieces in (True, False):
        deco_amd.set_stream_pieces(pieces)
        try:
            Ap = dev_csr(g).local_matrix(2, 0)
            G = deco_amd.AMG(tree, Ap)
            y64, yt = G.apply(x).numpy(), G.apply_t(x).numpy()
            G.set_precision(deco_amd.PREC_FP32)
            out[pieces] = (y64, yt, G.apply(x).numpy())
            if pieces:
                G.set_precision(deco_amd.PREC_FP64)
                Go, _, _ = oracle_amg(G, oracle.local_matrix(to_mat(g), 2, 0), tree)
                assert rel(y64, Go.apply(x)) <= 1e-12 and rel(yt, Go.apply_t(x)) <= 1e-12
        finally:
            deco_amd.set_stream_pieces(True)
    assert rel(out[True][0], out[False][0]) <= 1e-13 and rel(out[True][1], out[False][1]) <= 1e-13
    assert rel(out[True][2], out[False][2]) <= 1e-5


def _sorted_vs_row_order(G_factory, x):
    """V-cycles (FP64 forward and transposed, FP32 forward) with the CSR-stream products on the
    column-sorted chunk form wherever it applies (mode 2), by the library's rule (1), and on the
    row-order arrays (0), the row-piece form off."""
    out = {}
    deco_amd.set_stream_pieces(False)
    try:
        for mode in (2, 1, 0):
            deco_amd.set_stream_sorted(mode)
            G = G_factory()
            y64, yt = G.apply(x).numpy(), G.apply_t(x).numpy()
            G.set_precision(deco_amd.PREC_FP32)
            out[mode] = (y64, yt, G.apply(x).numpy())
    finally:
        deco_amd.set_stream_sorted(1)
        deco_amd.set_stream_pieces(True)
    return out


@pytest.mark.parametrize("relax", [18, -1])
def test_amg_sorted_chunks_bit_identical(gpu, relax):
    """The stream products on the column-sorted chunk form (entries by column, products written
    to their row-order slots) give the V-cycle of the row-order arrays bit for bit, in both
    precisions and transposed, with the Jacobi-type up-sweep operators (18) and the two-stage
    smoother's fused restriction operators (-1)."""
    g = gen_system(1)
    tree = {"amg_SoC_alpha": "0.5", "amg_coarse_type": "10", "amg_relax_type": str(relax),
            "amg_prol_trunc_fac": "1e-3", "amg_A_drop_tol": "1e-4", "amg_max_coarse": "50"}
    x = np.random.default_rng(17).uniform(-1, 1, g["n"] // 2)
    out = _sorted_vs_row_order(lambda: deco_amd.AMG(tree, dev_csr(g).local_matrix(2, 0)), x)
    for mode in (2, 1):
        for a, b in zip(out[mode], out[0]):
            assert np.array_equal(a, b)


def test_amg_wide_chunk_span_falls_back(gpu):
    """Couplings 80,000 rows apart make the restriction's row chunks span more than 65535
    columns: those products keep 32-bit columns (20-bit offsets in the column-sorted form), and
    the V-cycle still matches the oracle's."""
    import scipy.sparse as sp
    m = 300
    T = sp.diags([-1.0, 2.0, -1.0], [-1, 0, 1], shape=(m, m))
    L = (sp.kron(sp.eye(m), T) + sp.kron(T, sp.eye(m))).tocsr()
    n = L.shape[0]
    i = np.arange(0, 5000, 97)
    far = sp.csr_matrix((np.full(2 * len(i), -1.5), (np.concatenate([i, i + 80000]), np.concatenate([i + 80000, i]))),
                        shape=(n, n))
    S = (L + far + sp.diags(np.asarray(abs(far).sum(1)).ravel())).tocsr()
    S.sort_indices()
    tree = {"amg_SoC_alpha": "0.25", "amg_coarse_type": "10", "amg_relax_type": "18",
            "amg_prol_trunc_fac": "1e-3", "amg_A_drop_tol": "0", "amg_max_coarse": "300"}
    G = deco_amd.AMG(tree, deco_amd.CSR.from_scipy(S))
    Go, levels, _ = oracle_amg(G, oracle.Mat.from_scipy(S), tree)
    assert len(levels) >= 2
    x = np.random.default_rng(15).uniform(-1, 1, n)
    assert rel(G.apply(x).numpy(), Go.apply(x)) <= 1e-12
    # the column-sorted form takes these chunks too (20-bit offsets), with the same sums
    out = _sorted_vs_row_order(lambda: deco_amd.AMG(tree, deco_amd.CSR.from_scipy(S)), x)
    for mode in (2, 1):
        for a, b in zip(out[mode], out[0]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("relax,cover", [(18, 1), (7, 2), (-1, 1), (-1, 2), (12, 2), (14, 1)])
def test_amg_dense_tail_matches_sparse_subcycle(gpu, relax, cover):
    """The sub-cycle of the last `cover` levels with interpolation, formed as one dense operator at
    setup (on the device, level by level from the coarsest inverse), against the same hierarchy
    with the level-by-level cycle (dense tail disabled) and the oracle's level-by-level V-cycle,
    forward and transposed, in both precisions."""
    g = gen_system(1)
    A = dev_csr(g)
    Ap = A.local_matrix(2, 0)
    tree = {"amg_SoC_alpha": "0.5", "amg_coarse_type": "10", "amg_relax_type": str(relax),
            "amg_prol_trunc_fac": "1e-3", "amg_A_drop_tol": "1e-4", "amg_max_coarse": "50"}
    deco_amd.set_amg_dense_tail(0)
    try:
        G0 = deco_amd.AMG(tree, Ap)
    finally:
        deco_amd.set_amg_dense_tail(4608)
    Go, levels, _ = oracle_amg(G0, oracle.local_matrix(to_mat(g), 2, 0), tree)
    assert len(levels) >= 3
    rows = levels[-cover]["n"]
    deco_amd.set_amg_dense_tail(rows)
    try:
        G1 = deco_amd.AMG(tree, Ap)
    finally:
        deco_amd.set_amg_dense_tail(4608)
    assert G0.dense_tail() == 0 and G1.dense_tail() == rows > 0
    x = np.random.default_rng(7).uniform(-1, 1, Ap.n)
    y0, y1 = G0.apply(x).numpy(), G1.apply(x).numpy()
    assert rel(y1, y0) <= 1e-13 and not np.array_equal(y1, y0)   # the dense product did run
    assert rel(y1, Go.apply(x)) <= 1e-12
    t0, t1 = G0.apply_t(x).numpy(), G1.apply_t(x).numpy()
    assert rel(t1, t0) <= 1e-13
    assert rel(t1, Go.apply_t(x)) <= 1e-12
    for G in (G0, G1):
        G.set_precision(deco_amd.PREC_FP32)
    assert rel(G1.apply(x).numpy(), G0.apply(x).numpy()) <= 1e-5
    assert rel(G1.apply_t(x).numpy(), G0.apply_t(x).numpy()) <= 1e-5


@pytest.mark.parametrize("seq", [0, 3])
def test_cpr_dense_tail_is_the_whole_pressure_cycle(gpu, seq):
    """One level with interpolation (2048 pressure rows, two AMG levels): the dense tail is the
    whole pressure V-cycle, and in the fused CPR stage it also adds its result into the iterate."""
    g = gen_system(4)
    tree = small_params(2, seq=str(seq), **{"prec_loc.amg_max_levels": "2"})
    A = dev_csr(g)
    M = deco_amd.CPR(tree, A, None)
    Mo = oracle.CPR(deco_amd.cpr_params(tree), to_mat(g), None)
    Go, levels, _ = oracle_amg(M.amg(), Mo.pressure_matrix(), tree)
    assert len(levels) == 1 and M.amg().dense_tail() == levels[0]["n"]
    Mo.set_local(Go)
    x = np.random.default_rng(seq).uniform(-1, 1, g["n"])
    assert rel(M.apply(x).numpy(), Mo.apply(x)) <= 1e-12
    if seq == 3:                                          # the transposed CPR's sequence
        assert rel(M.apply_t(x).numpy(), Mo.apply_t(x)) <= 1e-11
    M.set_precision(deco_amd.PREC_FP32)
    assert rel(M.apply(x).numpy(), Mo.apply(x)) <= 2e-4
    M.set_precision(deco_amd.PREC_FP64)


@pytest.mark.parametrize("cfg", [4, 2])
def test_transposed_cpr_sequence_and_value_update(gpu, cfg):
    """The transposed CPR application (seq 3) against its definition assembled on the host from
    the device's FSAI and AMG transposes and scipy products with the decoupled A^T:
    M^T u = a + P e_p + G^T (s - A^T P e_p), a = G^T u, s = u - A^T a, e_p = AMG^T(s[bs i]).
    Then A's values change under the preconditioner (here to 1.5 A): the transposed application's
    products with A^T, its pressure columns included, follow the new values."""
    g = gen_system(cfg)
    bs = g["bs"]
    A = dev_csr(g)
    M = deco_amd.CPR(small_params(bs), A, None)
    F, G = M.fsai(), M.amg()
    u = np.random.default_rng(21).uniform(-1, 1, g["n"])

    def expected(At):
        a = F.apply_t(u).numpy()
        s = u - At @ a
        e = np.zeros_like(u)
        e[::bs] = G.apply_t(s[::bs]).numpy()
        return a + e + F.apply_t(s - At @ e).numpy()

    _, _, v = A.download()
    At = A.to_scipy().T.tocsr()
    t0 = M.apply_t(u).numpy()
    assert rel(t0, expected(At)) <= 1e-12
    A.update_values(v)                                    # the same values: the same operator
    assert rel(M.apply_t(u).numpy(), t0) <= 1e-14
    A.update_values(1.5 * v)
    t1 = M.apply_t(u).numpy()
    assert rel(t1, expected(1.5 * At)) <= 1e-12 and rel(t1, t0) > 1e-3
    M.set_precision(deco_amd.PREC_FP32)
    assert rel(M.apply_t(u).numpy(), t1) <= 2e-4
    M.set_precision(deco_amd.PREC_FP64)


@pytest.mark.parametrize("bs", [2, 3])
def test_sell_pairs_bit_identical(gpu, bs):
    """FP32 products on the uniform-offset sliced block shadows with two slices per wavefront
    (modes 1 and 2) against one slice per wavefront (mode 0): the CPR application, its transpose
    and an FP32-preconditioned solve agree bit for bit. 23 x 19 x 5 cells make 35 slices of 64
    cell rows, the last one partial, so the last wavefront's second slice lies past the end."""
    g = gen_system(0, custom=(23, 19, 5, bs))
    A = dev_csr(g)
    M = deco_amd.CPR(small_params(bs), A, None)
    M.set_layout(deco_amd.LAYOUT_BLOCK)
    M.set_precision(deco_amd.PREC_FP32)
    u = np.random.default_rng(5).uniform(-1, 1, g["n"])
    sol = {"type": "rgmres", "reltol": "1e-8", "restart": "30", "itmax": "200", "verbosity": "0"}
    out = {}
    try:
        for mode in (0, 1, 2):
            deco_amd.set_sell_pairs(mode)
            S = deco_amd.SOL(sol, A, M, g["b"])
            out[mode] = (M.apply(u).numpy(), M.apply_t(u).numpy(), S.x.numpy())
            assert S.flag == 0
    finally:
        deco_amd.set_sell_pairs(1)
    for mode in (1, 2):
        for a, b in zip(out[mode], out[0]):
            assert np.array_equal(a, b)
