"""GPU parity of the device CSR kernels against the CPU oracle (through the C ABI)."""
import numpy as np
import pytest
import scipy.sparse as sp

import deco_amd
import oracle
from conftest import gen_system, to_mat

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("cfg", [1, 4, 2])
def test_spmv_fp64_parity(gpu, cfg):
    g = gen_system(cfg)
    A = deco_amd.CSR(g["n"], g["n"], g["r"], g["c"], g["v"])
    M = to_mat(g)
    x = np.random.default_rng(cfg).uniform(-1, 1, g["n"])
    y = A.apply(x).numpy()
    yo = oracle.spmv(M, x)
    assert rel(y, yo) <= 1e-14
    # elementwise bound gamma_k (|A||x|)_i, k = row length
    absAx = oracle.spmv(oracle.Mat(M.n, M.m, M.r, M.c, np.abs(M.v)), np.abs(x))
    k = np.diff(M.r).max()
    assert np.all(np.abs(y - yo) <= 2 * k * 1.2e-16 * absAx + 1e-300)
    # b = A x* from the generator
    assert rel(A.apply(g["xstar"]).numpy(), g["b"]) <= 1e-14


def test_spmv_strided_residual_transpose_f32(gpu):
    g = gen_system(1)
    bs = g["bs"]
    A = deco_amd.CSR(g["n"], g["n"], g["r"], g["c"], g["v"])
    M = to_mat(g)
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, g["n"])
    z = rng.uniform(-1, 1, g["n"])
    assert rel(A.apply(x, bs=bs).numpy(), oracle.spmv(M, x, bs=bs)) <= 1e-14
    assert rel(A.residual(x, z).numpy(), z - oracle.spmv(M, x)) <= 1e-13
    S = M.to_scipy()
    assert rel(A.apply_transpose(x).numpy(), S.T @ x) <= 1e-14
    y32 = A.apply_f32(x.astype(np.float32)).numpy()
    assert rel(y32.astype(np.float64), S @ x) <= 1e-5


def test_blas_dot_nrm2_axpby(gpu):
    n = 1_000_003
    a = np.arange(1, n + 1, dtype=np.float64)
    assert abs(deco_amd.dot(a, a) - n * (n + 1) * (2 * n + 1) / 6) <= 1e-15 * n ** 3
    rng = np.random.default_rng(3)
    x, y = rng.standard_normal(n), rng.standard_normal(n)
    assert abs(deco_amd.nrm2(x) - np.linalg.norm(x)) <= 1e-13 * np.linalg.norm(x)
    yd = deco_amd.DeviceArray.from_numpy(y)
    deco_amd.axpby(0.5, x, -2.0, yd)
    assert np.array_equal(yd.numpy(), -2.0 * y + 0.5 * x)
    # deterministic: bitwise-identical reruns
    assert deco_amd.dot(x, y) == deco_amd.dot(x, y)


def test_read_bandwidth_probe(gpu):
    """The read-bandwidth measurement the bench reports beside the roofline: a positive rate
    below the published HBM peak, and argument errors reported as errors."""
    import ctypes
    L = deco_amd.lib()
    g = ctypes.c_double()
    deco_amd.check(L.deco_read_bandwidth(1 << 30, 3, ctypes.byref(g)))   # 1 GiB: past the 256 MB cache
    assert 0.0 < g.value < 8000.0
    assert L.deco_read_bandwidth(1024, 3, ctypes.byref(g)) != 0
    assert L.deco_read_bandwidth(64 << 20, 0, ctypes.byref(g)) != 0


def test_transpose_filters_local_matrix_bitexact(gpu):
    g = gen_system(2)
    bs = g["bs"]
    A = deco_amd.CSR(g["n"], g["n"], g["r"], g["c"], g["v"])
    M = to_mat(g)
    T = A.transpose()
    To = oracle.transpose(M)
    r, c, v = T.download()
    assert np.array_equal(r, To.r) and np.array_equal(c, To.c) and np.array_equal(v, To.v)
    for tol, proper in [(0.05, False), (0.2, True)]:
        B = A.copy()
        Mo = M.copy()
        if proper:
            B.filter_proper(tol)
            oracle.filter_proper(Mo, tol)
        else:
            B.filter(tol)
            oracle.filter(Mo, tol)
        r, c, v = B.download()
        assert np.array_equal(r, Mo.r) and np.array_equal(c, Mo.c) and np.array_equal(v, Mo.v)
    L = A.local_matrix(bs, 0)
    Lo = oracle.local_matrix(M, bs, 0)
    r, c, v = L.download()
    assert np.array_equal(r, Lo.r) and np.array_equal(c, Lo.c) and np.array_equal(v, Lo.v)


def test_empty_and_single_row_edges(gpu):
    A = deco_amd.CSR(1, 1, [0, 1], [0], [2.0])
    assert A.apply(np.array([3.0])).numpy()[0] == 6.0
    E = deco_amd.CSR(3, 3, [0, 0, 0, 0], np.zeros(0, np.int32), np.zeros(0))
    assert np.array_equal(E.apply(np.ones(3)).numpy(), np.zeros(3))
    assert E.transpose().nz == 0


def test_null_device_arrays_are_rejected(gpu):
    """A NULL device array for a non-empty operand returns DECO_ERR_ARG (no kernel runs on it);
    empty operands may be NULL."""
    import ctypes
    A = deco_amd.CSR(3, 3, [0, 1, 2, 3], [0, 1, 2], [1.0, 2.0, 3.0])
    x, y = deco_amd.DeviceArray.from_numpy(np.ones(3)), deco_amd.DeviceArray(3)
    lib = deco_amd.lib()
    null = None
    assert lib.deco_spmv(A.h, null, y.ptr, 1) == -1
    assert lib.deco_spmv(A.h, x.ptr, null, 1) == -1
    assert lib.deco_spmv_residual(A.h, x.ptr, null, y.ptr) == -1
    assert lib.deco_spmv_block(A.h, 1, 1, null, y.ptr) == -1
    assert lib.deco_dot(3, x.ptr, null, ctypes.byref(ctypes.c_double())) == -1
    assert lib.deco_axpby(3, 1.0, null, 1.0, y.ptr) == -1
    assert lib.deco_memcpy_d2d(y.ptr, null, 24) == -1
    assert lib.deco_grad_A_pattern(A.h, x.ptr, x.ptr, null) == -1
    assert lib.deco_spmv_time(A.h, 5, 1, ctypes.byref(ctypes.c_double())) == -1
    assert lib.deco_dot(0, null, null, ctypes.byref(ctypes.c_double())) == 0
    assert lib.deco_spmv(A.h, x.ptr, y.ptr, 1) == 0 and np.array_equal(y.numpy(), [1.0, 2.0, 3.0])


@pytest.mark.parametrize("cfg,bs,ro,co", [(2, 3, "213", "213"), (2, 3, "312", "132"), (1, 2, "21", "21"),
                                          (1, 2, "12", "12")])
def test_block_reorder_bitexact(gpu, cfg, bs, ro, co):
    """deco_block_reorder against the oracle's AD::block_reorder: r, c, v and b bit-identical."""
    g = gen_system(cfg)
    A = deco_amd.CSR(g["n"], g["n"], g["r"], g["c"], g["v"])
    b = deco_amd.DeviceArray.from_numpy(g["b"])
    A.block_reorder(bs, ro, co, b)
    Ao, bo = oracle.block_reorder(to_mat(g), g["b"], bs, ro, co)
    r, c, v = A.download()
    assert np.array_equal(r, Ao.r) and np.array_equal(c, Ao.c[:Ao.nnz]) and np.array_equal(v, Ao.v[:Ao.nnz])
    assert np.array_equal(b.numpy(), bo)
    with pytest.raises(deco_amd._lib.DecoError):
        A.block_reorder(bs, "1" * bs, co)


@pytest.mark.parametrize("pad,off16,dia", [(10.0, True, False), (10.0, False, False), (-1.0, True, False),
                                           (0.10, True, True)])
@pytest.mark.parametrize("cfg,shape", [(1, (2, 2)), (2, (3, 3)), (1, "prow"), (1, "tril"), (2, "tril")])
def test_block_f32_product(gpu, cfg, shape, pad, off16, dia):
    """Block-shadow products against the CSR product: full cell blocks, the pressure rows
    (1 x bs blocks), and lower-triangular patterns whose diagonal blocks are partly empty (the
    FSAI factor shape); in the sliced layout (pad 10: chosen whatever its padding) with 16-bit
    block-column offsets and with 32-bit columns, in the chunked one (pad -1), and with the
    uniform-offset form allowed (taken by the full 32x32x16 grid's stencil shapes, not by the
    black-oil system, whose fault links give more than 16 offsets)."""
    deco_amd.set_sell_padding(pad)
    deco_amd.set_sell_offsets16(off16)
    deco_amd.set_sell_dia(dia)
    try:
        _block_products(cfg, shape, None if dia else (0 if pad < 0 else (2 if off16 else 1)))
    finally:
        deco_amd.set_sell_padding(0.10)
        deco_amd.set_sell_offsets16(True)
        deco_amd.set_sell_dia(True)


def _block_shape(cfg, shape):
    g = gen_system(cfg)
    bs = g["bs"]
    S = to_mat(g).to_scipy()
    if shape == "prow":
        S = S[::bs].tocsr()
        R, C = 1, bs
    elif shape == "tril":
        S = sp.tril(S).tocsr()
        R, C = bs, bs
    else:
        R, C = shape
    S.sort_indices()
    return S, R, C


def _uniform_offsets_apply(S, R, C, pad=0.10):
    """Whether the uniform-offset layout applies: at most 16 distinct block-column offsets and
    64-row slices of that width storing at most (1 + pad) times the blocks."""
    rows = np.repeat(np.arange(S.shape[0]), np.diff(S.indptr)) // R
    cols = S.indices // C
    offs = np.unique(cols - rows)
    nnzb = len(np.unique(rows.astype(np.int64) * (S.shape[1] // C + 1) + cols))
    nb = S.shape[0] // R
    return len(offs) <= 16 and ((nb + 63) // 64) * 64 * len(offs) <= nnzb * (1 + pad)


def _block_products(cfg, shape, layout):
    S, R, C = _block_shape(cfg, shape)
    A = deco_amd.CSR.from_scipy(S)
    if layout is None:
        if _uniform_offsets_apply(S, R, C):
            assert A.block_info(R, C)[0] == 3
        else:
            assert A.block_info(R, C)[0] != 3
    else:
        if layout == 2:   # 16-bit offsets only where every block column lies within reach
            rows = np.repeat(np.arange(S.shape[0]), np.diff(S.indptr)) // R
            fits = np.abs(S.indices // C - rows).max() <= 32767 and S.shape[0] // R <= S.shape[1] // C
            layout = 2 if fits else 1
        assert A.block_info(R, C)[0] == layout
    x = np.random.default_rng(3).uniform(-1, 1, S.shape[1]).astype(np.float32)
    y = A.apply_f32(x, block=(R, C)).numpy().astype(np.float64)
    ref = S @ x.astype(np.float64)
    assert rel(y, ref) <= 2e-6
    # the CSR FP32 product agrees as well
    assert rel(A.apply_f32(x).numpy().astype(np.float64), ref) <= 2e-6
    with pytest.raises(deco_amd._lib.DecoError):
        A.apply_f32(x, block=(R, 5))
    # FP64 values on the same block structure: the oracle's CSR product to FP64 rounding (the
    # sums pair the two products of a block before adding it to the row)
    xd = np.random.default_rng(4).uniform(-1, 1, S.shape[1])
    yo = oracle.spmv(oracle.Mat.from_scipy(S), xd)
    assert rel(A.apply(xd, block=(R, C)).numpy(), yo) <= 1e-14
    with pytest.raises(deco_amd._lib.DecoError):
        A.apply(xd, block=(R, 5))


@pytest.mark.parametrize("cfg,shape", [(1, (2, 2)), (1, "prow"), (1, "tril"), (11, (2, 2))])
def test_uniform_offset_layout_equals_sliced(gpu, cfg, shape):
    """The uniform-offset layout sums each block row's blocks in the same column order as the
    sliced one, its zero blocks adding nothing: FP64 and FP32 products equal the sliced layout's
    bit for bit; the FP64 one agrees with the oracle to 1e-14."""
    S, R, C = _block_shape(cfg, shape)
    assert _uniform_offsets_apply(S, R, C)
    rng = np.random.default_rng(23)
    x = rng.uniform(-1, 1, S.shape[1])
    ys = {}
    for dia in (True, False):
        deco_amd.set_sell_dia(dia)
        try:
            A = deco_amd.CSR.from_scipy(S)
            assert A.block_info(R, C)[0] == (3 if dia else 2)
            ys[dia] = (A.apply(x, block=(R, C)).numpy(), A.apply_f32(x.astype(np.float32), block=(R, C)).numpy())
        finally:
            deco_amd.set_sell_dia(True)
    assert np.array_equal(ys[True][0], ys[False][0]) and np.array_equal(ys[True][1], ys[False][1])
    assert rel(ys[True][0], oracle.spmv(oracle.Mat.from_scipy(S), x)) <= 1e-14


def test_spmv_wide_column_offsets(gpu):
    """Rows coupling to columns far from the diagonal (60000 rows away) and near ones only: the
    FP64 product equals the oracle bit for bit, the FP32 one to FP32 accuracy."""
    n = 70000
    rng = np.random.default_rng(11)
    for far in (False, True):
        rows = np.repeat(np.arange(n), 3)
        i = np.arange(n)
        cols = np.stack([i, np.minimum(i + 1, n - 1), np.minimum(i + 7, n - 1)], 1).ravel()
        if far:
            cols[3 * 5 + 2] = 60005            # row 5 couples to a column 60000 rows away
        S = sp.csr_matrix((rng.uniform(-1, 1, 3 * n), (rows, cols)), shape=(n, n))
        S.sum_duplicates()
        S.sort_indices()
        A = deco_amd.CSR.from_scipy(S)
        M = oracle.Mat.from_scipy(S)
        x = rng.uniform(-1, 1, n)
        assert np.array_equal(A.apply(x).numpy(), oracle.spmv(M, x))
        y32 = A.apply_f32(x.astype(np.float32)).numpy().astype(np.float64)
        assert rel(y32, S @ x) <= 2e-6


def test_sliced_offsets_fall_back_to_32bit_columns(gpu):
    """The 16-bit block-column offsets are used only where they are exact: a block column more
    than 32767 block rows away, or more block rows than block columns (an offset 0 would leave
    x), keep 32-bit columns (layout 1); both products stay exact."""
    rng = np.random.default_rng(17)
    n = 40000
    i = np.arange(n)
    rows = np.concatenate([i, i[:-1], [5]])
    cols = np.concatenate([i, i[:-1] + 1, [n - 1]])   # row 5 reaches 39994 block rows away
    S = sp.csr_matrix((rng.uniform(-1, 1, len(rows)), (rows, cols)), shape=(n, n))
    S.sort_indices()
    A = deco_amd.CSR.from_scipy(S)
    assert A.block_info(1, 1)[0] == 1
    x = rng.uniform(-1, 1, n)
    assert rel(A.apply(x, block=(1, 1)).numpy(), oracle.spmv(oracle.Mat.from_scipy(S), x)) <= 1e-14
    near = S.copy()
    near.data[near.indices == n - 1] = 0.0
    near.eliminate_zeros()
    near = (near + sp.csr_matrix(([1.0], ([n - 1], [n - 1])), shape=(n, n))).tocsr()
    near.sort_indices()
    # two offsets (0, 1): the uniform-offset form, or with it disabled 16-bit offsets
    assert deco_amd.CSR.from_scipy(near).block_info(1, 1)[0] == 3
    deco_amd.set_sell_dia(False)
    try:
        assert deco_amd.CSR.from_scipy(near).block_info(1, 1)[0] == 2
    finally:
        deco_amd.set_sell_dia(True)
    # 600 block rows over 300 block columns (sliced whatever its padding)
    W = (sp.random(600, 300, density=0.02, random_state=rng, format="csr") + sp.eye(600, 300, format="csr")).tocsr()
    W.sort_indices()
    deco_amd.set_sell_padding(10.0)
    try:
        B = deco_amd.CSR.from_scipy(W)
        assert B.block_info(1, 1)[0] == 1
        xw = rng.uniform(-1, 1, 300)
        assert rel(B.apply(xw, block=(1, 1)).numpy(), oracle.spmv(oracle.Mat.from_scipy(W), xw)) <= 1e-14
    finally:
        deco_amd.set_sell_padding(0.10)


def test_invalid_csr_structures_are_rejected(gpu):
    """A column outside [0, m), a decreasing row pointer and one that does not span [0, nnz] are
    argument errors of deco_csr_create and deco_csr_create_dev (no kernel indexes with them);
    the CPR setup rejects a row without its diagonal."""
    import ctypes
    for n, m, r, c, v in [(3, 3, [0, 1, 2, 3], [0, 3, 2], [1.0] * 3),
                          (3, 3, [0, 2, 1, 3], [0, 1, 2], [1.0] * 3)]:
        with pytest.raises(deco_amd._lib.DecoError):
            deco_amd.CSR(n, m, r, c, v)
        rd = deco_amd.DeviceArray.from_numpy(np.array(r, np.int32))
        cd = deco_amd.DeviceArray.from_numpy(np.array(c, np.int32))
        vd = deco_amd.DeviceArray.from_numpy(np.array(v, np.float64))
        h = ctypes.c_void_p()
        assert deco_amd.lib().deco_csr_create_dev(n, m, len(c), rd.ptr, cd.ptr, vd.ptr, ctypes.byref(h)) == -1
    rd = deco_amd.DeviceArray.from_numpy(np.array([0, 1, 2, 2], np.int32))   # r[n] != nnz
    cd = deco_amd.DeviceArray.from_numpy(np.array([0, 1, 2], np.int32))
    vd = deco_amd.DeviceArray.from_numpy(np.ones(3))
    h = ctypes.c_void_p()
    assert deco_amd.lib().deco_csr_create_dev(3, 3, 3, rd.ptr, cd.ptr, vd.ptr, ctypes.byref(h)) == -1
    g = gen_system(4)
    S = to_mat(g).to_scipy().tolil()
    S[5, 5] = 0.0
    S = S.tocsr()
    S.eliminate_zeros()
    S.sort_indices()
    from conftest import small_params
    with pytest.raises(deco_amd._lib.DecoError):
        deco_amd.CPR(small_params(2), deco_amd.CSR.from_scipy(S), None)


def test_stream_product_row_lengths(gpu):
    """The CSR-stream product over rows of every length class: short rows (summed in column
    order), rows of 130-1600 entries (the chunk cap drops so a chunk fits the LDS window), rows
    longer than half the window (block reduction per row), and empty rows; a matrix without
    entries gives zeros."""
    rng = np.random.default_rng(11)
    m = 60_000
    lens = np.concatenate([rng.integers(0, 21, 2500), [0, 0, 130, 257, 394, 800, 1600, 1601, 2500, 4000],
                           rng.integers(100, 400, 40), rng.integers(0, 8, 700)])
    rng.shuffle(lens)
    n = len(lens)
    r = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    c = np.concatenate([np.sort(rng.choice(m, k, replace=False)) for k in lens]).astype(np.int32)
    v = rng.uniform(-1, 1, len(c))
    S = sp.csr_matrix((v, c, r), shape=(n, m))
    A = deco_amd.CSR(n, m, r, c, v)
    x = rng.uniform(-1, 1, m)
    z = rng.uniform(-1, 1, n)
    y = A.apply(x).numpy()
    ref = S @ x
    absAx = abs(S) @ np.abs(x)
    assert np.all(np.abs(y - ref) <= 2 * max(lens.max(), 1) * 1.2e-16 * absAx)
    assert np.array_equal(y[lens == 0], np.zeros(np.count_nonzero(lens == 0)))
    assert rel(A.residual(x, z).numpy(), z - ref) <= 1e-13
    y32 = A.apply_f32(x.astype(np.float32)).numpy().astype(np.float64)
    assert np.all(np.abs(y32 - ref) <= 1e-5 * (absAx + 1e-30) + 1e-6)
    E = deco_amd.CSR(5000, 100, np.zeros(5001, np.int32), np.zeros(0, np.int32), np.zeros(0))
    assert np.array_equal(E.apply(np.ones(100)).numpy(), np.zeros(5000))


@pytest.mark.parametrize("shape", ["ragged", "long_rows", "too_long", "wide"])
def test_column_sorted_chunks_bit_identical(gpu, shape):
    """CSR products on the column-sorted chunk form, forced wherever it applies, against the
    row-order arrays: the same values bit for bit in FP64 and FP32 (product and residual) for
    ragged rows with empty ones, rows up to half the chunk tile (every chunk fits its window), a
    row longer than that (the form does not apply) and chunks spanning more than 2^20 columns
    (the form does not apply); the FP64 product against scipy to the CSR-stream bound."""
    rng = np.random.default_rng({"ragged": 21, "long_rows": 22, "too_long": 23, "wide": 24}[shape])
    n = 30_000
    m = 3_000_000 if shape == "wide" else n
    if shape == "ragged":
        lens = rng.integers(0, 40, n)
        lens[rng.choice(n, 500, replace=False)] = 0
    else:
        lens = rng.integers(1, 12, n)
        if shape == "long_rows":
            lens[rng.choice(n, 20, replace=False)] = rng.integers(200, 1600, 20)
        elif shape == "too_long":
            lens[n // 2] = 2500
    cols = []
    for i, k in enumerate(lens):
        lo, hi = (0, m) if shape == "wide" else (max(0, i - 6000), min(m, i + 6000))
        cols.append(np.unique(lo + rng.integers(0, hi - lo, k)))   # sorted, duplicates dropped
    lens = np.array([len(c_) for c_ in cols])
    r = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    c = np.concatenate(cols).astype(np.int32)
    v = rng.uniform(-1, 1, len(c))
    x = rng.uniform(-1, 1, m)
    z = rng.uniform(-1, 1, n)
    out = {}
    try:
        for mode in (2, 0):
            deco_amd.set_stream_sorted(mode)
            A = deco_amd.CSR(n, m, r, c, v)
            out[mode] = (A.apply(x).numpy(), A.residual(x, z).numpy(),
                         A.apply_f32(x.astype(np.float32)).numpy())
    finally:
        deco_amd.set_stream_sorted(1)
    for a, b in zip(out[2], out[0]):
        assert np.array_equal(a, b)
    S = sp.csr_matrix((v, c, r), shape=(n, m))
    absAx = abs(S) @ np.abs(x)
    assert np.all(np.abs(out[2][0] - S @ x) <= 2 * max(lens.max(), 1) * 1.2e-16 * absAx)
