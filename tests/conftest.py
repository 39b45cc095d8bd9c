import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP path)")


_gen_cache = {}


def gen_system(cfg, seed=0, custom=None):
    """Generated system (host arrays), cached per session."""
    import deco_amd
    key = (cfg, seed, custom)
    if key not in _gen_cache:
        _gen_cache[key] = deco_amd.generate(cfg, seed, custom=custom)
    return _gen_cache[key]


def to_mat(g):
    import oracle
    return oracle.Mat(g["n"], g["n"], g["r"], g["c"], g["v"])


@pytest.fixture(scope="session")
def gpu():
    import deco_amd
    deco_amd.init(0)
    return True


def small_params(bs, **kw):
    """Reference-style "prec" tree (param_norne.json-like) for tests."""
    import deco_amd
    prec = {
        "block_size": str(bs), "pressure_id": "0", "seq": "3", "weights": "quasiimpes_GE",
        "use_drs": "1", "drs": {"e_dd": "0.2", "e_ps": "0.0"}, "CPD": "1", "tol_zero": "0.0",
        "prec_glo": {"type": "FSAI", "tol_rel": "0.2", "tol_abs": "1e-16", "power": "1", "block_size": "1"},
        "prec_loc": {"tol_rel": "1e-4", "amg_coarse_type": "10", "amg_relax_type": "18", "amg_agg_levs": "0",
                     "amg_agg_path": "1", "amg_A_drop_tol": "1e-4", "amg_prol_trunc_fac": "1e-3",
                     "amg_SoC_alpha": "0.5"},
    }
    for k, v in kw.items():
        node = prec
        keys = k.split(".")
        for kk in keys[:-1]:
            node = node[kk]
        node[keys[-1]] = str(v)
    return deco_amd.PTree(prec)


def diag_pos(g, i):
    """Index of the diagonal entry of row i in the generated system's CSR arrays."""
    import numpy as np
    b, e = g["r"][i], g["r"][i + 1]
    k = b + int(np.searchsorted(g["c"][b:e], i))
    assert g["c"][k] == i
    return k


def zero_diag_system(bs):
    """A small generated system with chosen zero entries that drive every branch of the CPR's
    tiny-diagonal fix (precond/CPR.hpp:753-840): returns (system with the modified values,
    {row: expected outcome}). bs 2: a zero pressure and a zero saturation diagonal. bs 3: a zero
    pressure diagonal; a zero diagonal in row i+1 whose right neighbour is nonzero (rows i+1 and
    i+2 swap); a zero diagonal in row i+1 whose neighbours (i+1, i+2) and (i+2, i+1) are zero too
    (the diagonal becomes -1e-20 and is flagged)."""
    g = dict(gen_system(0, custom=(4, 4, 3, bs)))
    v = g["v"].copy()
    if bs == 2:
        cases = {3 * 2: "flag", 5 * 2 + 1: "flag"}
        for i in cases:
            v[diag_pos(g, i)] = 0.0
    else:
        cases = {2 * 3: "flag", 4 * 3 + 1: "swap", 7 * 3 + 1: "flag"}
        v[diag_pos(g, 2 * 3)] = 0.0
        k = diag_pos(g, 4 * 3 + 1)
        assert v[k + 1] != 0.0
        v[k] = 0.0
        k = diag_pos(g, 7 * 3 + 1)
        k2 = diag_pos(g, 7 * 3 + 2)
        v[k] = v[k + 1] = v[k2 - 1] = 0.0
    g["v"] = v
    return g, cases


def amg_params(p):
    """deco_cpr_params of a CPR "prec" tree (PTree) or of a "prec_loc" dict (standalone AMG)."""
    import deco_amd
    if isinstance(p, dict):
        return deco_amd.cpr_params(deco_amd.PTree({"block_size": "1", "prec_loc": p}))
    if isinstance(p, deco_amd.PTree):
        return deco_amd.cpr_params(p)
    return p


def oracle_amg(G, A0, params, check=True):
    """The oracle's AMG stage on A0 (oracle.Mat, the stage's input matrix) built from only the
    device hierarchy G's interpolation operators P_l and C/F splittings: R_l, the Galerkin
    products, the drop rule, the smoother diagonals and the coarsest inverse are the oracle's own.
    With check, the device's setup products are compared with them: every level matrix (pattern
    and values), R_l and the smoother diagonals bit for bit, the coarsest inverse to 1e-12, and
    the drop rule against a scipy Galerkin product of the device's own level matrices (dropped
    entries below amg_A_drop_tol times the row maximum, kept ones at or above it)."""
    import oracle
    import scipy.sparse as sp
    p = amg_params(params)
    levels, cinv = G.export()
    Ps = [oracle.Mat(l["n"], l["nc"], *l["P"]) for l in levels]
    Go = oracle.amg(A0, Ps, [l["cf"] for l in levels], p.amg_relax_type, p.amg_A_drop_tol, p.amg_relax_weight)
    if not check:
        return Go, levels, cinv
    assert Go.nlevels == len(levels)
    mats = [l["A"] for l in levels] + [G.coarse_matrix()]
    for l in range(len(levels) + 1):
        Ao, Ro, do = Go.level(l)
        r, c, v = mats[l]
        assert np.array_equal(r, Ao.r) and np.array_equal(c, Ao.c[:Ao.nnz]), f"level {l} pattern"
        assert np.array_equal(v, Ao.v[:Ao.nnz]), f"level {l} values"
        if l < len(levels):
            rr, rc, rv = levels[l]["R"]
            assert np.array_equal(rr, Ro.r) and np.array_equal(rc, Ro.c[:Ro.nnz]) and np.array_equal(rv, Ro.v[:Ro.nnz])
            assert np.array_equal(levels[l]["d"], do), f"level {l} smoother diagonal"
        if l > 0 and p.amg_A_drop_tol > 0:
            lv = levels[l - 1]
            Al = sp.csr_matrix((lv["A"][2], lv["A"][1], lv["A"][0]), shape=(lv["n"], lv["n"]))
            P = sp.csr_matrix((lv["P"][2], lv["P"][1], lv["P"][0]), shape=(lv["n"], lv["nc"]))
            full = (P.T.tocsr() @ Al @ P).tocsr()
            full.sort_indices()
            nr = full.shape[0]
            rows = np.repeat(np.arange(nr), np.diff(full.indptr))
            mx = np.zeros(nr)
            np.maximum.at(mx, rows, np.abs(full.data))
            kept = np.isin(rows.astype(np.int64) * nr + full.indices,
                           np.repeat(np.arange(nr), np.diff(r)).astype(np.int64) * nr + c)
            thr = p.amg_A_drop_tol * mx[rows]
            off = full.indices != rows
            assert np.all(np.abs(full.data[~kept]) < thr[~kept] * (1 + 1e-10)), f"level {l}: a dropped entry is large"
            assert np.all(off[~kept]), f"level {l}: a diagonal was dropped"
            big = kept & off
            assert np.all(np.abs(full.data[big]) >= thr[big] * (1 - 1e-10)), f"level {l}: a kept entry is small"
    ci = Go.coarse_inverse()
    assert np.linalg.norm(cinv - ci) <= 1e-12 * np.linalg.norm(ci)
    return Go, levels, cinv
