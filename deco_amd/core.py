"""Host-side mirror of deco's C++ API over the C ABI of libdeco_hip.so.

Names follow the reference: CSR (core/CSR.hpp), CPR (precond/CPR.hpp), FSAI
(precond/FSAI/FSAI.hpp), Jacobi (precond/Jacobi/Jacobi.hpp), AMG in the role of hypre_AMG
(precond/hypre_AMG/hypre_AMG.hpp) and SOL (solver/SOL.hpp). Constructors take a PTree subtree
with the reference's keys. All numerical work runs in the HIP library; this module only moves
buffers and parameters.
"""
import ctypes as C
import numpy as np

from . import _lib
from ._lib import check, CprParams, SolParams, SolInfo
from .ptree import PTree


def lib():
    return _lib.load()


def init(device=0):
    check(lib().deco_init(device))


def _vp(a):
    return a.ctypes.data_as(C.c_void_p)


class DeviceArray:
    """A device buffer owned by this object (hipMalloc / hipFree through the library)."""

    def __init__(self, n, dtype=np.float64):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(lib().deco_malloc(max(self.nbytes, 8), C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        if a.size:
            check(lib().deco_memcpy_h2d(d.ptr, _vp(a), d.nbytes))
        return d

    @classmethod
    def zeros(cls, n, dtype=np.float64):
        d = cls(n, dtype)
        check(lib().deco_memset(d.ptr, 0, max(d.nbytes, 8)))
        return d

    def numpy(self):
        a = np.empty(self.n, self.dtype)
        if self.n:
            check(lib().deco_memcpy_d2h(_vp(a), self.ptr, self.nbytes))
        return a

    def copy_from(self, a):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.size != self.n:
            raise ValueError(f"copy_from: {a.size} values for a device array of {self.n}")
        check(lib().deco_memcpy_h2d(self.ptr, _vp(a), self.nbytes))

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().deco_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def as_dev(x, dtype=np.float64):
    """x as a device array of dtype: a DeviceArray is passed through (it must hold dtype), host
    data is uploaded."""
    if isinstance(x, DeviceArray):
        if x.dtype != np.dtype(dtype):
            raise TypeError(f"device array of {x.dtype} where {np.dtype(dtype)} is needed")
        return x
    return DeviceArray.from_numpy(np.asarray(x, dtype=dtype))


# ---------------------------------------------------------------------------- CSR ---------
class CSR:
    """Device CSR<int,int,double> (core/CSR.hpp:3-75)."""

    def __init__(self, n=None, m=None, r=None, c=None, v=None, handle=None, owner=True):
        if handle is not None:
            self.h = handle
        else:
            r = np.ascontiguousarray(r, np.int32)
            c = np.ascontiguousarray(c, np.int32)
            v = np.ascontiguousarray(v, np.float64)
            h = C.c_void_p()
            check(lib().deco_csr_create(int(n), int(m), int(c.size), _vp(r), _vp(c), _vp(v), C.byref(h)))
            self.h = h.value
        self.owner = owner

    @classmethod
    def from_scipy(cls, S):
        S = S.tocsr()
        S.sort_indices()
        return cls(S.shape[0], S.shape[1], S.indptr, S.indices, S.data)

    def info(self):
        n, m, nz = C.c_int(), C.c_int(), C.c_int()
        check(lib().deco_csr_info(self.h, C.byref(n), C.byref(m), C.byref(nz)))
        return n.value, m.value, nz.value

    @property
    def n(self):
        return self.info()[0]

    @property
    def m(self):
        return self.info()[1]

    @property
    def nz(self):
        return self.info()[2]

    def download(self):
        n, m, nz = self.info()
        r = np.empty(n + 1, np.int32)
        c = np.empty(nz, np.int32)
        v = np.empty(nz, np.float64)
        check(lib().deco_csr_download(self.h, _vp(r), _vp(c), _vp(v)))
        return r, c, v

    def update_values(self, v):
        """New values on the same pattern (host array of nz entries); drops what was derived from
        the old ones (shadows, cached transpose)."""
        v = np.ascontiguousarray(v, np.float64)
        if v.size != self.nz:
            raise ValueError(f"{v.size} values for a pattern of {self.nz} entries")
        check(lib().deco_csr_update_values(self.h, _vp(v)))

    def to_scipy(self):
        import scipy.sparse as sp
        r, c, v = self.download()
        n, m, _ = self.info()
        return sp.csr_matrix((v, c, r), shape=(n, m))

    def apply(self, x, bs=1, block=None):
        """CSR::apply(x, b, bs) (core/CSR.hpp:132-136); block=(R, C) runs the product on the
        FP64 block shadow of A."""
        n, m, _ = self.info()
        xd = as_dev(x)
        if block is not None:
            y = DeviceArray(n)
            check(lib().deco_spmv_block(self.h, int(block[0]), int(block[1]), xd.ptr, y.ptr))
            return y
        y = DeviceArray(n // bs)
        check(lib().deco_spmv(self.h, xd.ptr, y.ptr, bs))
        return y

    def residual(self, x, z):
        n, _, _ = self.info()
        xd, zd = as_dev(x), as_dev(z)
        y = DeviceArray(n)
        check(lib().deco_spmv_residual(self.h, xd.ptr, zd.ptr, y.ptr))
        return y

    def apply_transpose(self, x):
        n, m, _ = self.info()
        xd = as_dev(x)
        y = DeviceArray(m)
        check(lib().deco_spmv_transpose(self.h, xd.ptr, y.ptr))
        return y

    def apply_f32(self, x, block=None):
        """y = A x with FP32 values; block=(R, C) uses the FP32 block shadow of A."""
        n, m, _ = self.info()
        xd = as_dev(x, np.float32)
        y = DeviceArray(n, np.float32)
        if block is None:
            check(lib().deco_spmv_f32(self.h, xd.ptr, y.ptr))
        else:
            check(lib().deco_spmv_block_f32(self.h, int(block[0]), int(block[1]), xd.ptr, y.ptr))
        return y

    def spmv_time(self, precision=_lib.PREC_FP64, iters=20, block=None):
        """Average product duration in ms (hipEvents); with block=(R, C), (ms, blocks) of the
        block-shadow product."""
        ms = C.c_double()
        if block is None:
            check(lib().deco_spmv_time(self.h, precision, iters, C.byref(ms)))
            return ms.value
        nb = C.c_int()
        check(lib().deco_spmv_block_time(self.h, int(block[0]), int(block[1]), precision, iters, C.byref(ms),
                                         C.byref(nb)))
        return ms.value, nb.value

    def block_info(self, R, C_):
        """(layout, blocks, stored blocks) of the R x C_ block shadow (created if needed); layout
        0 chunked, 1 sliced, 2 sliced with 16-bit block-column offsets, 3 sliced with uniform
        offsets (no stored columns)."""
        sl, nb, ent = C.c_int(), C.c_int(), C.c_longlong()
        check(lib().deco_block_info(self.h, int(R), int(C_), C.byref(sl), C.byref(nb), C.byref(ent)))
        return sl.value, nb.value, ent.value

    def copy(self):
        h = C.c_void_p()
        check(lib().deco_csr_copy(self.h, C.byref(h)))
        return CSR(handle=h.value)

    def block_reorder(self, bs, row_order, col_order, b=None):
        """AD::block_reorder (adapter/csr2coo.cpp:19-88) in place; b (DeviceArray) is permuted
        like the rows when given."""
        check(lib().deco_block_reorder(self.h, bs, row_order.encode(), col_order.encode(),
                                       b.ptr if b is not None else None))

    def filter(self, tol):
        check(lib().deco_csr_filter(self.h, tol))

    def filter_proper(self, tol):
        check(lib().deco_csr_filter_proper(self.h, tol))

    def local_matrix(self, bs=3, p=0, weights=None):
        h = C.c_void_p()
        wd = as_dev(weights) if weights is not None else None
        check(lib().deco_csr_local_matrix(self.h, bs, p, wd.ptr if wd else None, C.byref(h)))
        return CSR(handle=h.value)

    def transpose(self):
        h = C.c_void_p()
        check(lib().deco_csr_transpose(self.h, C.byref(h)))
        return CSR(handle=h.value)

    def __matmul__(self, other):
        h = C.c_void_p()
        check(lib().deco_spgemm(self.h, other.h, C.byref(h)))
        return CSR(handle=h.value)

    def tril_square_pattern(self):
        h = C.c_void_p()
        check(lib().deco_spgemm_tril_pattern(self.h, C.byref(h)))
        return CSR(handle=h.value)

    def __del__(self):
        try:
            if self.owner and getattr(self, "h", None):
                lib().deco_csr_destroy(self.h)
                self.h = None
        except Exception:
            pass


# ---------------------------------------------------------------------------- BLAS -------
def dot(x, y):
    xd, yd = as_dev(x), as_dev(y)
    out = C.c_double()
    check(lib().deco_dot(xd.n, xd.ptr, yd.ptr, C.byref(out)))
    return out.value


def nrm2(x):
    xd = as_dev(x)
    out = C.c_double()
    check(lib().deco_nrm2(xd.n, xd.ptr, C.byref(out)))
    return out.value


def axpby(a, x, b, y):
    """y = b*y + a*x in place on the DeviceArray y."""
    xd = as_dev(x)
    check(lib().deco_axpby(y.n, a, xd.ptr, b, y.ptr))
    return y


# ---------------------------------------------------------------------------- generator --
def generate(cfg, seed=0, custom=None, grid=None):
    """Seeded synthetic reservoir system (host arrays): dict with r, c, v, b, xstar, n, nnz, bs.
    custom=(nx,ny,nz,bs) uses the 32x32x16 physics; grid=(nx,ny,nz) keeps cfg's physics."""
    L = lib()
    h = C.c_void_p()
    if grid is not None:
        check(L.deco_gen_scaled(cfg, grid[0], grid[1], grid[2], seed, C.byref(h)))
    elif custom is not None:
        nx, ny, nz, bs = custom
        check(L.deco_gen_custom(nx, ny, nz, bs, seed, C.byref(h)))
    else:
        check(L.deco_gen_create(cfg, seed, C.byref(h)))
    try:
        n, nnz, bs, nc = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(L.deco_gen_info(h, C.byref(n), C.byref(nnz), C.byref(bs), C.byref(nc)))
        pr, pc, pv, pb, px = (C.c_void_p() for _ in range(5))
        check(L.deco_gen_arrays(h, C.byref(pr), C.byref(pc), C.byref(pv), C.byref(pb), C.byref(px)))
        n, nnz = n.value, nnz.value

        def grab(p, cnt, ct, dt):
            if cnt == 0:
                return np.empty(0, dt)
            return np.ctypeslib.as_array(C.cast(p, C.POINTER(ct)), shape=(cnt,)).astype(dt, copy=True)

        out = dict(
            r=grab(pr, n + 1, C.c_int, np.int32), c=grab(pc, nnz, C.c_int, np.int32),
            v=grab(pv, nnz, C.c_double, np.float64), b=grab(pb, n, C.c_double, np.float64),
            xstar=grab(px, n, C.c_double, np.float64), n=n, nnz=nnz, bs=bs.value, ncells=nc.value)
    finally:
        L.deco_gen_destroy(h)
    return out


# ---------------------------------------------------------------------------- params -----
_WEIGHTS = {"none": _lib.W_NONE, "quasiimpes": _lib.W_QUASIIMPES, "quasiimpes_GE": _lib.W_QUASIIMPES_GE}
_LAYOUTS = {"block": _lib.LAYOUT_BLOCK, "csr": _lib.LAYOUT_CSR}


def cpr_params(prec, precision=_lib.PREC_FP64):
    """deco_cpr_params from the reference's "prec" subtree (keys read at precond/CPR.hpp:49-65,
    118,130; FSAI.hpp:19-20,34; Jacobi.hpp:18; hypre_AMG.tpp:48-66)."""
    if isinstance(prec, dict):
        prec = PTree(prec)
    p = CprParams()
    p.block_size = prec.get("block_size", int)
    p.pressure_id = prec.get("pressure_id", int, 0)
    p.seq = prec.get("seq", int, 3)
    wt = prec.get("weights", str, "quasiimpes_GE")
    if wt not in _WEIGHTS:
        raise ValueError("Unknown weights type " + wt)
    p.weights = _WEIGHTS[wt]
    p.use_drs = prec.get("use_drs", int, 0)
    p.drs_e_dd = prec.get("drs.e_dd", float, 0.2)
    p.drs_e_ps = prec.get("drs.e_ps", float, 0.0)
    p.cpd = prec.get("CPD", int, 1)
    p.tol_zero = prec.get("tol_zero", float, 0.0)
    glo = prec.get_child("prec_glo") if prec.has("prec_glo") else PTree({})
    gtype = glo.get("type", str, "FSAI")
    if gtype not in ("FSAI", "Jacobi"):
        raise ValueError("Unknown global prec.prec_glo.type: " + gtype)
    p.glo_type = _lib.GLO_FSAI if gtype == "FSAI" else _lib.GLO_JACOBI
    p.glo_tol_rel = glo.get("tol_rel", float, 0.0)
    p.glo_tol_abs = glo.get("tol_abs", float, 0.0)
    p.glo_power = glo.get("power", int, 1)
    p.glo_block_size = glo.get("block_size", int, 1)
    loc = prec.get_child("prec_loc") if prec.has("prec_loc") else PTree({})
    p.loc_tol_rel = loc.get("tol_rel", float, 0.0)
    p.amg_coarse_type = loc.get("amg_coarse_type", int, 10)
    p.amg_relax_type = loc.get("amg_relax_type", int, -1)
    p.amg_agg_levs = loc.get("amg_agg_levs", int, 0)
    p.amg_agg_path = loc.get("amg_agg_path", int, 1)
    p.amg_A_drop_tol = loc.get("amg_A_drop_tol", float, 0.0)
    p.amg_prol_trunc_fac = loc.get("amg_prol_trunc_fac", float, 0.0)
    p.amg_SoC_alpha = loc.get("amg_SoC_alpha", float, 0.25)
    p.amg_max_coarse = loc.get("amg_max_coarse", int, 300)
    p.amg_max_levels = loc.get("amg_max_levels", int, 25)
    p.amg_P_max_elmts = loc.get("amg_P_max_elmts", int, 4)
    p.amg_max_row_sum = loc.get("amg_max_row_sum", float, 0.9)
    p.amg_relax_weight = loc.get("amg_relax_weight", float, 0.0)
    p.precision = precision
    lay = prec.get("layout", str, "block")
    if lay not in _LAYOUTS:
        raise ValueError("Unknown layout " + lay)
    p.layout = _LAYOUTS[lay]
    return p


def sol_params(sol):
    if isinstance(sol, dict):
        sol = PTree(sol)
    s = SolParams()
    t = sol.get("type", str)
    if t not in _lib.SOL_TYPES:
        raise ValueError("Invalid solver type: " + t)
    s.type = _lib.SOL_TYPES[t]
    s.reltol = sol.get("reltol", float)
    s.restart = sol.get("restart", int, 30)
    s.itmax = sol.get("itmax", int)
    s.verbosity = sol.get("verbosity", int, 0)
    s.use_x0 = sol.get("use_x0", int, 0)
    ortho = sol.get("ortho", str, "dcgs2")
    if ortho not in _lib.ORTHO:
        raise ValueError("Invalid Gram-Schmidt variant: " + ortho)
    s.ortho = _lib.ORTHO[ortho]
    s.flexible = sol.get("flexible", int, 0)
    return s


# ---------------------------------------------------------------------------- preconditioners
def _view(h, parent):
    v = CSR(handle=h, owner=False)
    v._parent = parent
    return v


class Prec:
    """A device preconditioner handle: apply(r) = M^{-1} r (CPRBase::apply, precond/CPR.hpp:3-8)."""

    def __init__(self, h, owner=True, parent=None):
        self.h = h
        self.owner = owner
        self._parent = parent

    @property
    def n(self):
        v = C.c_int()
        check(lib().deco_prec_size(self.h, C.byref(v)))
        return v.value

    def apply(self, r, out=None):
        rd = as_dev(r)
        x = out if out is not None else DeviceArray(self.n)
        check(lib().deco_prec_apply(self.h, rd.ptr, x.ptr))
        return x

    def apply_t(self, r, out=None):
        rd = as_dev(r)
        x = out if out is not None else DeviceArray(self.n)
        check(lib().deco_prec_apply_transpose(self.h, rd.ptr, x.ptr))
        return x

    def set_precision(self, precision):
        check(lib().deco_prec_set_precision(self.h, precision))

    def set_layout(self, layout):
        """Storage of the products in either precision: LAYOUT_BLOCK (cell-block shadows, the
        default) or LAYOUT_CSR (the CSR arrays, the reference's format and summation order)."""
        check(lib().deco_prec_set_layout(self.h, layout))

    def __del__(self):
        try:
            if self.owner and getattr(self, "h", None):
                lib().deco_prec_destroy(self.h)
                self.h = None
        except Exception:
            pass


class Identity(Prec):
    def __init__(self, n):
        h = C.c_void_p()
        check(lib().deco_identity_create(int(n), C.byref(h)))
        super().__init__(h.value)


class Jacobi(Prec):
    """Jacobi(param, A) (precond/Jacobi/Jacobi.hpp:15-25); param is a prec_glo-like subtree."""

    def __init__(self, param, A):
        if isinstance(param, dict):
            param = PTree(param)
        p = CprParams()
        p.glo_block_size = param.get("block_size", int, 1)
        h = C.c_void_p()
        check(lib().deco_jacobi_create(C.byref(p), A.h, C.byref(h)))
        super().__init__(h.value, parent=A)


class FSAI(Prec):
    """FSAI(param, A) (precond/FSAI/FSAI.hpp:14-37); param is the whole "prec" tree or the
    "prec_glo" subtree."""

    def __init__(self, param, A, handle=None, parent=None):
        if handle is not None:
            super().__init__(handle, owner=False, parent=parent)
            return
        if isinstance(param, dict):
            param = PTree(param)
        tree = param if param.has("prec_glo") else PTree({"block_size": "1", "prec_glo": param.to_dict()})
        p = cpr_params(tree)
        h = C.c_void_p()
        check(lib().deco_fsai_create(C.byref(p), A.h, C.byref(h)))
        super().__init__(h.value, parent=A)

    def factors(self):
        fl, fu = C.c_void_p(), C.c_void_p()
        check(lib().deco_fsai_factors(self.h, C.byref(fl), C.byref(fu)))
        return _view(fl.value, self), _view(fu.value, self)


class AMG(Prec):
    """BoomerAMG-like pressure preconditioner in the role of hypre_AMG(param, Ap)
    (precond/hypre_AMG/hypre_AMG.hpp:16-21)."""

    def __init__(self, param, Ap, handle=None, parent=None):
        if handle is not None:
            super().__init__(handle, owner=False, parent=parent)
            return
        if isinstance(param, dict):
            param = PTree(param)
        tree = param if param.has("prec_loc") else PTree({"block_size": "1", "prec_loc": param.to_dict()})
        p = cpr_params(tree)
        h = C.c_void_p()
        check(lib().deco_amg_create(C.byref(p), Ap.h, C.byref(h)))
        super().__init__(h.value, parent=Ap)

    def export(self):
        """Host copy of the hierarchy: (levels, coarse_inverse) for the CPU V-cycle."""
        L = lib()
        nl = C.c_int()
        check(L.deco_amg_num_levels(self.h, C.byref(nl)))
        levels = []
        for lvl in range(nl.value):
            a, p, r, d, cf = (C.c_void_p() for _ in range(5))
            check(L.deco_amg_level(self.h, lvl, C.byref(a), C.byref(p), C.byref(r), C.byref(d), C.byref(cf)))
            A = _view(a.value, self)
            n = A.n
            dd = np.empty(n)
            cc = np.empty(n, np.int32)
            check(L.deco_memcpy_d2h(_vp(dd), d.value, 8 * n))
            check(L.deco_memcpy_d2h(_vp(cc), cf.value, 4 * n))
            levels.append(dict(A=A.download(), P=_view(p.value, self).download(), R=_view(r.value, self).download(),
                               n=n, nc=_view(p.value, self).m, d=dd, cf=cc))
        nc = C.c_int()
        inv = C.c_void_p()
        check(L.deco_amg_coarse_inverse(self.h, C.byref(nc), C.byref(inv)))
        ci = np.empty(nc.value * nc.value)
        if nc.value:
            check(L.deco_memcpy_d2h(_vp(ci), inv.value, 8 * ci.size))
        return levels, ci.reshape(nc.value, nc.value)

    def coarse_matrix(self):
        """Host copy (r, c, v) of the coarsest matrix."""
        L = lib()
        nl = C.c_int()
        check(L.deco_amg_num_levels(self.h, C.byref(nl)))
        a = C.c_void_p()
        check(L.deco_amg_level(self.h, nl.value, C.byref(a), None, None, None, None))
        return _view(a.value, self).download()

    def dense_tail(self):
        """Rows of the dense operator that replaces the last level's sub-cycle (0: none)."""
        rows = C.c_int()
        check(lib().deco_amg_dense_tail(self.h, C.byref(rows)))
        return rows.value


class CPR(Prec):
    """CPR<int,int,double,FSAI|Jacobi,AMG>(param, A, b) (precond/CPR.hpp:11-172).

    As in the reference the setup decouples A and b in place (b may be None)."""

    def __init__(self, param, A, b=None, precision=_lib.PREC_FP64):
        if isinstance(param, dict):
            param = PTree(param)
        p = cpr_params(param, precision)
        h = C.c_void_p()
        check(lib().deco_cpr_create(C.byref(p), A.h, b.ptr if b is not None else None, C.byref(h)))
        super().__init__(h.value, parent=A)
        self.bs = p.block_size
        self.A = A

    def pressure_matrix(self):
        h = C.c_void_p()
        check(lib().deco_cpr_pressure_matrix(self.h, C.byref(h)))
        return _view(h.value, self)

    def fsai(self):
        h = C.c_void_p()
        check(lib().deco_cpr_fsai(self.h, C.byref(h)))
        return FSAI(None, None, handle=h.value, parent=self)

    def amg(self):
        h = C.c_void_p()
        check(lib().deco_cpr_amg(self.h, C.byref(h)))
        return AMG(None, None, handle=h.value, parent=self)

    def left_transform(self):
        W = np.empty(self.A.n * self.bs)
        check(lib().deco_cpr_left_transform(self.h, _vp(W)))
        return W

    def transform_transpose(self, mu, out=None):
        """lambda = W^T mu (W the decoupling's left transform), into out (a DeviceArray) or a new one."""
        md = as_dev(mu)
        lam = out if out is not None else DeviceArray(md.n)
        check(lib().deco_cpr_transform_transpose(self.h, md.ptr, lam.ptr))
        return lam


# ---------------------------------------------------------------------------- solver -----
class SOL:
    """SOL(param, A, precond, b, x) (solver/SOL.hpp:19-39): the constructor runs the solve on the
    device. Results: iter, res_abs, res_rel, flag, x (DeviceArray), info (timings, counts)."""

    def __init__(self, param, A, precond, b, x=None, transpose=False):
        sp = sol_params(param)
        bd = as_dev(b)
        if x is not None:
            xd = as_dev(x)
            if not isinstance(x, DeviceArray):
                sp.use_x0 = 1
        else:
            xd = DeviceArray(A.n)
            sp.use_x0 = 0
        info = SolInfo()
        fn = lib().deco_solve_transpose if transpose else lib().deco_solve
        check(fn(C.byref(sp), A.h, precond.h if precond is not None else None, bd.ptr, xd.ptr, C.byref(info)))
        self.x = xd
        self.info = info
        self.iter, self.res_abs, self.res_rel, self.flag = info.iter, info.res_abs, info.res_rel, info.flag

    def get_result(self):
        return self.x


def set_sell_padding(max_fraction):
    """Block shadows made from now on take the sliced layout when it pads the blocks by at most
    max_fraction (default 0.10); a negative value selects the chunked block CSR always."""
    check(lib().deco_set_sell_padding(float(max_fraction)))


def set_sell_offsets16(enable):
    """Sliced block shadows made from now on store their block columns as 16-bit offsets from
    the block row when every offset fits (default), else (or when disabled) 32-bit columns."""
    check(lib().deco_set_sell_offsets16(1 if enable else 0))


def set_stream_pieces(enable):
    """Stream products of matrices without a block shadow made from now on use the row-piece
    sliced form where it fits (default), else the CSR-stream kernel."""
    check(lib().deco_set_stream_pieces(1 if enable else 0))


def set_sell_pairs(mode):
    """FP32 products on uniform-offset sliced block shadows with two slices per wavefront: 0
    never, 1 when a batch holds at most 8 values per lane (default), 2 always; the same sums
    bit for bit."""
    check(lib().deco_set_sell_pairs(int(mode)))


def set_stream_sorted(mode):
    """When CSR-stream products of matrices without a block shadow read the column-sorted chunk
    form (the same sums bit for bit, gathers of one wavefront on a few cache lines): 1 or True
    by the library's rule (default), 2 wherever it applies, 0 or False never."""
    check(lib().deco_set_stream_sorted(int(mode)))


def set_sell_dia(enable):
    """Block shadows made from now on take the uniform-offset sliced layout (no stored columns)
    where the matrix has at most 16 distinct block-column offsets within the padding threshold
    (default), else the plain sliced or chunked one."""
    check(lib().deco_set_sell_dia(1 if enable else 0))


def set_amg_dense_tail(max_rows):
    """AMG hierarchies set up from now on replace the whole sub-cycle of their highest level of at
    most max_rows rows with one dense operator (default 4608, at most 8192; 0 disables)."""
    check(lib().deco_set_amg_dense_tail(int(max_rows)))


# ---------------------------------------------------------------------------- profiler ---
def profile_enable(on=True):
    """Phase timer tree under the reference's labels (PROFILE, core/profiler.hpp); device waits
    at every phase boundary while enabled."""
    check(lib().deco_profile_enable(1 if on else 0))


def profile_reset():
    check(lib().deco_profile_reset())


def profile_start(name):
    check(lib().deco_profile_start(name.encode()))


def profile_stop():
    check(lib().deco_profile_stop())


def profile_time(name):
    t = C.c_double()
    check(lib().deco_profile_time(name.encode(), C.byref(t)))
    return t.value


def count_bytes(on=True):
    """Start (or stop) the library's algorithmic byte count of the device work it queues; both
    reset the count."""
    check(lib().deco_count_bytes(1 if on else 0))


def counted_bytes():
    """(bytes, launches) counted since count_bytes()."""
    b, n = C.c_double(), C.c_longlong()
    check(lib().deco_counted_bytes(C.byref(b), C.byref(n)))
    return b.value, n.value


def time_outer_products(on=True):
    """Start (or stop) event timing of the solvers' products with the system matrix; both reset."""
    check(lib().deco_time_outer_products(1 if on else 0))


def outer_product_times():
    """(total ms, count) of the products with the system matrix timed since time_outer_products()."""
    t, n = C.c_double(), C.c_longlong()
    check(lib().deco_outer_product_times(C.byref(t), C.byref(n)))
    return t.value, n.value


def profile_report():
    need = C.c_size_t()
    check(lib().deco_profile_report(None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value)
    check(lib().deco_profile_report(buf, need.value, None))
    return buf.value.decode()


def grad_A_pattern(A, lam, x, out=None):
    """dL/dA_ij = -lambda_i x_j on A's pattern (device), into out (a DeviceArray) or a new one."""
    ld, xd = as_dev(lam), as_dev(x)
    dA = out if out is not None else DeviceArray(A.nz)
    check(lib().deco_grad_A_pattern(A.h, ld.ptr, xd.ptr, dA.ptr))
    return dA


def amg_interp_host(param, n, r, c, v):
    """One AMG setup level on the host (strength, C/F split, ext+i interpolation)."""
    if isinstance(param, dict):
        param = PTree(param)
    tree = param if param.has("prec_loc") else PTree({"block_size": "1", "prec_loc": param.to_dict()})
    p = cpr_params(tree)
    r = np.ascontiguousarray(r, np.int32)
    c = np.ascontiguousarray(c, np.int32)
    v = np.ascontiguousarray(v, np.float64)
    cf = np.empty(n, np.int32)
    nc, pnnz = C.c_int(), C.c_int()
    pr, pc, pv = C.c_void_p(), C.c_void_p(), C.c_void_p()
    L = lib()
    check(L.deco_amg_interp_host(C.byref(p), n, _vp(r), _vp(c), _vp(v), _vp(cf), C.byref(nc), C.byref(pnnz),
                                 C.byref(pr), C.byref(pc), C.byref(pv)))
    k = pnnz.value
    Pr = np.ctypeslib.as_array(C.cast(pr, C.POINTER(C.c_int)), shape=(n + 1,)).copy()
    Pc = np.ctypeslib.as_array(C.cast(pc, C.POINTER(C.c_int)), shape=(max(k, 1),))[:k].copy()
    Pv = np.ctypeslib.as_array(C.cast(pv, C.POINTER(C.c_double)), shape=(max(k, 1),))[:k].copy()
    for q in (pr, pc, pv):
        L.deco_host_free(q)
    return cf, nc.value, (Pr, Pc, Pv)
