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
        assert a.size == self.n
        check(lib().deco_memcpy_h2d(self.ptr, _vp(a), self.nbytes))

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().deco_free(self.ptr)
                self.ptr = None
        except Exception:
            pass


def as_dev(x, dtype=np.float64):
    if isinstance(x, DeviceArray):
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

    def to_scipy(self):
        import scipy.sparse as sp
        r, c, v = self.download()
        n, m, _ = self.info()
        return sp.csr_matrix((v, c, r), shape=(n, m))

    def apply(self, x, bs=1):
        """CSR::apply(x, b, bs) (core/CSR.hpp:132-136)."""
        n, m, _ = self.info()
        xd = as_dev(x)
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

    def apply_f32(self, x):
        n, m, _ = self.info()
        xd = as_dev(x, np.float32)
        y = DeviceArray(n, np.float32)
        check(lib().deco_spmv_f32(self.h, xd.ptr, y.ptr))
        return y

    def spmv_time(self, precision=_lib.PREC_FP64, iters=20):
        ms = C.c_double()
        check(lib().deco_spmv_time(self.h, precision, iters, C.byref(ms)))
        return ms.value

    def copy(self):
        h = C.c_void_p()
        check(lib().deco_csr_copy(self.h, C.byref(h)))
        return CSR(handle=h.value)

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
def generate(cfg, seed=0, custom=None):
    """Seeded synthetic reservoir system (host arrays): dict with r, c, v, b, xstar, n, nnz, bs."""
    L = lib()
    h = C.c_void_p()
    if custom is not None:
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
    p.precision = precision
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
    return s
