"""ctypes binding of libdeco_hip.so (the C ABI declared in include/deco_hip.h).

The shared library is loaded from this package directory only; there is no fallback: if it is
missing the import of the binding raises, so nothing can silently run a CPU path.
"""
import ctypes as C
import importlib.util
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libdeco_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "deco_hip.h")

DECO_OK = 0
ERRORS = {
    -1: "DECO_ERR_ARG", -2: "DECO_ERR_DEVICE", -3: "DECO_ERR_SINGULAR", -4: "DECO_ERR_BREAKDOWN",
    -5: "DECO_ERR_UNSUPPORTED", -6: "DECO_ERR_NOMEM", -7: "DECO_ERR_STATE",
}

W_NONE, W_QUASIIMPES, W_QUASIIMPES_GE = 0, 1, 2
GLO_FSAI, GLO_JACOBI = 0, 1
PREC_FP64, PREC_FP32 = 0, 1
LAYOUT_BLOCK, LAYOUT_CSR = 0, 1
SOL_TYPES = {"rgmres": 0, "lgmres": 1, "bicgstab": 2, "lbicgstab": 3, "pcg": 4}
ORTHO = {"dcgs2": 0, "mgs": 1}


class DecoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class CprParams(C.Structure):
    _fields_ = [
        ("block_size", C.c_int), ("pressure_id", C.c_int), ("seq", C.c_int), ("weights", C.c_int),
        ("use_drs", C.c_int), ("drs_e_dd", C.c_double), ("drs_e_ps", C.c_double), ("cpd", C.c_int),
        ("tol_zero", C.c_double),
        ("glo_type", C.c_int), ("glo_tol_rel", C.c_double), ("glo_tol_abs", C.c_double),
        ("glo_power", C.c_int), ("glo_block_size", C.c_int),
        ("loc_tol_rel", C.c_double), ("amg_coarse_type", C.c_int), ("amg_relax_type", C.c_int),
        ("amg_agg_levs", C.c_int), ("amg_agg_path", C.c_int), ("amg_A_drop_tol", C.c_double),
        ("amg_prol_trunc_fac", C.c_double), ("amg_SoC_alpha", C.c_double),
        ("amg_max_coarse", C.c_int), ("amg_max_levels", C.c_int), ("amg_P_max_elmts", C.c_int),
        ("amg_max_row_sum", C.c_double), ("precision", C.c_int), ("amg_relax_weight", C.c_double),
        ("layout", C.c_int),
    ]


class SolParams(C.Structure):
    _fields_ = [("type", C.c_int), ("reltol", C.c_double), ("restart", C.c_int), ("itmax", C.c_int),
                ("verbosity", C.c_int), ("use_x0", C.c_int), ("ortho", C.c_int), ("flexible", C.c_int)]


class SolInfo(C.Structure):
    _fields_ = [("iter", C.c_int), ("res_abs", C.c_double), ("res_rel", C.c_double), ("flag", C.c_int),
                ("bnorm", C.c_double), ("time_ms", C.c_double), ("n_prec", C.c_int), ("n_spmv", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


vp = C.c_void_p
ip = C.POINTER(C.c_int)
dp = C.POINTER(C.c_double)
i = C.c_int
d = C.c_double
sz = C.c_size_t
pp = C.POINTER(C.c_void_p)

# name -> (restype, argtypes); opaque handles and device pointers travel as void*
SIGNATURES = {
    "deco_init": (i, [i]),
    "deco_last_error": (C.c_char_p, []),
    "deco_sync": (i, []),
    "deco_stream": (vp, []),
    "deco_malloc": (i, [sz, pp]),
    "deco_free": (i, [vp]),
    "deco_memcpy_h2d": (i, [vp, vp, sz]),
    "deco_memcpy_d2h": (i, [vp, vp, sz]),
    "deco_memcpy_d2d": (i, [vp, vp, sz]),
    "deco_memset": (i, [vp, i, sz]),
    "deco_csr_create": (i, [i, i, i, vp, vp, vp, pp]),
    "deco_csr_create_dev": (i, [i, i, i, vp, vp, vp, pp]),
    "deco_csr_destroy": (i, [vp]),
    "deco_csr_info": (i, [vp, ip, ip, ip]),
    "deco_csr_download": (i, [vp, vp, vp, vp]),
    "deco_csr_device_ptrs": (i, [vp, pp, pp, pp]),
    "deco_csr_update_values": (i, [vp, vp]),
    "deco_csr_copy": (i, [vp, pp]),
    "deco_csr_filter": (i, [vp, d]),
    "deco_csr_filter_proper": (i, [vp, d]),
    "deco_csr_local_matrix": (i, [vp, i, i, vp, pp]),
    "deco_csr_transpose": (i, [vp, pp]),
    "deco_block_reorder": (i, [vp, i, C.c_char_p, C.c_char_p, vp]),
    "deco_spmv": (i, [vp, vp, vp, i]),
    "deco_spmv_residual": (i, [vp, vp, vp, vp]),
    "deco_spmv_transpose": (i, [vp, vp, vp]),
    "deco_spmv_f32": (i, [vp, vp, vp]),
    "deco_spmv_block_f32": (i, [vp, i, i, vp, vp]),
    "deco_spmv_time": (i, [vp, i, i, dp]),
    "deco_spmv_block": (i, [vp, i, i, vp, vp]),
    "deco_read_bandwidth": (i, [C.c_size_t, i, dp]),
    "deco_set_sell_padding": (i, [C.c_double]),
    "deco_block_info": (i, [vp, i, i, ip, ip, C.POINTER(C.c_longlong)]),
    "deco_set_sell_offsets16": (i, [i]),
    "deco_set_sell_dia": (i, [i]),
    "deco_set_stream_pieces": (i, [i]),
    "deco_set_sell_pairs": (i, [i]),
    "deco_set_stream_sorted": (i, [i]),
    "deco_spmv_block_time": (i, [vp, i, i, i, i, dp, ip]),
    "deco_spgemm": (i, [vp, vp, pp]),
    "deco_spgemm_tril_pattern": (i, [vp, pp]),
    "deco_dot": (i, [i, vp, vp, dp]),
    "deco_dot_noalloc": (i, [i, vp, vp, dp]),
    "deco_nrm2": (i, [i, vp, dp]),
    "deco_axpby": (i, [i, d, vp, d, vp]),
    "deco_cpr_create": (i, [C.POINTER(CprParams), vp, vp, pp]),
    "deco_fsai_create": (i, [C.POINTER(CprParams), vp, pp]),
    "deco_jacobi_create": (i, [C.POINTER(CprParams), vp, pp]),
    "deco_amg_create": (i, [C.POINTER(CprParams), vp, pp]),
    "deco_identity_create": (i, [i, pp]),
    "deco_prec_destroy": (i, [vp]),
    "deco_prec_apply": (i, [vp, vp, vp]),
    "deco_prec_apply_transpose": (i, [vp, vp, vp]),
    "deco_prec_size": (i, [vp, ip]),
    "deco_prec_set_precision": (i, [vp, i]),
    "deco_prec_set_layout": (i, [vp, i]),
    "deco_cpr_pressure_matrix": (i, [vp, pp]),
    "deco_cpr_fsai": (i, [vp, pp]),
    "deco_cpr_amg": (i, [vp, pp]),
    "deco_cpr_left_transform": (i, [vp, vp]),
    "deco_fsai_factors": (i, [vp, pp, pp]),
    "deco_amg_num_levels": (i, [vp, ip]),
    "deco_amg_level": (i, [vp, i, pp, pp, pp, pp, pp]),
    "deco_amg_coarse_inverse": (i, [vp, ip, pp]),
    "deco_amg_dense_tail": (i, [vp, ip]),
    "deco_set_amg_dense_tail": (i, [i]),
    "deco_amg_interp_host": (i, [C.POINTER(CprParams), i, vp, vp, vp, vp, ip, ip, pp, pp, pp]),
    "deco_host_free": (None, [vp]),
    "deco_workspace_release": (i, []),
    "deco_profile_enable": (i, [i]),
    "deco_profile_reset": (i, []),
    "deco_profile_start": (i, [C.c_char_p]),
    "deco_profile_stop": (i, []),
    "deco_profile_time": (i, [C.c_char_p, dp]),
    "deco_profile_report": (i, [C.c_char_p, sz, C.POINTER(C.c_size_t)]),
    "deco_count_bytes": (i, [i]),
    "deco_counted_bytes": (i, [C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "deco_time_outer_products": (i, [i]),
    "deco_outer_product_times": (i, [C.POINTER(C.c_double), C.POINTER(C.c_longlong)]),
    "deco_solve": (i, [C.POINTER(SolParams), vp, vp, vp, vp, C.POINTER(SolInfo)]),
    "deco_solve_transpose": (i, [C.POINTER(SolParams), vp, vp, vp, vp, C.POINTER(SolInfo)]),
    "deco_grad_A_pattern": (i, [vp, vp, vp, vp]),
    "deco_cpr_transform_transpose": (i, [vp, vp, vp]),
    "deco_gen_create": (i, [i, C.c_uint64, pp]),
    "deco_gen_custom": (i, [i, i, i, i, C.c_uint64, pp]),
    "deco_gen_scaled": (i, [i, i, i, i, C.c_uint64, pp]),
    "deco_gen_info": (i, [vp, ip, ip, ip, ip]),
    "deco_gen_arrays": (i, [vp, pp, pp, pp, pp, pp]),
    "deco_gen_destroy": (i, [vp]),
}


def header_functions(path=HEADER):
    """Names of every function the C header declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(deco_[A-Za-z0-9_]+)\s*\(", text)))


_lib = None


def load():
    """Load libdeco_hip.so from the package directory (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    # PyTorch-ROCm opens its bundled libamdhip64.so by RPATH. When torch is installed it is
    # loaded first, so that this library binds to the same HIP runtime (same soname) instead
    # of the system one: two runtimes in one process cannot share the device.
    if importlib.util.find_spec("torch") is not None:
        import torch  # noqa: F401
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def check(code):
    if code != DECO_OK:
        msg = load().deco_last_error()
        raise DecoError(code, msg.decode() if msg else "")
    return code
