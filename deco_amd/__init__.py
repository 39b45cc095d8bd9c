"""deco_amd: MI355X-native CPR-preconditioned Krylov solve path of deco.

The numerical work lives in libdeco_hip.so (hand-written HIP kernels for gfx950, C ABI in
include/deco_hip.h). Importing this package does not touch the GPU; the library is loaded on
first use and its absence is an error.
"""
from .ptree import PTree, remove_json_comments  # noqa: F401
from .core import (  # noqa: F401
    CSR, DeviceArray, as_dev, init, dot, nrm2, axpby, generate, cpr_params, sol_params, lib,
    Prec, Identity, Jacobi, FSAI, AMG, CPR, SOL, grad_A_pattern, amg_interp_host,
    profile_enable, profile_reset, profile_start, profile_stop, profile_time, profile_report,
    set_sell_padding, set_sell_offsets16, set_sell_dia, set_stream_pieces, set_stream_sorted, set_sell_pairs, set_amg_dense_tail, count_bytes, counted_bytes,
    time_outer_products, outer_product_times,
)
from ._lib import PREC_FP64, PREC_FP32, LAYOUT_BLOCK, LAYOUT_CSR, check  # noqa: F401
