"""The C-ABI library loads on a machine without a GPU and exports every declared symbol."""
import ctypes

import pytest

from deco_amd import _lib


def test_library_exports_every_header_symbol():
    lib = _lib.load()
    names = _lib.header_functions()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert sorted(_lib.SIGNATURES) == names


def test_error_string_is_available_without_gpu():
    lib = _lib.load()
    assert isinstance(lib.deco_last_error(), (bytes, type(None)))


def test_struct_layouts_match_c(tmp_path):
    """ctypes mirrors of the parameter/result structs have the C compiler's layout."""
    import os
    import subprocess
    src = tmp_path / "layout.c"
    structs = {"deco_cpr_params": _lib.CprParams, "deco_sol_params": _lib.SolParams,
               "deco_sol_info": _lib.SolInfo}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "deco_hip.h"', "int main(void){"]
    for cname, py in structs.items():
        lines.append(f'printf("%zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("%zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0;}")
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    inc = os.path.join(os.path.dirname(_lib.HEADER))
    subprocess.run(["gcc", "-I", inc, str(src), "-o", str(exe)], check=True)
    out = [int(t) for t in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = []
    for _, py in structs.items():
        want.append(ctypes.sizeof(py))
        want += [getattr(py, f).offset for f, _ in py._fields_]
    assert out == want


def test_profile_tree_report_format():
    """The phase tree re-enters phases by name and prints PROFILE::report's layout
    (core/profiler.hpp:89-112); no device is touched while none is initialised."""
    import deco_amd
    deco_amd.profile_reset()
    deco_amd.profile_enable(True)
    for _ in range(2):
        deco_amd.profile_start("setup & solve")
        deco_amd.profile_start("Prec")
        deco_amd.profile_stop()
        deco_amd.profile_start("Solve")
        deco_amd.profile_start("AMG apply")
        deco_amd.profile_stop()
        deco_amd.profile_stop()
        deco_amd.profile_stop()
    deco_amd.profile_enable(False)
    rep = deco_amd.profile_report()
    lines = [l for l in rep.splitlines() if "└" in l]
    assert rep.startswith("\n--------------------------------\nProfile of Execution Times (seconds):\n")
    assert [l.strip().split()[1] for l in lines] == ["setup", "Prec", "Solve", "AMG"]
    assert lines[0].startswith("└── setup & solve") and "100.00%" in lines[0]
    assert lines[1].startswith("    └── Prec") and lines[3].startswith("        └── AMG apply")
    t = deco_amd.profile_time("Solve")
    assert t >= deco_amd.profile_time("AMG apply") >= 0.0 and deco_amd.profile_time("nothing") == 0.0
    deco_amd.profile_reset()
    with pytest.raises(deco_amd._lib.DecoError):
        deco_amd.profile_enable(True)
        deco_amd.profile_stop()
    deco_amd.profile_enable(False)


def test_global_layout_switches_validate_without_a_device():
    """The library-wide switches are plain host settings: in-range values are accepted (and
    restored), out-of-range ones rejected with DECO_ERR_ARG, with no device present."""
    import deco_amd
    deco_amd.set_amg_dense_tail(0)
    deco_amd.set_amg_dense_tail(8192)
    deco_amd.set_amg_dense_tail(4608)
    for bad in (-1, 8193):
        with pytest.raises(deco_amd._lib.DecoError) as e:
            deco_amd.set_amg_dense_tail(bad)
        assert deco_amd._lib.ERRORS[e.value.code] == "DECO_ERR_ARG"
    deco_amd.set_sell_offsets16(False)
    deco_amd.set_sell_offsets16(True)
