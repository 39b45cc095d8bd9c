"""deco_dot_noalloc (the allocation-free dot product) through the C ABI against deco_dot: the same
kernels and reduction grid give the same value bit for bit, on two calls in a row, the second with
a smaller input (the scratch array keeps the first call's partials, which the second must not
read), and an error for a negative length."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_dot_noalloc_matches_dot(gpu):
    import deco_amd
    from deco_amd._lib import DecoError

    L = deco_amd.lib()
    rng = np.random.default_rng(77)
    for n in (1_000_003, 4_097):
        x = deco_amd.DeviceArray.from_numpy(rng.standard_normal(n))
        y = deco_amd.DeviceArray.from_numpy(rng.standard_normal(n))
        a, b = C.c_double(), C.c_double()
        deco_amd.check(L.deco_dot(n, x.ptr, y.ptr, C.byref(a)))
        deco_amd.check(L.deco_dot_noalloc(n, x.ptr, y.ptr, C.byref(b)))
        assert a.value == b.value, (n, a.value, b.value)
        ref = float(np.dot(x.numpy(), y.numpy()))
        assert abs(b.value - ref) <= 1e-12 * np.sqrt(n) * np.abs(x.numpy()).max() * np.abs(y.numpy()).max()
    z = C.c_double()
    assert L.deco_dot_noalloc(0, None, None, C.byref(z)) == 0 and z.value == 0.0
    with pytest.raises(DecoError):
        deco_amd.check(L.deco_dot_noalloc(-1, x.ptr, y.ptr, C.byref(z)))
