"""hpcgx_dot_noalloc (the dot product over the library's fixed device scratch
array, no allocation) against hpcgx_dot on a problem (the same kernels over
the problem's own partials): bit for bit, on two calls in a row, the second
with a smaller input, and then a zero-length call.  numpy's f64 dot is a
sanity bound (1e-13 relative) for both."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu] {msg}", flush=True)


def test_dot_noalloc_matches_dot(hip):
    G = hip.Problem(32, 32, 32, nlevels=1, stores=hip.STORE_F64)
    rng = np.random.default_rng(5)
    n1 = 3_000_017                     # 2930 reduction blocks
    n2 = 40_001                        # 40 blocks: the first call's later partials stay in the array
    x = rng.uniform(-1.0, 1.0, n1)
    y = rng.uniform(-1.0, 1.0, n1)
    dx, dy = hip.DeviceArray.from_host(x), hip.DeviceArray.from_host(y)
    for n in (n1, n2):
        _say(f"dot_noalloc n={n}")
        got = hip.dot_noalloc(dx, dy, n)
        want = G.dot(dx, dy, n=n)
        ref = float(np.dot(x[:n], y[:n]))
        _say(f"n={n}: noalloc {got!r}, problem dot {want!r}, numpy {ref!r}")
        assert got == want
        assert abs(got - ref) <= 1e-13 * float(np.dot(np.abs(x[:n]), np.abs(y[:n])))
    assert hip.dot_noalloc(None, None, 0) == 0.0
    assert G.dot(dx, dy, n=0) == 0.0
    dx.free()
    dy.free()
    G.close()


def test_dot_noalloc_rejects_bad_arguments(hip):
    res = ctypes.c_double()
    assert hip.lib.hpcgx_dot_noalloc(-1, None, None, ctypes.byref(res), None) == -2
    assert b"negative length" in hip.lib.hpcgx_last_error()
    assert hip.lib.hpcgx_dot_noalloc(8, None, None, ctypes.byref(res), None) == -2
    assert b"null pointer" in hip.lib.hpcgx_last_error()
    assert hip.lib.hpcgx_dot_noalloc(0, None, None, None, None) == -2
