"""The double-precision sin/cos/acos shared by the oracle and the FP64 device
path (include/vx3mi_math.h) against the platform libm, in ulps."""
import numpy as np
import pytest

import oracle as O


def ulps(a, b):
    return np.abs(a - b) / np.spacing(np.abs(b))


@pytest.mark.parametrize("name,lo,hi", [("sin", -70.0, 70.0), ("cos", -70.0, 70.0), ("sin", -1e-3, 1e-3),
                                         ("cos", -0.8, 0.8), ("acos", -1.0, 1.0), ("acos", 0.99, 1.0),
                                         ("acos", -1.0, -0.99)])
def test_within_two_ulp_of_libm(name, lo, hi):
    L = O.lib()
    f = getattr(L, "vxo_" + name)
    ref = getattr(np, "arccos" if name == "acos" else name)
    xs = np.random.default_rng(0).uniform(lo, hi, 20000)
    got = np.array([f(float(x)) for x in xs])
    want = ref(xs)
    mask = np.abs(want) > 1e-300
    assert np.max(ulps(got[mask], want[mask])) <= 2.0


def test_special_values():
    L = O.lib()
    assert L.vxo_sin(0.0) == 0.0 and L.vxo_cos(0.0) == 1.0
    assert L.vxo_acos(1.0) == 0.0 and L.vxo_acos(-1.0) == np.pi
    assert L.vxo_acos(0.0) == np.pi / 2
    assert np.isnan(L.vxo_sin(float("inf"))) and np.isnan(L.vxo_acos(float("nan")))
