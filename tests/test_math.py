"""The double-precision sin/cos/acos shared by the oracle and the FP64 device
path (include/vx3mi_math.h) against the platform libm, in ulps; the FP32
rotation primitives of the timestep against their FP64 forms."""
import os

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


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


def test_fp32_rotation_primitives_agree_with_fp64(tmp_path):
    """The FP32 forms of the rotation primitives (Maclaurin series for small
    rotations, trigonometry-free align-to-x rotation, reciprocal square roots)
    agree with the FP64 forms, which follow the reference's formulas."""
    import ctypes as C
    import subprocess
    out = tmp_path / "fp32_primitives_check.so"
    csrc = os.path.join(ROOT, "voxcraft-sim_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", "-I" + csrc,
                    "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "fp32_primitives_check.cpp"),
                    "-o", str(out)], check=True)
    lib = C.CDLL(str(out))
    res = (C.c_double * 3)()
    lib.check_primitives(C.c_ulonglong(7), 200000, res)
    assert res[0] < 1e-6 and res[1] < 1e-6 and res[2] < 1e-6, list(res)  # summed over 4 components: ~1 ulp each
    # the actuation sine of the FP32 path: about 1 ulp of the sine of the same
    # float argument, unbiased; a reduction to the wrong side or a dropped
    # series term is off by 1e-2 or more, rounded constants alone bias it
    lib.check_sin_cycles.restype = C.c_double
    mean = C.c_double()
    worst = lib.check_sin_cycles(1 << 20, C.byref(mean))
    assert worst < 2e-7 and abs(mean.value) < 3e-9, (worst, mean.value)
    # the rotation-vector angle of the FP32 differentiable mode: a few float ulps
    # of atan2 over the half circle (an octant folded the wrong way is off by
    # 1e-1 or more)
    lib.check_atan2.restype = C.c_double
    worst = lib.check_atan2(1 << 20)
    assert worst < 5e-7, worst
