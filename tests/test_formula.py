"""MathTree formulas: the postfix program the device evaluates (formula_eval,
after mt_postfix) gives the same bits as the FIFO evaluation of the reference's
token order (VX3_MathTree::eval, src/Utils/VX3_MathTree.h:50-190) on random
formula trees, operand order of every non-commutative operator included."""
import ctypes as C
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    out = tmp_path_factory.mktemp("formula") / "formula_check.so"
    csrc = os.path.join(ROOT, "voxcraft-sim_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off", "-I" + csrc,
                    "-I" + os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "formula_check.cpp"),
                    os.path.join(csrc, "vx3mi_formula.cpp"), "-o", str(out)], check=True)
    lib = C.CDLL(str(out))
    lib.check_random.restype = C.c_int
    lib.check_random.argtypes = [C.c_uint64, C.c_int, C.c_int, C.POINTER(C.c_int)]
    return lib


@pytest.mark.parametrize("depth", [1, 3, 6])
def test_postfix_matches_fifo_evaluation(harness, depth):
    skipped = C.c_int()
    bad = harness.check_random(1000 + depth, 3000, depth, C.byref(skipped))
    assert bad == 0
    assert skipped.value == 0  # trees of depth <= 6 fit the 16-value stack
