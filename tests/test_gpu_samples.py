"""GPU CycleStep samples (angle, targetCloseness, numClosePairs; VX3_VoxelyzeKernel.cu:
314-334, 545-562) read by force fields and attach conditions, against the FP64
oracle: bit-identical states in flat and plain modes, and the runner's fitness
from the device samples.
"""
import numpy as np
import pytest

import oracle as O
import vx3mi
from vx3mi import vxa as W
from test_attach_oracle import stacked_cubes
from test_samples_oracle import free_bar

pytestmark = pytest.mark.gpu

FEEDBACK = {"z": ("mtMUL", ("mtCONST", 0.05), ("mtVAR", "angle")),
            "x": ("mtADD", ("mtCONST", 0.02), ("mtMUL", ("mtCONST", 1e-5), ("mtVAR", "targetCloseness"))),
            "y": ("mtADD", ("mtMUL", ("mtCONST", 40.0), ("mtVAR", "t")),
                  ("mtMUL", ("mtCONST", 0.003), ("mtVAR", "numClosePairs")))}


def log(msg):
    print(msg, flush=True)


def assert_same(b, ref, what):
    s, r = b.state(0), ref.state()
    for key in ("pos", "quat", "linmom", "angmom"):
        if not np.array_equal(s[key], r[key]):
            raise AssertionError(f"{what}: {key} differs (max {np.abs(s[key] - r[key]).max():.3e})")


@pytest.mark.parametrize("mode", [vx3mi.MODE_FLAT, vx3mi.MODE_PLAIN])
def test_force_fields_reading_the_samples_fp64_bit_exact(mode):
    prob = free_bar(FEEDBACK)
    ref = O.OracleModel(prob)
    cycle = int(0.002 / np.float64(ref.step_dt()))
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    for k in range(10):
        ref.step(cycle)
        b.step(cycle, mode=mode)
        s = ref.samples()
        log(f"cycle {k + 1}: angle {s['angle']:.3e} closeness {s['closeness']:.4f} pairs {s['pairs']}")
        assert_same(b, ref, f"cycle {k + 1}")
    assert ref.samples()["angle"] > 0 and ref.samples()["pairs"] == 1


def test_attach_condition_reading_the_angle_fp64_bit_exact():
    prob = stacked_cubes(temp_period=0.002, attach_conditions=[("mtLESSTHAN", ("mtVAR", "angle"), ("mtCONST", 2.0))])
    ref = O.OracleModel(prob)
    nl0 = ref.nl
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    ref.step(1200)
    b.step(1200, mode=vx3mi.MODE_FLAT)
    assert_same(b, ref, "attach on angle")
    links = b.attach_state(0)["links"]
    t = ref.topology()
    np.testing.assert_array_equal(links[:, 0], t["link_neg"][nl0:])
    np.testing.assert_array_equal(links[:, 1], t["link_pos"][nl0:])
    log(f"attach on angle: {len(links)} links")


def test_runner_fitness_from_device_samples():
    """vx3mi_run on a robot whose force field reads the angle: the stop and
    fitness formulas see the device's samples, which equal the oracle's."""
    prob = free_bar(FEEDBACK)
    ref = O.OracleModel(prob)
    dt = ref.step_dt()
    cycle = int(0.002 / np.float64(dt))
    n = 7 * cycle + 3
    t_stop = float(np.float64(dt) * (n - 0.5))
    text = W.write_vxa(prob, stop=("mtSUB", ("mtVAR", "t"), ("mtCONST", t_stop)),
                       fitness=("mtADD", ("mtVAR", "angle"), ("mtVAR", "targetCloseness")))
    m = vx3mi.Model.from_xml(text)
    b = vx3mi.Batch([m], vx3mi.FP64)
    res, _, _ = vx3mi.run(b, [m], history=False)
    ref.step(res[0]["steps"])
    s = ref.samples()
    assert res[0]["steps"] >= 7 * cycle
    assert res[0]["fitness_score"] == s["angle"] + s["closeness"]
    assert res[0]["num_close_pairs"] == s["pairs"]
    assert res[0]["angle_sample_times"] == s["count"]
    b.close()
