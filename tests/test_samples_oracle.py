"""The CycleStep samples of doTimeStep in the oracle (VX3_VoxelyzeKernel.cu:314-334,
545-562): which steps sample, recentAngle and targetCloseness / numClosePairs
against numpy on the oracle's own positions, and force fields reading them.
"""
import numpy as np
import pytest

import oracle as O

L0 = 0.01


def free_bar(force_field=None, **kw):
    """A 3x1x1 bar without gravity or floor, pushed along x with a y force that
    grows with time (its centre of mass bends away from a straight line); the
    last voxel is a target, as is the first."""
    mats = [dict(E=1e6, rho=1000), dict(E=1e6, rho=1000, is_target=1)]
    ff = {"x": ("mtCONST", 0.02), "y": ("mtMUL", ("mtCONST", 40.0), ("mtVAR", "t"))}
    ff.update(force_field or {})
    p = dict(dims=(3, 1, 1), cells=np.array([2, 1, 2], dtype=np.int8), materials=mats, lattice_dim=L0,
             floor_enabled=0, grav_acc=0.0, slow_damp_z=0.0, temp_period=0.002, enable_collision=0,
             force_field=ff, max_dist_pairs=2.5)
    p.update(kw)
    return p


def com_of(m):
    return m.state()["pos"].mean(axis=0)  # one material: the mass-weighted mean


def test_samples_come_every_cycle_step_with_the_reference_formulas():
    m = O.OracleModel(free_bar())
    dt = m.step_dt()
    cycle = int(0.002 / np.float64(dt))
    assert cycle > 5
    coms = [com_of(m)]  # InitializeCenterOfMass, then one per sample
    for k in range(1, 6 * cycle + 1):
        m.step(1)
        s = m.samples()
        assert s["count"] == k // cycle, k
        if k % cycle:
            continue
        coms.append(com_of(m))
        np.testing.assert_allclose(s["com"], coms[-1], rtol=1e-13, atol=1e-18)
        if len(coms) < 4:
            assert s["angle"] == 0.0  # the first two samples (angleSampleTimes < 3)
        else:
            A, B, C = coms[-3], coms[-2], coms[-1]
            u, w = B - A, C - B
            want = np.arccos(u @ w / (np.linalg.norm(u) * np.linalg.norm(w)))
            assert want > 1e-4
            assert s["angle"] == pytest.approx(want, rel=1e-9, abs=1e-12)
        p = m.state()["pos"]
        d = np.linalg.norm(p[0] - p[2])
        assert s["closeness"] == pytest.approx(1 / d, rel=1e-14)
        assert s["pairs"] == int(d < 2.5 * L0)


def test_no_closeness_without_a_pair_radius():
    m = O.OracleModel(free_bar(max_dist_pairs=0.0))
    m.step(200)
    s = m.samples()
    assert s["count"] > 0 and s["closeness"] == 0.0 and s["pairs"] == 0


def test_a_force_field_reads_the_sampled_angle():
    # F_z = 0.05 * angle: nothing until the third sample sets a non-zero angle
    plain = O.OracleModel(free_bar())
    fed = O.OracleModel(free_bar({"z": ("mtMUL", ("mtCONST", 0.05), ("mtVAR", "angle"))}))
    cycle = int(0.002 / np.float64(plain.step_dt()))
    plain.step(3 * cycle)
    fed.step(3 * cycle)
    np.testing.assert_array_equal(plain.state()["pos"], fed.state()["pos"])
    assert fed.samples()["angle"] > 0
    plain.step(1)
    fed.step(1)
    assert np.abs(fed.state()["linmom"][:, 2]).max() > 0
    assert np.abs(plain.state()["linmom"][:, 2]).max() == 0
