"""Launch configurations and the autotuner on the GPU: every configuration
(large and small tiles, persistent workgroups of 128/256/512 threads, flat
graphs) produces the same bits; the autotuner leaves the state untouched and
its choice is what MODE_AUTO then runs."""
import numpy as np
import pytest

import vx3mi
from vx3mi import configs

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)


def run(models, precision, steps, mode, tile=0, threads=0):
    b = vx3mi.Batch(models, precision)
    if mode in (vx3mi.MODE_TILED, vx3mi.MODE_PERSISTENT):
        b.configure(mode, tile, threads)
        mode = vx3mi.MODE_AUTO
    b.step(steps, mode=mode)
    s = b.state()
    t = b.time()
    b.close()
    return s, t


@pytest.mark.parametrize("precision", [vx3mi.FP32, vx3mi.FP64])
def test_every_launch_configuration_gives_the_same_bits(precision):
    # robots small enough for the small tile in both precisions (FP64: <= 384
    # voxels / 512 links per tile)
    probs = configs.config2(seed=11, n=96, lo=4, hi=7)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    ref = run(models, precision, 250, vx3mi.MODE_FLAT)
    cases = [(vx3mi.MODE_TILED, 0, 0), (vx3mi.MODE_TILED, 1, 0), (vx3mi.MODE_PERSISTENT, 0, 128),
             (vx3mi.MODE_PERSISTENT, 0, 256), (vx3mi.MODE_PERSISTENT, 0, 512)]
    for mode, tile, threads in cases:
        log(f"config mode {mode} tile {tile} threads {threads} fp{precision}")
        got = run(models, precision, 250, mode, tile, threads)
        for k in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(ref[0][k], got[0][k]), (mode, tile, threads, k)
        assert np.array_equal(ref[1], got[1])


def test_autotune_restores_state_and_auto_runs_the_choice():
    probs = configs.config2(seed=12, n=128, lo=4, hi=10)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    b = vx3mi.Batch(models, vx3mi.FP32)
    b.step(50)
    before, tb = b.state(), b.time()
    info = b.autotune(60)
    log(f"autotune {info}")
    after, ta = b.state(), b.time()
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    assert np.array_equal(tb, ta)
    assert info["ms"]["persistent-256"] is not None and info["ms"]["flat-graph"] is not None
    assert min(v for v in info["ms"].values() if v is not None) > 0
    # AUTO now runs the chosen configuration: same bits as running it explicitly
    b.step(100)
    auto = b.state()
    b.close()
    ref = vx3mi.Batch(models, vx3mi.FP32)
    ref.step(50)
    ref.configure(info["mode"], info["tile"], info["threads"])
    ref.step(100)
    for k in auto:
        assert np.array_equal(auto[k], ref.state()[k]), k


def test_configure_rejects_an_impossible_tile():
    probs = [configs.config1(seed=3, dims=(12, 12, 6))] * 2  # collisions on: the large tile only
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
    b.configure(vx3mi.MODE_TILED, 0, 0)
    with pytest.raises(vx3mi.VX3MIError):
        b.configure(vx3mi.MODE_TILED, 1, 0)
    big = [configs.config1(seed=3, dims=(20, 20, 8))] * 2  # beyond a collision tile's 1,248 voxels
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in big], vx3mi.FP32)
    for tile in (0, 1):
        with pytest.raises(vx3mi.VX3MIError):
            b.configure(vx3mi.MODE_TILED, tile, 0)
