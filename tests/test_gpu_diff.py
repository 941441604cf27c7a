"""Differentiable mode on the GPU: forward parity with the FP64 oracle's
differentiable mode, adjoint gradients against FP64 central finite
differences (fixture tests/golden/diff_fd_4x4x4.npz, written by
tests/golden/make_diff_fd.py from the oracle), exactness of checkpointing with
recomputation, batching, and the torch.autograd.Function wrapper.

Gradient error norm: ||g - g_fd|| / ||g_fd|| over all parameters of a family
(phase in radians, log10 E); bar 1e-4.
"""
import math
import os
import sys

import numpy as np
import pytest

import oracle as O
import vx3mi
from vx3mi import configs

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_diff_fd as FD  # noqa: E402

FIX = np.load(os.path.join(HERE, "golden", "diff_fd_4x4x4.npz"))
FIX_C4 = np.load(os.path.join(HERE, "golden", "diff_fd_c4_robot0.npz"))


def log(msg):
    print(msg, flush=True)


def e_pos(p, ref, L=0.01):
    return float(np.max(np.linalg.norm(p - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), L)))


def make(prob, precision, **opts):
    m = vx3mi.Model.from_problem(prob)
    return m, vx3mi.Diff(m, precision, **opts)


@pytest.mark.parametrize("precision,tol", [(vx3mi.FP64, 1e-10), (vx3mi.FP32, 1e-5)])
@pytest.mark.parametrize("dims,nsteps", [((4, 4, 4), 300), ((8, 8, 8), 1000)])
def test_forward_matches_oracle(precision, tol, dims, nsteps):
    log(f"diff forward parity {dims} {nsteps} steps fp{precision}")
    prob = configs.config3(dims=dims)
    phase, E = np.asarray(prob["phase"]), np.asarray(prob["voxel_E"], dtype=np.float64)
    m, d = make(prob, precision, dt=float(FIX["dt"]))
    d.set_params(phase, E)
    d.forward(nsteps)
    o = O.OracleDiff(prob, phase, E, dt=float(FIX["dt"]))
    o.step(nsteps)
    s, r = d.state(), o.state()
    err = e_pos(s["pos"], r["pos"])
    qerr = float(np.max(np.minimum(np.linalg.norm(s["quat"] - r["quat"], axis=1),
                                   np.linalg.norm(s["quat"] + r["quat"], axis=1))))
    log(f"  e_pos {err:.3e} e_q {qerr:.3e}")
    assert err < tol and qerr < 10 * tol


def adjoint_grads(precision, K=0):
    prob, psi, log10E = FD.subset_problem()
    m, d = make(prob, precision, dt=float(FIX["dt"]), checkpoint_every=K)
    E = 10.0 ** log10E
    d.set_params(psi / (2 * math.pi), E)
    d.forward(int(FIX["nsteps"]))
    masses = np.full(m.nv, 1.0)
    gpos = np.zeros((m.nv, 3))
    gpos[:, 0] = -masses / masses.sum()
    gphase, gE = d.backward(grad_pos=gpos)
    return gphase / (2 * math.pi), gE * E * math.log(10.0), d


@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
def test_adjoint_matches_finite_differences(precision):
    log(f"adjoint vs FD fp{precision}")
    g_psi, g_l, d = adjoint_grads(precision)
    rel_psi = np.linalg.norm(g_psi - FIX["grad_psi"]) / np.linalg.norm(FIX["grad_psi"])
    rel_l = np.linalg.norm(g_l - FIX["grad_log10E"]) / np.linalg.norm(FIX["grad_log10E"])
    log(f"  rel err phase {rel_psi:.3e} log10E {rel_l:.3e} info {d.info()}")
    assert rel_psi < 1e-4 and rel_l < 1e-4


@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
def test_checkpointing_with_recompute_is_exact(precision):
    """K = whole run (tape written by the forward) and K = 7 (recompute per
    segment from checkpoints) give bit-identical gradients."""
    a = adjoint_grads(precision, K=0)
    b = adjoint_grads(precision, K=7)
    assert b[2].info()["checkpoint_every"] == 7
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


def test_backward_leaves_final_state_and_repeats():
    prob = configs.config3(dims=(4, 4, 4))
    m, d = make(prob, vx3mi.FP64, dt=float(FIX["dt"]), checkpoint_every=9)
    d.forward(50)
    s0 = d.state()["pos"].copy()
    g = np.ones((m.nv, 3))
    r1 = d.backward(grad_pos=g)
    assert np.array_equal(d.state()["pos"], s0)
    r2 = d.backward(grad_pos=g)
    assert np.array_equal(r1[0], r2[0]) and np.array_equal(r1[1], r2[1])


def test_batched_equals_single():
    """Two robots of different shapes in one batch: states and gradients equal
    the ones of each robot alone, bit for bit."""
    probs = [configs.config3(seed=1, dims=(4, 3, 3)), configs.config3(seed=2, dims=(3, 5, 2))]
    models = [vx3mi.Model.from_problem(p) for p in probs]
    params = [(np.asarray(p["phase"]), np.asarray(p["voxel_E"], np.float64)) for p in probs]
    both = vx3mi.Diff(models, vx3mi.FP32, dt=float(FIX["dt"]))
    both.set_params(np.concatenate([a for a, _ in params]), np.concatenate([b for _, b in params]))
    both.forward(120)
    g = np.random.default_rng(0).normal(size=(both.nv, 3))
    gb = both.backward(grad_pos=g)
    off = both.voxel_offsets
    for r, mdl in enumerate(models):
        one = vx3mi.Diff(mdl, vx3mi.FP32, dt=float(FIX["dt"]))
        one.set_params(*params[r])
        one.forward(120)
        assert np.array_equal(one.state()["pos"], both.state(r)["pos"])
        go = one.backward(grad_pos=g[off[r]:off[r + 1]])
        assert np.array_equal(go[0], gb[0][off[r]:off[r + 1]]) and np.array_equal(go[1], gb[1][off[r]:off[r + 1]])


def test_batched_checkpointed_equals_whole_tape():
    """Two robots, K = 13 over 120 steps (recompute on a second stream beside
    the reverse sweep, two tapes in turn) against the whole run's tape."""
    probs = [configs.config3(seed=3, dims=(4, 4, 3)), configs.config3(seed=4, dims=(5, 3, 2))]
    models = [vx3mi.Model.from_problem(p) for p in probs]
    phase = np.concatenate([np.asarray(p["phase"]) for p in probs])
    E = np.concatenate([np.asarray(p["voxel_E"], np.float64) for p in probs])
    out = []
    for K in (0, 13):
        d = vx3mi.Diff(models, vx3mi.FP32, dt=float(FIX["dt"]), checkpoint_every=K)
        d.set_params(phase, E)
        d.forward(120)
        g = np.random.default_rng(1).normal(size=(d.nv, 3))
        out.append(d.backward(grad_pos=g))
        log(f"  K={K} info {d.info()}")
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_reverse_sweep_width_does_not_change_gradients():
    """512 and 1024 threads per robot in the reverse sweep, and the flat sweep
    (every robot's reverse step at once, one item per thread, -1), give the
    same bits (each work item is one thread's; the gather order is fixed)."""
    prob = configs.config3(dims=(5, 4, 4))
    m = vx3mi.Model.from_problem(prob)
    out = []
    for rt in (512, 1024, -1):
        d = vx3mi.Diff(m, vx3mi.FP32, dt=float(FIX["dt"]), checkpoint_every=17, reverse_threads=rt)
        d.set_params(np.asarray(prob["phase"]), np.asarray(prob["voxel_E"], np.float64))
        d.forward(150)
        out.append(d.backward(grad_pos=np.ones((m.nv, 3))))
    for o in out[1:]:
        assert np.array_equal(out[0][0], o[0]) and np.array_equal(out[0][1], o[1])


@pytest.mark.parametrize("precision", [vx3mi.FP32, vx3mi.FP64])
def test_flat_reverse_sweep_equals_per_robot_sweep(precision):
    """A batch of robots of different sizes (the config-4 draw): the flat reverse
    sweep gives the per-robot sweep's gradients bit for bit, with recompute from
    checkpoints (K = 13 over 60 steps)."""
    probs = configs.config4(n=6, seed=3)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    phase = np.concatenate([np.asarray(p["phase"], np.float64) for p in probs])
    E = np.concatenate([np.asarray(p["voxel_E"], np.float64) for p in probs])
    out = []
    for rt in (0, -1):
        d = vx3mi.Diff(models, precision, checkpoint_every=13, reverse_threads=rt)
        d.set_params(phase, E)
        d.forward(60)
        g = np.random.default_rng(4).normal(size=(d.nv, 3))
        out.append(d.backward(grad_pos=g))
        d.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.abs(out[0][0]).max() > 0


def test_torch_function_gradients():
    torch = pytest.importorskip("torch")
    from vx3mi.torch_diff import simulate

    prob, psi, log10E = FD.subset_problem()
    m, d = make(prob, vx3mi.FP64, dt=float(FIX["dt"]))
    psi_t = torch.tensor(psi, requires_grad=True)
    l_t = torch.tensor(log10E, requires_grad=True)
    pos = simulate(d, psi_t / (2 * math.pi), 10.0 ** l_t, int(FIX["nsteps"]))
    lat = m.topology()["lattice"][:, 0] * 0.01
    loss = -(pos[:, 0].mean() - float(lat.mean()))
    loss.backward()
    assert float(loss) == pytest.approx(float(FIX["loss"]), rel=1e-8)
    rel_psi = np.linalg.norm(psi_t.grad.numpy() - FIX["grad_psi"]) / np.linalg.norm(FIX["grad_psi"])
    rel_l = np.linalg.norm(l_t.grad.numpy() - FIX["grad_log10E"]) / np.linalg.norm(FIX["grad_log10E"])
    assert rel_psi < 1e-4 and rel_l < 1e-4


# ---- configuration 4: the batched adjoint workload ---------------------------------
@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
def test_config4_robot_subset_matches_finite_differences(precision):
    """The 4x4x4 part of configuration 4's robot 0 (52 voxels of its largest
    component there, the robot's own phases and moduli), 200 steps: adjoint
    gradients against the FP64 central differences of the oracle."""
    prob, psi, log10E = FD.c4_subset_problem()
    assert np.array_equal(np.asarray(prob["cells"]), FIX_C4["cells"])
    m, d = make(prob, precision, dt=float(FIX_C4["dt"]))
    E = 10.0 ** log10E
    d.set_params(psi / (2 * math.pi), E)
    d.forward(int(FIX_C4["nsteps"]))
    gpos = np.zeros((m.nv, 3))
    gpos[:, 0] = -1.0 / m.nv  # one material: equal masses
    gphase, gE = d.backward(grad_pos=gpos)
    g_psi, g_l = gphase / (2 * math.pi), gE * E * math.log(10.0)
    rel_psi = np.linalg.norm(g_psi - FIX_C4["grad_psi"]) / np.linalg.norm(FIX_C4["grad_psi"])
    rel_l = np.linalg.norm(g_l - FIX_C4["grad_log10E"]) / np.linalg.norm(FIX_C4["grad_log10E"])
    log(f"config4 subset fp{precision}: rel err phase {rel_psi:.3e} log10E {rel_l:.3e}")
    assert rel_psi < 1e-4 and rel_l < 1e-4


def _com_x_grad(nv):
    g = np.zeros((nv, 3))
    g[:, 0] = -1.0 / nv
    return g


@pytest.mark.parametrize("nrobots,steps", [(16, 400), (256, 150)])
def test_config4_batched_equals_single(nrobots, steps):
    """Configuration 4 (dims 6..16, per-voxel phase and modulus): the first 16
    robots over 400 steps with checkpoints every 64 and recompute, and the full
    256-robot batch over 150 steps; every robot's final state and gradients of
    its own centre-of-mass loss equal a run of that robot alone, bit for bit
    (16 robots: all of them; 256: four spread over the batch)."""
    probs = configs.config4()[:nrobots]
    models = [vx3mi.Model.from_problem(p) for p in probs]
    params = [(np.asarray(p["phase"]), np.asarray(p["voxel_E"], np.float64)) for p in probs]
    K = 64
    both = vx3mi.Diff(models, vx3mi.FP32, dt=float(FIX["dt"]), checkpoint_every=K)
    both.set_params(np.concatenate([a for a, _ in params]), np.concatenate([b for _, b in params]))
    both.forward(steps)
    off = both.voxel_offsets
    g = np.concatenate([_com_x_grad(m.nv) for m in models])
    gb = both.backward(grad_pos=g)
    log(f"config4 {nrobots} robots, {both.nv} voxels: info {both.info()}")
    check = range(nrobots) if nrobots <= 16 else (0, 85, 170, nrobots - 1)
    for r in check:
        one = vx3mi.Diff(models[r], vx3mi.FP32, dt=float(FIX["dt"]), checkpoint_every=K)
        one.set_params(*params[r])
        one.forward(steps)
        assert np.array_equal(one.state()["pos"], both.state(r)["pos"]), r
        go = one.backward(grad_pos=g[off[r]:off[r + 1]])
        assert np.array_equal(go[0], gb[0][off[r]:off[r + 1]]), r
        assert np.array_equal(go[1], gb[1][off[r]:off[r + 1]]), r
        assert np.all(np.isfinite(go[0])) and np.linalg.norm(go[0]) > 0, r
