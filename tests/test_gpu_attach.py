"""GPU attach (handle_collision_attachment, VX3_VoxelyzeKernel.cu:729-831) against the
FP64 oracle, through the C ABI.

FP64: positions, orientations and momenta bit-identical to the oracle, and the
same links (ends, axis, updates left at 0.01 force) made in the same step, in
flat and plain modes.  FP32: the plain and flat modes agree bit for bit, make
the oracle's links, and stay within 1e-5 relative of it (positions, the
e_pos norm of test_gpu_parity).
"""
import numpy as np
import pytest

import oracle as O
import vx3mi
from test_attach_oracle import stacked_cubes, u_chain

pytestmark = pytest.mark.gpu

L0 = 0.01


def log(msg):
    print(msg, flush=True)


def moved_u(w, h):
    """A U chain with its right tip put next to its left one (positions as the
    oracle and the batch take them)."""
    prob = u_chain(w, h)
    ref = O.OracleModel(prob)
    lat = ref.topology()["lattice"]
    t1 = int(np.flatnonzero((lat[:, 0] == 0) & (lat[:, 1] == h))[0])
    t2 = int(np.flatnonzero((lat[:, 0] == w) & (lat[:, 1] == h))[0])
    pos = ref.state()["pos"].copy()
    pos[t2] = pos[t1] + np.array([L0, 0.0, 0.0])
    ref.set_state(pos=pos)
    return prob, ref, pos


def oracle_links(ref, nl0):
    t = ref.topology()
    _, left = ref.attach_state()
    return np.stack([t["link_neg"][nl0:], t["link_pos"][nl0:], t["link_axis"][nl0:], left[nl0:]], axis=1)


def assert_same_state(b, ref, robot, what):
    s, r = b.state(robot), ref.state()
    for key in ("pos", "quat", "linmom", "angmom"):
        if not np.array_equal(s[key], r[key]):
            bad = np.argwhere(s[key] != r[key])
            raise AssertionError(f"{what}: {key} differs at {bad[:4].tolist()} "
                                 f"(max {np.abs(s[key] - r[key]).max():.3e})")


@pytest.mark.parametrize("mode", [vx3mi.MODE_FLAT, vx3mi.MODE_PLAIN])
def test_falling_cube_attach_fp64_bit_exact(mode):
    prob = stacked_cubes()
    ref = O.OracleModel(prob)
    nl0 = ref.nl
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    made = None
    for chunk in range(12):  # 1,200 steps: the links come at ~860, then 340 updates at 0.01 force
        ref.step(100)
        b.step(100, mode=mode)
        log(f"mode {mode} steps {100 * (chunk + 1)}: oracle links {ref.nl - nl0}")
        assert_same_state(b, ref, 0, f"step {100 * (chunk + 1)}")
        st = b.attach_state(0)
        assert st["overflow"] == 0
        np.testing.assert_array_equal(st["links"], oracle_links(ref, nl0))
        np.testing.assert_array_equal(st["flags"], ref.attach_state()[0])
        if made is None and ref.nl > nl0:
            made = chunk
    assert made is not None and len(b.attach_state(0)["links"]) == 4


@pytest.mark.parametrize("w,h", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_touching_tips_fp64_bit_exact(w, h):
    prob, ref, pos = moved_u(w, h)
    nl0 = ref.nl
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    b.set_state(0, pos=pos)
    ref.step(1)
    b.step(1, mode=vx3mi.MODE_FLAT)
    np.testing.assert_array_equal(b.attach_state(0)["links"], oracle_links(ref, nl0))
    ref.step(299)
    b.step(299, mode=vx3mi.MODE_FLAT)
    assert_same_state(b, ref, 0, f"U {w}x{h}")
    np.testing.assert_array_equal(b.attach_state(0)["links"], oracle_links(ref, nl0))
    log(f"U {w}x{h}: {len(oracle_links(ref, nl0))} links")


@pytest.mark.parametrize("case", ["no_collision", "condition_t", "attach_off"])
def test_attach_variants_fp64_bit_exact(case):
    """Attach without contact forces (EnableCollision off), an AttachCondition
    on time, and attach off with a sticky material."""
    free = O.OracleModel(stacked_cubes())
    dt = free.step_dt()
    kw = dict(no_collision=dict(enable_collision=0),
              condition_t=dict(attach_conditions=[("mtGREATERTHAN", ("mtVAR", "t"), ("mtCONST", 1000 * dt))]),
              attach_off=dict(enable_attach=0))[case]
    prob = stacked_cubes(**kw)
    ref = O.OracleModel(prob)
    nl0 = ref.nl
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    ref.step(1300)
    b.step(1300, mode=vx3mi.MODE_FLAT)
    assert_same_state(b, ref, 0, case)
    np.testing.assert_array_equal(b.attach_state(0)["links"], oracle_links(ref, nl0))
    log(f"{case}: {ref.nl - nl0} links")


def test_mixed_batch_equals_single_runs_and_the_oracle():
    """Attach robots beside robots without attach in one flat batch: each robot
    as alone, and as the oracle."""
    pu, ref_u, pos_u = moved_u(3, 2)
    probs = [stacked_cubes(), stacked_cubes(mat=dict(sticky=0)), pu, stacked_cubes(enable_attach=0)]
    refs = [O.OracleModel(p) for p in probs[:2]] + [ref_u, O.OracleModel(probs[3])]
    nl0 = [r.nl for r in refs]
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    b.set_state(2, pos=pos_u)
    singles = [vx3mi.Batch([vx3mi.Model.from_problem(p)], vx3mi.FP64) for p in probs]
    singles[2].set_state(0, pos=pos_u)
    for r in refs:
        r.step(1100)
    b.step(1100, mode=vx3mi.MODE_FLAT)
    for s in singles:
        s.step(1100, mode=vx3mi.MODE_FLAT)
    for i, r in enumerate(refs):
        assert_same_state(b, r, i, f"robot {i}")
        assert_same_state(singles[i], r, 0, f"single {i}")
        np.testing.assert_array_equal(b.attach_state(i)["links"], oracle_links(r, nl0[i]))
    assert [len(b.attach_state(i)["links"]) for i in range(4)] == [4, 0, 1, 0]


def test_fp32_attach_modes_agree_and_follow_the_oracle():
    prob = stacked_cubes()
    ref = O.OracleModel(prob)
    nl0 = ref.nl
    flat = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP32)
    plain = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP32)
    worst = 0.0
    for chunk in range(12):
        ref.step(100)
        flat.step(100, mode=vx3mi.MODE_FLAT)
        plain.step(100, mode=vx3mi.MODE_PLAIN)
        sf, sp, r = flat.state(0), plain.state(0), ref.state()
        for key in ("pos", "quat", "linmom", "angmom"):
            np.testing.assert_array_equal(sf[key], sp[key])
        e = float(np.max(np.linalg.norm(sf["pos"] - r["pos"], axis=1)
                         / np.maximum(np.linalg.norm(r["pos"], axis=1), L0)))
        worst = max(worst, e)
        log(f"FP32 steps {100 * (chunk + 1)}: e_pos {e:.3e}, links {len(flat.attach_state(0)['links'])} "
            f"(oracle {ref.nl - nl0})")
    np.testing.assert_array_equal(flat.attach_state(0)["links"], oracle_links(ref, nl0))
    assert worst <= 1e-5


def test_autotune_keeps_the_attach_state():
    """The autotuner's trial steps are undone, links made by them included."""
    prob = stacked_cubes()
    a = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    a.step(850, mode=vx3mi.MODE_FLAT)
    b.step(850, mode=vx3mi.MODE_FLAT)
    tune = a.autotune(100)  # trial steps cross the attach step
    assert tune["mode"] == vx3mi.MODE_FLAT
    a.step(200)
    b.step(200, mode=vx3mi.MODE_FLAT)
    for key in ("pos", "quat", "linmom", "angmom"):
        np.testing.assert_array_equal(a.state(0)[key], b.state(0)[key])
    np.testing.assert_array_equal(a.attach_state(0)["links"], b.attach_state(0)["links"])


def test_attach_overflow_is_reported():
    """More attach candidates than a surface voxel keeps in one step (8 partners
    of lower index): a checkerboard of unlinked sticky voxels squeezed to 0.6 of
    the lattice spacing, up to 9 such partners each.  The robot's overflow flag
    is set; the links made are distinct, join voxels inside the envelope at the
    step's start, use each voxel slot once, and are no more than the oracle
    (which keeps every candidate) makes in that step."""
    d = 6
    z, y, x = np.meshgrid(np.arange(d), np.arange(d), np.arange(d), indexing="ij")
    cells = (((x + y + z) % 2) == 0).astype(np.int8).reshape(-1)
    prob = dict(dims=(d, d, d), cells=cells, materials=[dict(E=1e6, rho=1000, sticky=1)], enable_attach=1,
                lattice_dim=L0, floor_enabled=0, grav_acc=0.0, slow_damp_z=0.0, vary_temp=0)
    ref = O.OracleModel(prob)
    pos = ref.state()["pos"]
    pos = pos.mean(axis=0) + 0.6 * (pos - pos.mean(axis=0))
    ref.set_state(pos=pos)
    nl0 = ref.nl
    ref.step(1)
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    b.set_state(0, pos=pos)
    b.step(1, mode=vx3mi.MODE_FLAT)
    st = b.attach_state(0)
    links = st["links"]
    log(f"overflow {st['overflow']}, links {len(links)} (oracle {ref.nl - nl0})")
    assert st["overflow"] == 1
    assert 0 < len(links) <= ref.nl - nl0
    pairs = {(int(n), int(p)) for n, p in links[:, :2]}
    assert len(pairs) == len(links)
    dist = np.linalg.norm(pos[links[:, 0]] - pos[links[:, 1]], axis=1)
    assert (dist < 2 * 0.625 * L0).all()
    slots = [(int(n), 2 * int(a)) for n, _, a, _ in links] + [(int(p), 2 * int(a) + 1) for _, p, a, _ in links]
    assert len(set(slots)) == len(slots)
    assert not b.diverged().any()
