"""GPU parity: the HIP timestep (through the C ABI) against the FP64 oracle.

Tolerances (written per test): FP64 device state within 1e-12 relative of the
oracle, FP32 within 1e-5, over the first steps of each configuration; the
relative error norms are
  e_pos = max_i |p_i - p_i*| / max(|p_i*|, L_vox)
  e_q   = max_i min(|q_i - q_i*|, |q_i + q_i*|).
"""
import numpy as np
import pytest

import oracle as O
import vx3mi
from vx3mi import configs

pytestmark = pytest.mark.gpu

TOL = {vx3mi.FP64: 1e-12, vx3mi.FP32: 1e-5}


def log(msg):
    print(msg, flush=True)


def pos_err(p, ref, L=0.01):
    d = np.linalg.norm(p - ref, axis=1)
    return float(np.max(d / np.maximum(np.linalg.norm(ref, axis=1), L))) if len(p) else 0.0


def quat_err(q, ref):
    return float(np.max(np.minimum(np.linalg.norm(q - ref, axis=1), np.linalg.norm(q + ref, axis=1)))) if len(q) else 0.0


def small_problem(seed=3, dims=(3, 3, 3), p=0.8, **kw):
    rng = np.random.default_rng(seed)
    prob = configs.random_robot(rng, dims, p, enable_collision=0)
    prob.update(kw)
    return prob


def random_pose(nv, rng, disp=2e-4, ang=0.02):
    pos_noise = rng.normal(scale=disp, size=(nv, 3))
    axis = rng.normal(size=(nv, 3))
    axis /= np.linalg.norm(axis, axis=1, keepdims=True)
    th = rng.uniform(0, ang, size=nv)
    q = np.concatenate([np.cos(th / 2)[:, None], axis * np.sin(th / 2)[:, None]], axis=1)
    return pos_noise, q


@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
@pytest.mark.parametrize("ang", [0.01, 0.3])
def test_link_forces_random_poses(precision, ang):
    """Per-link forces and moments for random relative poses, small- and large-angle branches."""
    prob = small_problem()
    ref = O.OracleModel(prob)
    batch = vx3mi.Batch([vx3mi.Model.from_problem(prob)], precision)
    rng = np.random.default_rng(11)
    # one step first so the voxels carry a previous dt (link damping divides by it)
    ref.step(1)
    batch.step(1)
    s0 = ref.state()
    noise, q = random_pose(ref.nv, rng, ang=ang)
    pos = s0["pos"] + noise
    for k in range(3):  # three evaluations: exercises the velocity-damping path and hysteresis
        ref.set_state(pos=pos, quat=q)
        batch.set_state(pos=pos, quat=q)
        fr = ref.update_links()
        fg = batch.eval_link_forces()
        scale = np.max(np.abs(fr))
        err = np.max(np.abs(fg - fr)) / scale
        tol = 1e-12 if precision == vx3mi.FP64 else 2e-5
        assert err <= tol, (k, err)
        noise, q = random_pose(ref.nv, rng, ang=ang)
        pos = s0["pos"] + noise


@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
def test_config0_trajectory(precision):
    """Config 0 (4x4x4 cube, thermal actuation, floor, friction): first 1000 steps."""
    prob = configs.config0()
    ref = O.OracleModel(prob)
    batch = vx3mi.Batch([vx3mi.Model.from_problem(prob)], precision)
    worst_p = worst_q = 0.0
    for k in range(10):
        batch.step(100, mode=vx3mi.MODE_FLAT)
        ref.step(100)
        s, r = batch.state(), ref.state()
        worst_p = max(worst_p, pos_err(s["pos"], r["pos"]))
        worst_q = max(worst_q, quat_err(s["quat"], r["quat"]))
    print(f"config0 FP{precision}: e_pos {worst_p:.3e} e_q {worst_q:.3e}")
    assert worst_p <= TOL[precision] and worst_q <= TOL[precision]
    assert abs(batch.time()[0] - ref.time) <= 1e-15


@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
def test_config3_per_voxel_modulus(precision):
    """Per-voxel E and phase (config 3 robot), collisions off: 500 steps."""
    prob = configs.config3()
    ref = O.OracleModel(prob)
    batch = vx3mi.Batch([vx3mi.Model.from_problem(prob)], precision)
    batch.step(500)
    ref.step(500)
    s, r = batch.state(), ref.state()
    ep, eq = pos_err(s["pos"], r["pos"]), quat_err(s["quat"], r["quat"])
    print(f"config3 FP{precision}: e_pos {ep:.3e} e_q {eq:.3e}")
    assert ep <= TOL[precision] and eq <= TOL[precision]


def test_modes_bitwise_identical():
    """Plain (per-robot launches), flat and persistent kernels give identical bits."""
    probs = configs.config2(seed=5, n=12)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    states = []
    for mode in (vx3mi.MODE_PLAIN, vx3mi.MODE_FLAT, vx3mi.MODE_PERSISTENT, vx3mi.MODE_TILED):
        b = vx3mi.Batch(models, vx3mi.FP32)
        b.step(130, mode=mode)
        states.append(b.state())
        b.close()
    for s in states[1:]:
        for k in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(s[k], states[0][k]), k


@pytest.mark.parametrize("mode", [vx3mi.MODE_PERSISTENT, vx3mi.MODE_TILED])
@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
def test_batched_robots_match_oracle(precision, mode):
    """Uneven robots in one batch, each checked against its own oracle run."""
    probs = configs.config2(seed=1, n=6)
    batch = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], precision)
    batch.step(150, mode=mode)
    batch.step(250, mode=mode)
    for r, p in enumerate(probs):
        ref = O.OracleModel(p)
        ref.step(400)
        s, o = batch.state(r), ref.state()
        ep, eq = pos_err(s["pos"], o["pos"]), quat_err(s["quat"], o["quat"])
        assert ep <= TOL[precision] and eq <= TOL[precision], (r, ep, eq)


def test_recommended_dt_reduction():
    """Device max-reduction equals the host evaluation and the oracle (exact)."""
    probs = configs.config2(seed=2, n=20) + [configs.config0(), configs.config3()]
    models = [vx3mi.Model.from_problem(p) for p in probs]
    b = vx3mi.Batch(models, vx3mi.FP32)
    dev = b.recommended_dt()
    host = np.array([m.recommended_dt() for m in models])
    orc = np.array([O.OracleModel(p).recommended_dt() for p in probs])
    assert np.array_equal(dev, host)
    assert np.array_equal(dev, orc)


def test_center_of_mass():
    probs = configs.config2(seed=4, n=5)
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    b.step(50)
    com = b.com()
    for r, p in enumerate(probs):
        ref = O.OracleModel(p)
        ref.step(50)
        assert np.allclose(com[r], ref.com(), rtol=1e-12, atol=1e-15)


def _perturbed(prob, seed=0, scale=1.6e-3):
    """Oracle model plus an initial state with voxels jostled by ~0.16 L so that
    many unlinked surface pairs start inside each other's envelope."""
    ref = O.OracleModel(prob)
    rng = np.random.default_rng(seed)
    s = ref.state()
    pos = s["pos"] + rng.normal(scale=scale, size=s["pos"].shape)
    ref.set_state(pos=pos)
    return ref, pos


@pytest.mark.parametrize("precision", [vx3mi.FP64, vx3mi.FP32])
def test_collisions_config1_block(precision):
    """Collisions on (config 1 block, 14x12x6): candidate lists + per-voxel contact
    gather against the oracle's pair loop, 300 steps from a jostled state."""
    prob = configs.config1(seed=2, dims=(14, 12, 6))
    ref, pos = _perturbed(prob)
    off = dict(prob, enable_collision=0)
    ref_off = O.OracleModel(off)
    ref_off.set_state(pos=pos)
    batch = vx3mi.Batch([vx3mi.Model.from_problem(prob)], precision)
    batch.set_state(pos=pos)
    for _ in range(3):
        batch.step(100, mode=vx3mi.MODE_FLAT)
        ref.step(100)
        ref_off.step(100)
    s, r = batch.state(), ref.state()
    ep, eq = pos_err(s["pos"], r["pos"]), quat_err(s["quat"], r["quat"])
    # the contacts matter: the run without them is far away
    assert pos_err(ref_off.state()["pos"], r["pos"]) > 1e-4
    print(f"collisions FP{precision}: e_pos {ep:.3e} e_q {eq:.3e}")
    assert ep <= TOL[precision] and eq <= TOL[precision]
    assert batch.diverged().sum() == 0


def _oracle_pairs(ref):
    o = np.sort(ref.contact_pairs(), axis=1)
    return o[np.lexsort((o[:, 1], o[:, 0]))] if len(o) else o.reshape(0, 2)


@pytest.mark.parametrize("mode", [vx3mi.MODE_FLAT, vx3mi.MODE_PLAIN])
def test_contact_pairs_equal_the_oracle(mode):
    """The pairs the collision phase finds in contact, step by step, are the
    oracle's all-pairs contacts (VX3_VoxelyzeKernel.cu:448-458), FP64, on the
    jostled config-1 block, while the candidate lists are rebuilt as the voxels
    move; the states stay bit-identical."""
    prob = configs.config1(seed=2, dims=(14, 12, 6))
    ref, pos = _perturbed(prob)
    batch = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    batch.set_state(pos=pos)
    batch.record_contacts()
    total = 0
    for k in range(80):
        batch.step(1, mode=mode)
        ref.step(1)
        g, o = batch.contact_pairs(0), _oracle_pairs(ref)
        assert np.array_equal(g, o), (k, len(g), len(o))
        total += len(g)
    s, r = batch.state(), ref.state()
    for key in ("pos", "quat", "linmom", "angmom"):
        assert np.array_equal(s[key], r[key]), key
    lst = int(batch.listings()[0])
    log(f"contact pairs over 80 steps: {total}, listings {lst}")
    assert total > 1000 and 1 < lst < 80  # rebuilt more than once, not every step


@pytest.mark.parametrize("mode", [vx3mi.MODE_FLAT, vx3mi.MODE_PLAIN])
def test_contact_pairs_large_body_several_count_workgroups(mode):
    """A 40x40x8 config-1 body (8,946 voxels, 8,222 surface voxels): its count
    pass runs in three workgroups, the last of which takes the prefix sum and the
    scatter, and the lists are built inside the contact kernel.  FP64 from a
    jostled state: the pairs in contact equal the oracle's all-pairs contacts at
    every step and the states stay bit-identical."""
    prob = configs.config1(seed=2, dims=(40, 40, 8))
    ref, pos = _perturbed(prob)
    model = vx3mi.Model.from_problem(prob)
    assert model.ns > 2 * 4096
    batch = vx3mi.Batch([model], vx3mi.FP64)
    batch.set_state(pos=pos)
    batch.record_contacts()
    total = 0
    for k in range(60):
        batch.step(1, mode=mode)
        ref.step(1)
        g, o = batch.contact_pairs(0), _oracle_pairs(ref)
        assert np.array_equal(g, o), (k, len(g), len(o))
        total += len(g)
    s, r = batch.state(), ref.state()
    for key in ("pos", "quat", "linmom", "angmom"):
        assert np.array_equal(s[key], r[key]), key
    lst = int(batch.listings()[0])
    log(f"large body: contact pairs over 60 steps {total}, listings {lst}")
    assert total > 1000 and 1 < lst < 60


def _squeezed_block(factor, dims=(6, 6, 5)):
    """A solid block squeezed to `factor` of its lattice spacing about its centre
    (no actuation, no gravity): at 0.45 every surface voxel has more candidates
    than a list holds, so the collision phase searches the grid directly."""
    prob = dict(configs.config0(), dims=dims, cells=np.ones(int(np.prod(dims)), np.int8), phase=np.zeros(int(np.prod(dims))),
                enable_collision=1, vary_temp=0, grav_acc=0.0, floor_enabled=0)
    ref = O.OracleModel(prob)
    pos = ref.state()["pos"]
    c = pos.mean(axis=0)
    pos = c + factor * (pos - c) + np.random.default_rng(5).normal(scale=2e-5, size=pos.shape)
    ref.set_state(pos=pos)
    return prob, ref, pos


def _listing_steps(prob, pos, steps):
    """The steps that start with a listing, replayed on the oracle's trajectory
    from pos: the first one, then each step after one that leaves a surface voxel
    more than 0.49 of the skin from where it was at the last listing (the skin is
    0.15 voxel lengths for robots of at most 1,280 voxels, 0.25 otherwise; the
    squared threshold is kept in float)."""
    ref = O.OracleModel(prob)
    ref.set_state(pos=pos)
    model = vx3mi.Model.from_problem(prob)
    topo = model.topology()
    nlinks = np.bincount(np.concatenate([topo["link_neg"], topo["link_pos"]]), minlength=model.nv)
    surface = nlinks < 6
    skin = 0.15 if model.nv <= 1280 else 0.25
    thr2 = float(np.float32((0.49 * skin * prob["lattice_dim"]) ** 2))
    base, out = pos[surface].copy(), [0]
    for k in range(1, steps):
        ref.step(1)
        p = ref.state()["pos"][surface]
        if (((p - base) ** 2).sum(axis=1) > thr2).any():
            base, out = p.copy(), out + [k]
    return out


@pytest.mark.parametrize("mode", [vx3mi.MODE_FLAT, vx3mi.MODE_PERSISTENT, vx3mi.MODE_TILED])
@pytest.mark.parametrize("factor", [0.45, 0.8])
def test_dense_pack_fp64_bit_exact(factor, mode):
    """Squeezed blocks: at 0.45 of the spacing the lists overflow and every step
    searches the rebuilt grid (listings == steps); at 0.8 the lists hold and are
    rebuilt as the staleness rule says.  Pairs and states equal the oracle's bit
    for bit."""
    prob, ref, pos = _squeezed_block(factor)
    models = [vx3mi.Model.from_problem(prob)] * 2  # two robots: the persistent kernel runs
    batch = vx3mi.Batch(models, vx3mi.FP64)
    batch.set_state(0, pos=pos)
    batch.set_state(1, pos=pos)
    batch.record_contacts()
    for k in range(6):
        batch.step(1, mode=mode)
        ref.step(1)
        o = _oracle_pairs(ref)
        assert len(o) > 0
        assert np.array_equal(batch.contact_pairs(0), o), k
        assert np.array_equal(batch.contact_pairs(1), o), k
    for r in (0, 1):
        s = batch.state(r)
        for key in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(s[key], ref.state()[key]), (r, key)
    lst = batch.listings()
    if factor < 0.5:
        assert (lst == 6).all(), lst
    else:  # the lists hold: a listing exactly when the staleness rule calls for one
        assert (lst == len(_listing_steps(prob, pos, 6))).all(), (lst, _listing_steps(prob, pos, 6))


@pytest.mark.parametrize("mode", [vx3mi.MODE_FLAT, vx3mi.MODE_PERSISTENT, vx3mi.MODE_TILED])
def test_divergence_stops_only_that_robot(mode):
    """doTimeStep returns false once a link's strain passes 100
    (VX3_VoxelyzeKernel.cu:271-280): at dt = 1e-5 s the robot of 1e9 / 1e10 Pa
    materials (stable below ~2e-7 s) blows up and stops; the two 1e6 / 1e7 Pa
    robots keep advancing.  Times and states equal the oracle's run of each
    robot alone, bit for bit (FP64); vx3mi_batch_diverged reports DIVERGED."""
    soft = small_problem(seed=3, dims=(4, 3, 3))
    stiff = dict(small_problem(seed=4, dims=(4, 3, 3)),
                 materials=[dict(configs.SOFT, E=1e9), dict(configs.STIFF, E=1e10)])
    probs = [soft, stiff, small_problem(seed=5, dims=(3, 4, 3))]
    batch = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    dt = 1e-5
    batch.step(3000, dt=dt, mode=mode)
    d = np.zeros(3, dtype=np.int32)
    rc = vx3mi.lib().vx3mi_batch_diverged(batch._h, d.ctypes.data_as(vx3mi.C.POINTER(vx3mi.C.c_int)))
    assert rc == vx3mi.DIVERGED and d.tolist() == [0, 1, 0]
    t = batch.time()
    for r, p in enumerate(probs):
        ref = O.OracleModel(p)
        done = ref.step(3000, dt)
        assert (done < 3000) == (r == 1), (r, done)
        assert t[r] == ref.time, (r, t[r], ref.time)
        s, o = batch.state(r), ref.state()
        for key in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(s[key], o[key]), (r, key)


def test_config1_fp32_first_1000_steps():
    """Full config 1 (64x64x16, 45,800 voxels, collisions on), FP32 on the device
    against the FP64 oracle over the first 1000 steps, the window of the 1e-5
    bar.  The reference's equations are piecewise: a link switches between its
    small- and large-angle formulations at 0.05 rad / 50 % with 1.2x hysteresis
    (VX3_Link.cu:117-131), ToRotationVector between its square-root and acos
    forms at 2.4e-3 (VX3_Quat3D.h:347-359), FromAngleToPosX between its
    first-order and exact forms (:384-432), and a contact's force jumps by its
    damping term when a pair enters the envelope (VX3_Collision.cu:17-31).
    Where FP32 takes such a branch a step apart from FP64 the two part; with the
    FP64 decisions forced the FP32 equations stay within 1e-6 of FP64 (the host
    twin, tests/test_fp32_attribution.py).  Every link whose decisions (link
    flags: small angle, the two rotation-vector forms, the FromAngleToPosX form)
    and every voxel whose contact pairs differ between device and oracle name
    their voxels, step by step.  Then:
      * e_pos <= 1e-5 on every voxel farther than one lattice cell from every
        voxel named so far (and on every voxel before the first naming);
      * e_q <= 1e-5 on every voxel outside the sound cones of the namings: a
        naming at step s covers the voxels within 1 + c (k - s) lattice cells at
        step k, c = sqrt(E_max / rho) dt / L = 0.143 cells per step, the distance
        an elastic wave of the stiffer material travels (a light voxel's
        orientation is excited by the wave from a mismatch some cells away)."""
    prob = configs.config1()
    model = vx3mi.Model.from_problem(prob)
    topo = model.topology()
    lat, ln, lp = topo["lattice"], topo["link_neg"], topo["link_pos"]
    ref = O.OracleModel(prob)
    batch = vx3mi.Batch([model], vx3mi.FP32)
    batch.record_contacts()
    dt = float(np.float32(0.9 * model.recommended_dt()))
    c = float(np.sqrt(max(m["E"] for m in prob["materials"]) / 1000.0) * dt / 0.01)
    first = np.full(model.nv, -1)    # step at which each voxel was first named
    near = np.zeros(model.nv, bool)  # named voxels and their 26 neighbours
    grid = -np.ones(lat.max(axis=0) + 3, np.int64)
    grid[lat[:, 0] + 1, lat[:, 1] + 1, lat[:, 2] + 1] = np.arange(model.nv)
    off = np.array([(i, j, k) for i in (-1, 0, 1) for j in (-1, 0, 1) for k in (-1, 0, 1)])
    worst_pos = worst_q = 0.0
    link_diff = pair_diff = over = 0
    for k in range(1000):
        batch.step(1, mode=vx3mi.MODE_FLAT)
        ref.step(1)
        gf, _ = batch.link_flags()
        _, of = ref.flags()
        bad = np.nonzero((gf ^ of) & 30)[0]  # small angle, rotation-vector and FromAngleToPosX forms
        gs, os_ = set(map(tuple, batch.contact_pairs(0))), set(map(tuple, _oracle_pairs(ref)))
        pv = np.array(sorted({v for pr in gs ^ os_ for v in pr}), dtype=np.int64)
        link_diff += len(bad)
        pair_diff += len(pv)
        new = np.concatenate([ln[bad], lp[bad], pv]).astype(np.int64)
        new = np.unique(new[first[new] < 0])
        if len(new):
            first[new] = k
            nb = grid[tuple((lat[new][:, None, :] + off[None, :, :] + 1).reshape(-1, 3).T)]
            near[nb[nb >= 0]] = True
        st, r = batch.state(), ref.state()
        d = np.linalg.norm(st["pos"] - r["pos"], axis=1) / np.maximum(np.linalg.norm(r["pos"], axis=1), 0.01)
        q = np.minimum(np.linalg.norm(st["quat"] - r["quat"], axis=1), np.linalg.norm(st["quat"] + r["quat"], axis=1))
        worst_pos = max(worst_pos, float(d[~near].max()))
        hot = np.nonzero(q > 1e-5)[0]
        over += len(hot)
        named = np.nonzero(first >= 0)[0]
        for v in hot:  # outside every sound cone?
            dist = np.abs(lat[named] - lat[v]).max(axis=1) if len(named) else np.array([])
            if not np.any(dist <= 1 + c * (k - first[named])):
                worst_q = max(worst_q, float(q[v]))
    log(f"config1 FP32 1000 steps: e_pos beyond one cell of a naming {worst_pos:.3e}, e_q outside the sound "
        f"cones {worst_q:.3e}; link decision differences {link_diff}, pair differences {pair_diff}, "
        f"voxels named {int((first >= 0).sum())} (first at step {int(first[first >= 0].min()) if (first >= 0).any() else -1}), "
        f"within a cell of one {near.mean():.3f}; voxel-steps with e_q > 1e-5: {over}")
    assert worst_pos <= 1e-5 and worst_q <= 1e-5
    assert near.mean() < 0.2  # the position criterion leaves most of the body named by nothing
    assert batch.diverged().sum() == 0


@pytest.mark.parametrize("precision", [vx3mi.FP32, vx3mi.FP64])
def test_flat_graph_chunks_agree(precision):
    """Flat steps run as 1000-step graphs, 64-step graphs or single launches
    give the same bits (collisions on, both step parities)."""
    probs = configs.config2(seed=5, n=6, collisions=1)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    a, b = vx3mi.Batch(models, precision), vx3mi.Batch(models, precision)
    a.step(1, mode=vx3mi.MODE_FLAT)
    b.step(1, mode=vx3mi.MODE_FLAT)
    a.step(2130, mode=vx3mi.MODE_FLAT)         # two 1000-step graphs, two 64-step graphs, 2 single steps
    for n in [63] + [64] * 32 + [19]:           # odd starts: single steps, 64-step graphs
        b.step(n, mode=vx3mi.MODE_FLAT)
    assert a.last_step_launches() > 0
    sa, sb = a.state(), b.state()
    for key in ("pos", "quat", "linmom", "angmom"):
        np.testing.assert_array_equal(sa[key], sb[key])
    assert not a.diverged().any()


@pytest.mark.parametrize("precision", [vx3mi.FP32, vx3mi.FP64])
def test_collision_modes_bitwise_identical(precision):
    """Collisions on: plain, flat, persistent and tiled kernels agree bit for bit
    (states, contact pairs of the last step, target hits), over 150 steps from a
    jostled state in which the candidate lists are rebuilt several times (FP64:
    robots of at most 7^3 cells, the FP64 collision tile holds 640 voxels)."""
    probs = configs.config2(seed=8, n=6, hi=12 if precision == vx3mi.FP32 else 7, collisions=1)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    rng = np.random.default_rng(1)
    out = []
    base = vx3mi.Batch(models, precision)
    pos = base.state()["pos"] + rng.normal(scale=1.5e-3, size=(base.nv, 3))
    base.close()
    for mode in (vx3mi.MODE_PLAIN, vx3mi.MODE_FLAT, vx3mi.MODE_PERSISTENT, vx3mi.MODE_TILED):
        b = vx3mi.Batch(models, precision)
        b.set_state(pos=pos)
        b.step(149, mode=mode)
        b.record_contacts()
        b.step(1, mode=mode)
        out.append((b.state(), b.contact_pairs(), b.listings()))
        assert b.diverged().sum() == 0
        b.close()
    log(f"listings per robot {out[0][2].tolist()}, pairs in the last step {len(out[0][1])}")
    assert out[0][2].min() > 1 and len(out[0][1]) > 0
    for st, pr, lst in out[1:]:
        for k in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(st[k], out[0][0][k]), k
        assert np.array_equal(pr, out[0][1])


def test_tiled_collisions_fp64_match_oracle():
    """Collisions on, FP64, tiled kernel (contacts from LDS poses and momenta,
    candidate lists rebuilt inside the launch): every robot of a jostled batch
    equals the oracle bit for bit after 200 steps (two launches of 100)."""
    probs = configs.config2(seed=9, n=5, hi=7, collisions=1)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    b = vx3mi.Batch(models, vx3mi.FP64)
    rng = np.random.default_rng(2)
    refs = []
    for r, p in enumerate(probs):
        ref = O.OracleModel(p)
        pos = ref.state()["pos"] + rng.normal(scale=1.5e-3, size=(ref.nv, 3))
        ref.set_state(pos=pos)
        b.set_state(r, pos=pos)
        refs.append(ref)
    for _ in range(2):
        b.step(100, mode=vx3mi.MODE_TILED)
    assert b.diverged().sum() == 0
    for r, ref in enumerate(refs):
        ref.step(200)
        s, o = b.state(r), ref.state()
        for k in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(s[k], o[k]), (r, k)


def test_tiled_full_batch_matches_flat():
    """The whole 1024-robot workload: tiled (LPT-packed tiles) and flat kernels agree bit for bit."""
    probs = configs.config2(seed=0, n=1024)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    out = []
    for mode in (vx3mi.MODE_FLAT, vx3mi.MODE_TILED):
        b = vx3mi.Batch(models, vx3mi.FP32)
        b.step(300, mode=mode)
        out.append((b.state(), b.time()))
        b.close()
    for k in ("pos", "quat", "linmom", "angmom"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])


def test_tiled_full_batch_with_collisions_matches_flat():
    """The whole 1024-robot workload with collisions on (the reference default):
    the collision tiles (packed candidate lists, listing inside the launch) and
    the flat kernels agree bit for bit over 600 steps in which the robots list
    their candidates again; so do the contact pairs of the last step."""
    probs = configs.config2(seed=0, n=1024, collisions=1)
    models = [vx3mi.Model.from_problem(p) for p in probs]
    out = []
    for mode in (vx3mi.MODE_FLAT, vx3mi.MODE_TILED):
        b = vx3mi.Batch(models, vx3mi.FP32)
        b.step(599, mode=mode)
        b.record_contacts()
        b.step(1, mode=mode)
        out.append((b.state(), b.time(), b.contact_pairs(), b.listings()))
        assert b.diverged().sum() == 0
        b.close()
    log(f"listings {int(out[0][3].sum())} (flat) {int(out[1][3].sum())} (tiled), pairs {len(out[0][2])}")
    assert out[0][3].sum() > 1024 and len(out[0][2]) > 0
    for k in ("pos", "quat", "linmom", "angmom"):
        assert np.array_equal(out[0][0][k], out[1][0][k]), k
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])


def _fixed_bottom():
    p = configs.config0()
    cells = np.ones(64, dtype=np.int8)
    cells[:16] = 2  # z = 0 layer
    return dict(p, cells=cells, materials=[configs.SOFT, dict(configs.SOFT, fixed=1)])


FEATURES = {
    # Poisson's ratio active (VolumeEffectsEnabled): eHat stress with the frozen transverse strain sum
    "volume_effects": lambda: dict(configs.config0(), volume_effects=1),
    # a fixed bottom layer: fixed voxels do not move, links between two fixed voxels are skipped
    "fixed_voxels": _fixed_bottom,
    # linear-with-failure material: links fail once their strain exceeds Fail_Stress / E
    # (random phases so that thermal strains differ between neighbours)
    "failure": lambda: dict(configs.config0(), phase=np.random.default_rng(4).uniform(0, 1, 64),
                            materials=[dict(configs.SOFT, fail_stress=2e4)]),
}


@pytest.mark.parametrize("name", list(FEATURES))
def test_material_features_fp64_bit_exact(name):
    """Material features outside the five workloads, FP64 against the oracle bit for bit."""
    log(f"feature {name}")
    prob = FEATURES[name]()
    b = vx3mi.Batch([vx3mi.Model.from_problem(prob)], vx3mi.FP64)
    o = O.OracleModel(prob)
    for _ in range(3):
        b.step(200)
        o.step(200)
        s, r = b.state(), o.state()
        assert np.array_equal(s["pos"], r["pos"]) and np.array_equal(s["quat"], r["quat"])
    if name == "failure":  # some links did fail (zero force after exceeding the failure strain)
        ref = O.OracleModel(dict(prob, materials=[configs.SOFT]))
        ref.step(600)
        assert np.abs(ref.state()["pos"] - o.state()["pos"]).max() > 1e-6


def _floor_cube_with_externals():
    """config 0 (floor, friction, actuation) with externals: a corner fully fixed
    and displaced, a top voxel pushed sideways with a moment, a voxel whose z
    rotation is fixed and one held at a prescribed y."""
    p = dict(configs.config0(), phase=np.random.default_rng(5).uniform(0, 1, 64))
    p["externals"] = [dict(voxel=0, dof=vx3mi.DOF_ALL, translation=(0.0, 0.0, 1e-4), rotation=(0.0, 0.0, 0.05)),
                      dict(voxel=63, force=(0.02, 0.0, 0.0), moment=(0.0, 1e-4, 0.0)),
                      dict(voxel=21, dof=vx3mi.DOF_TZ),
                      dict(voxel=42, dof=vx3mi.DOF_Y, translation=(0.0, -2e-4, 0.0))]
    return p


def _external_problems():
    """Voxel externals (two robots), a bilinear-material robot and a plain cube."""
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_golden import beam_with_externals, bilinear_robot
    return [beam_with_externals(), _floor_cube_with_externals(), bilinear_robot(), configs.config0()]


@pytest.mark.parametrize("mode,tile,threads", [(vx3mi.MODE_PLAIN, 0, 0), (vx3mi.MODE_FLAT, 0, 0),
                                               (vx3mi.MODE_PERSISTENT, 0, 256), (vx3mi.MODE_TILED, 0, 0),
                                               (vx3mi.MODE_TILED, 1, 0)])
def test_externals_fp64_bit_exact_every_mode(mode, tile, threads):
    """Voxel externals (Boundary_Conditions: fixed / displaced / rotated degrees
    of freedom, forces, moments) and a bilinear material (yield, unload,
    reload) in every launch mode, FP64 against the oracle bit for bit, in one
    batch with a plain robot."""
    log(f"externals mode {mode} tile {tile} threads {threads}")
    probs = _external_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    run_mode = mode
    if mode in (vx3mi.MODE_TILED, vx3mi.MODE_PERSISTENT):
        b.configure(mode, tile, threads)
        run_mode = vx3mi.MODE_AUTO
    oracles = [O.OracleModel(p) for p in probs]
    for _ in range(3):
        b.step(200, mode=run_mode)
        for r, o in enumerate(oracles):
            o.step(200)
            s, ref = b.state(r), o.state()
            assert np.array_equal(s["pos"], ref["pos"]) and np.array_equal(s["quat"], ref["quat"]), r
            assert np.array_equal(s["linmom"], ref["linmom"]) and np.array_equal(s["angmom"], ref["angmom"]), r
    b.close()
    # the externals act: the fully fixed voxel sits at its prescribed pose
    s = oracles[1].state()
    assert np.allclose(s["pos"][0], [0.0, 0.0, 1e-4], atol=0) and abs(s["quat"][0][3] - np.sin(0.025)) < 1e-15


def test_externals_fp32_tiled():
    """FP32 tiled kernel with externals and a bilinear material within 1e-5 of
    the oracle over 1000 steps."""
    probs = _external_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
    b.configure(vx3mi.MODE_TILED, 0, 0)
    oracles = [O.OracleModel(p) for p in probs]
    worst = 0.0
    for _ in range(10):
        b.step(100)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            worst = max(worst, pos_err(s["pos"], ref["pos"]), quat_err(s["quat"], ref["quat"]))
    b.close()
    log(f"  fp32 externals worst {worst:.3e}")
    assert worst <= 1e-5


# force fields: only operations whose results are the same bits on the device
# and in the oracle (arithmetic, comparisons, sqrt, and the shared sin / cos)
FORCE_FIELD = {
    "x": ("mtMUL", ("mtCONST", 0.002), ("mtSIN", ("mtMUL", ("mtCONST", 40.0), ("mtVAR", "t")))),
    "y": ("mtDIV", ("mtSUB", ("mtVAR", "x"), ("mtVAR", "y")), ("mtCONST", 50.0)),
    "z": ("mtMUL", ("mtGREATERTHAN", ("mtVAR", "z"), ("mtCONST", 0.02)),
          ("mtMUL", ("mtCONST", 0.001), ("mtCOS", ("mtSQRT", ("mtVAR", "num_voxel"))))),
}


def _force_field_problems():
    rng = np.random.default_rng(21)
    withcol = configs.random_robot(rng, (5, 4, 3), 0.8, enable_collision=1)
    return [dict(configs.config0(), force_field=FORCE_FIELD), configs.config0(),
            dict(withcol, force_field={"x": FORCE_FIELD["x"]})]


@pytest.mark.parametrize("mode", [vx3mi.MODE_PLAIN, vx3mi.MODE_FLAT, vx3mi.MODE_AUTO])
def test_force_fields_fp64_bit_exact(mode):
    """MathTree force fields (Simulator/ForceField) evaluated on the device from
    postfix programs: FP64 equals the oracle's FIFO evaluation bit for bit, in a
    batch with a robot without a field and one with collisions."""
    log(f"force fields mode {mode}")
    probs = _force_field_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    oracles = [O.OracleModel(p) for p in probs]
    for _ in range(3):
        b.step(150, mode=mode)
        for r, o in enumerate(oracles):
            o.step(150)
            s, ref = b.state(r), o.state()
            assert np.array_equal(s["pos"], ref["pos"]) and np.array_equal(s["quat"], ref["quat"]), r
    b.close()
    # the field acts: the same cube without it ends elsewhere
    assert np.abs(oracles[0].state()["pos"] - oracles[1].state()["pos"]).max() > 1e-6


def test_force_fields_fp32_and_persistent_refused():
    probs = _force_field_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
    with pytest.raises(vx3mi.VX3MIError):
        b.configure(vx3mi.MODE_PERSISTENT, 0, 256)
    oracles = [O.OracleModel(p) for p in probs]
    b.step(450)
    worst = 0.0
    for r, o in enumerate(oracles):
        o.step(450)
        s, ref = b.state(r), o.state()
        worst = max(worst, pos_err(s["pos"], ref["pos"]))
    b.close()
    log(f"  fp32 force fields worst e_pos {worst:.3e}")
    assert worst <= 1e-5


def _cilia_removal_problems():
    """Three robots for one batch: cilia and removal together (collisions on, a
    material removed mid-run, the other with cilia switched on after a delay),
    cilia only, and a plain robot."""
    rng = np.random.default_rng(21)
    a = configs.random_robot(rng, (6, 5, 4), 0.8, enable_collision=1)
    nv = int((np.asarray(a["cells"]) > 0).sum())
    a["materials"] = [dict(configs.SOFT, cilia=1.5, cilia_on_after=0.002), dict(configs.STIFF, remove_after=0.003)]
    a.update(enable_cilia=1, secondary_experiment=1, cilia_base=rng.normal(size=(nv, 3)) * 2e-3,
             cilia_shift=rng.normal(size=(nv, 3)) * 1e-3)
    b = configs.random_robot(rng, (4, 4, 3), 0.8, enable_collision=0)
    nb = int((np.asarray(b["cells"]) > 0).sum())
    b["materials"] = [dict(configs.SOFT, cilia=-0.7), dict(configs.STIFF, cilia=2.0)]
    b.update(enable_cilia=1, cilia_base=rng.normal(size=(nb, 3)) * 2e-3)
    c = configs.random_robot(rng, (4, 3, 3), 0.8, enable_collision=0)
    return [a, b, c]


@pytest.mark.parametrize("mode", [vx3mi.MODE_PLAIN, vx3mi.MODE_FLAT])
def test_cilia_and_removal_fp64_bit_exact(mode):
    """Cilia forces (surface voxels, turn-on time, rotated with the voxel) and
    SecondaryExperiment removal (a material removed mid-run with its links; the
    exposed voxels join the collision and cilia surface) in one batch with a
    plain robot: FP64 bit for bit with the oracle, removal steps equal."""
    log(f"cilia/removal mode {mode}")
    probs = _cilia_removal_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    oracles = [O.OracleModel(p) for p in probs]
    for _ in range(6):
        b.step(100, mode=mode)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            for k in ("pos", "quat", "linmom", "angmom"):
                assert np.array_equal(s[k], ref[k]), (r, k)
            vr, lr = b.removed(r)
            ovr, olr = o.removed()
            assert np.array_equal(vr, ovr) and np.array_equal(lr, olr), r
    vr, _ = b.removed(0)
    assert (vr > 0).any() and (vr == 0).any()  # the removal happened, and not to everything
    assert not b.removed(1)[0].any() and not b.removed(2)[0].any()
    with pytest.raises(vx3mi.VX3MIError):
        b.step(1, mode=vx3mi.MODE_PERSISTENT)  # removal needs a kernel between steps
    b.close()


def test_cilia_and_removal_fp32():
    """The same batch in FP32 (flat mode) within 1e-5 of the oracle over 600 steps."""
    probs = _cilia_removal_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
    oracles = [O.OracleModel(p) for p in probs]
    worst = 0.0
    for _ in range(6):
        b.step(100)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            worst = max(worst, pos_err(s["pos"], ref["pos"]), quat_err(s["quat"], ref["quat"]))
            assert np.array_equal(b.removed(r)[0], o.removed()[0]), r
    b.close()
    log(f"  fp32 cilia/removal worst {worst:.3e}")
    assert worst <= TOL[vx3mi.FP32]


def test_cilia_in_persistent_mode_fp64():
    """Cilia without removal run in the persistent kernel too, bit for bit."""
    probs = _cilia_removal_problems()[1:]
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    b.configure(vx3mi.MODE_PERSISTENT, 0, 256)
    oracles = [O.OracleModel(p) for p in probs]
    b.step(400)
    for r, o in enumerate(oracles):
        o.step(400)
        s, ref = b.state(r), o.state()
        assert np.array_equal(s["pos"], ref["pos"]) and np.array_equal(s["angmom"], ref["angmom"]), r
    b.close()


def _data_model_problems():
    import os
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_golden import golden_problems
    a = golden_problems()["data_model"]
    b = dict(a, materials=[dict(m, ss_data=(np.cumsum(np.linspace(1e-4, 2e-4, 32)), np.cumsum(np.linspace(200.0, 20.0, 32))))
                           for m in a["materials"]])
    return [a, b, small_problem(seed=5)]


@pytest.mark.parametrize("mode,tile,threads", [(vx3mi.MODE_PLAIN, 0, 0), (vx3mi.MODE_FLAT, 0, 0),
                                               (vx3mi.MODE_PERSISTENT, 0, 256), (vx3mi.MODE_TILED, 0, 0)])
def test_data_materials_fp64_bit_exact_every_mode(mode, tile, threads):
    """MatModel 3 stress-strain data (links running along 4- to 63-point series
    curves, failing past the last point) in every launch mode: FP64 bit for bit
    with the oracle, which the data-model golden pins to the reference CPU path."""
    log(f"data materials mode {mode} tile {tile} threads {threads}")
    probs = _data_model_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    run_mode = mode
    if mode in (vx3mi.MODE_TILED, vx3mi.MODE_PERSISTENT):
        b.configure(mode, tile, threads)
        run_mode = vx3mi.MODE_AUTO
    oracles = [O.OracleModel(p) for p in probs]
    for _ in range(3):
        b.step(200, mode=run_mode)
        for r, o in enumerate(oracles):
            o.step(200)
            s, ref = b.state(r), o.state()
            for k in ("pos", "quat", "linmom", "angmom"):
                assert np.array_equal(s[k], ref[k]), (r, k)
    b.close()


def test_data_materials_fp32_tiled():
    """The same robots in FP32 (tiled) within 1e-5 of the oracle over 600 steps,
    actuated so that links run along several segments of their curves without
    reaching the failure point (failure is a threshold that FP32 crosses at its
    own step)."""
    a, b, c = _data_model_problems()
    probs = [dict(a, temp_amplitude=8.0), dict(b, temp_amplitude=2.0), c]
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
    b.configure(vx3mi.MODE_TILED, 0, 0)
    oracles = [O.OracleModel(p) for p in probs]
    worst = 0.0
    for _ in range(6):
        b.step(100)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            worst = max(worst, pos_err(s["pos"], ref["pos"]), quat_err(s["quat"], ref["quat"]))
    b.close()
    log(f"  fp32 data materials worst {worst:.3e}")
    assert worst <= TOL[vx3mi.FP32]


def _signal_problems():
    """A robot with signals driving cilia (pacemaker material, ShiftCiliaForce on
    every voxel, collisions on) and the same robot with removal as well, plus
    a plain robot."""
    rng = np.random.default_rng(31)
    a = configs.random_robot(rng, (6, 5, 3), 0.85, enable_collision=1)
    nv = int((np.asarray(a["cells"]) > 0).sum())
    sig = dict(signal_decay=0.85, signal_delay=0.0015, inactive_period=0.002)
    a["materials"] = [dict(configs.SOFT, cilia=1.0, **sig),
                      dict(configs.STIFF, cilia=1.0, is_pacemaker=1, pacemaker_period=0.004, **sig)]
    a.update(enable_signals=1, enable_cilia=1, cilia_base=rng.normal(size=(nv, 3)) * 1e-4,
             cilia_shift=rng.normal(size=(nv, 3)) * 2e-5)
    b = dict(a, secondary_experiment=1,
             materials=[a["materials"][0], dict(a["materials"][1], remove_after=0.005)])
    return [a, b, small_problem(seed=8)]


@pytest.mark.parametrize("mode", [vx3mi.MODE_PLAIN, vx3mi.MODE_FLAT])
def test_signals_fp64_bit_exact(mode):
    """Signals (pacemakers, propagation with delay / inactive period / value decay,
    local-signal decay) modulating cilia forces, with collisions and removal:
    FP64 state and every voxel's local signal bit for bit with the oracle."""
    log(f"signals mode {mode}")
    probs = _signal_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    oracles = [O.OracleModel(p) for p in probs]
    seen = 0
    for _ in range(8):
        b.step(100, mode=mode)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            for k in ("pos", "quat", "linmom", "angmom"):
                assert np.array_equal(s[k], ref[k]), (r, k)
            assert np.array_equal(b.signals(r), o.signals()), r
        seen = max(seen, int((b.signals(0) > 0).sum()))
    assert seen > 10  # the pulses spread
    b.close()


def test_signals_fp32():
    """The same batch in FP32 within 1e-5 of the oracle over 800 steps; the local
    signals (a function of time alone) equal the oracle's."""
    probs = _signal_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
    oracles = [O.OracleModel(p) for p in probs]
    worst = 0.0
    for _ in range(8):
        b.step(100)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            worst = max(worst, pos_err(s["pos"], ref["pos"]), quat_err(s["quat"], ref["quat"]))
            assert np.array_equal(b.signals(r), o.signals()), r
    b.close()
    log(f"  fp32 signals worst {worst:.3e}")
    assert worst <= TOL[vx3mi.FP32]


def _detach_problems():
    """A robot whose first material fails at a strain of 0.01 under 39 degC
    actuation (random phases), with EnableDetach and collisions on, signals and
    cilia as well; and the same robot with detach off (failed links stay, with
    zero force)."""
    rng = np.random.default_rng(41)
    a = configs.random_robot(rng, (5, 4, 3), 0.9, enable_collision=1)
    nv = int((np.asarray(a["cells"]) > 0).sum())
    a["materials"] = [dict(configs.SOFT, fail_stress=1e4, cilia=1.0),
                      dict(configs.STIFF, is_pacemaker=1, pacemaker_period=0.003, signal_delay=0.001,
                           inactive_period=0.001)]
    a.update(enable_detach=1, enable_signals=1, enable_cilia=1, cilia_shift=rng.normal(size=(nv, 3)) * 1e-5)
    return [a, dict(a, enable_detach=0)]


@pytest.mark.parametrize("mode", [vx3mi.MODE_PLAIN, vx3mi.MODE_FLAT])
def test_detach_fp64_bit_exact(mode):
    """EnableDetach: links past their failure strain leave both voxels (the
    surface, cilia, signals and collisions follow): FP64 state and detach steps
    bit for bit with the oracle.  (FP32 is not compared here: failure is a
    threshold that FP32 crosses at its own step.)"""
    log(f"detach mode {mode}")
    probs = _detach_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    oracles = [O.OracleModel(p) for p in probs]
    for _ in range(6):
        b.step(100, mode=mode)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            for k in ("pos", "quat", "linmom", "angmom"):
                assert np.array_equal(s[k], ref[k]), (r, k)
            assert np.array_equal(b.removed(r)[1], o.removed()[1]), r
            assert np.array_equal(b.signals(r), o.signals()), r
    lr = b.removed(0)[1]
    assert (lr > 0).sum() >= 3 and not b.removed(1)[1].any()
    b.close()


def _target_problems():
    """A non-target voxel pushed into a linked pair of target voxels (collisions
    and signals on), and a random robot whose stiff material is a target."""
    a = dict(configs.BASE, dims=(4, 1, 1), cells=np.array([1, 0, 2, 2], dtype=np.int8), phase=None,
             enable_collision=1, enable_signals=1, vary_temp=0, floor_enabled=0, grav_acc=0.0,
             materials=[dict(configs.SOFT), dict(configs.SOFT, is_target=1)],
             externals=[dict(voxel=0, force=(0.2, 0, 0)), dict(voxel=1, force=(-0.2, 0, 0))], slow_damp_z=0.0)
    rng = np.random.default_rng(51)
    b = configs.random_robot(rng, (6, 6, 4), 0.7, enable_collision=1)
    b["materials"] = [b["materials"][0], dict(b["materials"][1], is_target=1)]
    return [a, b]


@pytest.mark.parametrize("mode", [vx3mi.MODE_PLAIN, vx3mi.MODE_FLAT])
def test_targets_fp64_bit_exact(mode):
    """Target contacts: collisionCount (each target / non-target pair in the
    contact envelope, once per step) and the forced signal of the non-target
    voxel, FP64 bit for bit with the oracle."""
    probs = _target_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    oracles = [O.OracleModel(p) for p in probs]
    for _ in range(4):
        b.step(100, mode=mode)
        for r, o in enumerate(oracles):
            o.step(100)
            s, ref = b.state(r), o.state()
            for k in ("pos", "quat", "linmom", "angmom"):
                assert np.array_equal(s[k], ref[k]), (r, k)
            assert np.array_equal(b.signals(r), o.signals()), r
        assert b.hits().tolist() == [o.collision_count() for o in oracles]
    assert b.hits()[0] > 0
    b.close()


def _ragged_problems():
    """Robots of very different sizes in one batch: one voxel (no links), two
    voxels (one link), a 1 x 1 x 40 bar standing on the floor, a 3 x 3 x 3
    cube and a random robot."""
    base = dict(configs.config0(), phase=None)
    one = dict(base, dims=(1, 1, 1), cells=np.ones(1, dtype=np.int8))
    two = dict(base, dims=(2, 1, 1), cells=np.ones(2, dtype=np.int8), phase=np.array([0.0, 0.5]))
    bar = dict(base, dims=(1, 1, 40), cells=np.ones(40, dtype=np.int8), phase=np.linspace(0, 1, 40))
    cube = dict(base, dims=(3, 3, 3), cells=np.ones(27, dtype=np.int8), phase=np.linspace(0, 1, 27))
    return [one, two, bar, cube, small_problem(seed=9)]


@pytest.mark.parametrize("mode,tile,threads", [(vx3mi.MODE_PLAIN, 0, 0), (vx3mi.MODE_FLAT, 0, 0),
                                               (vx3mi.MODE_PERSISTENT, 0, 128), (vx3mi.MODE_TILED, 0, 0),
                                               (vx3mi.MODE_TILED, 1, 0)])
def test_ragged_batch_every_mode_fp64(mode, tile, threads):
    """Ragged batches (a voxel without links, a single link, a 40-voxel chain)
    give the oracle's bits in every launch mode; the robots' clocks advance
    independently (each its own recommended dt)."""
    log(f"ragged mode {mode} tile {tile}")
    probs = _ragged_problems()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    run_mode = mode
    if mode in (vx3mi.MODE_TILED, vx3mi.MODE_PERSISTENT):
        b.configure(mode, tile, threads)
        run_mode = vx3mi.MODE_AUTO
    oracles = [O.OracleModel(p) for p in probs]
    b.step(500, mode=run_mode)
    for r, o in enumerate(oracles):
        o.step(500)
        s, ref = b.state(r), o.state()
        for k in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(s[k], ref[k]), (r, k)
    assert np.array_equal(b.time(), np.array([o.time for o in oracles]))
    b.close()


def test_robot_beyond_tile_capacity_runs_elsewhere():
    """A robot larger than the large tile (2,352 voxels > 1,536) makes the tiled
    mode unavailable (configure raises) and the automatic choice runs it in
    another mode, still bit for bit in FP64."""
    big = dict(configs.config0(), dims=(14, 14, 12), cells=np.ones(14 * 14 * 12, dtype=np.int8), phase=None)
    probs = [big, small_problem(seed=10)]
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP64)
    with pytest.raises(vx3mi.VX3MIError):
        b.configure(vx3mi.MODE_TILED, 0, 0)
    b.step(100)
    for r, p in enumerate(probs):
        o = O.OracleModel(p)
        o.step(100)
        assert np.array_equal(b.state(r)["pos"], o.state()["pos"]), r
    b.close()
