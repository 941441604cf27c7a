"""Pins the FP64 oracle: golden trajectories of the reference's own CPU Voxelyze,
closed forms and invariants of the update equations.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign or axis, a transposed operand) fails it:
  * golden vectors: every term of the link and voxel update on real robots;
  * recommendedTimeStep: the closed form 1/(2 pi sqrt(E L / m));
  * free fall: the exact symplectic-Euler sequence z_n = z_0 - a dt^2 n(n+1)/2;
  * floor: rest penetration m g / (2 E L);
  * two-voxel axial oscillation: period 2 pi / sqrt(2 E L / m);
  * uniform temperature: equilibrium spacing L (1 + alpha T);
  * internal forces conserve linear momentum;
  * link forces in voxel frames are objective under a rigid motion;
  * two overlapping unlinked voxels: penalty force k (R - d) (k = 2 E L);
  * cilia: a turned free voxel follows the rotated force exactly; a cube gains
    momentum from its 26 surface voxels only, from the turn-on time on;
  * removal: the step it happens, frozen removed voxels, free remainder, and
    the regenerated surface (the exposed former centre gets cilia force);
  * signals: a pacemaker pulse runs down a bar at one signalTimeDelay per hop
    with value 100 rho^k, never echoes; cilia force scales with the signal.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from vx3mi import configs

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")


def quiet(dims, cells=None, **kw):
    """A robot with no gravity, floor, damping or thermal actuation unless asked."""
    n = int(np.prod(dims))
    p = dict(configs.BASE, dims=dims, cells=np.ones(n, dtype=np.int8) if cells is None else np.asarray(cells, np.int8),
             phase=None, materials=[dict(E=1e6, rho=1000.0, cte=0.01, mu_s=1.0, mu_k=0.5)],
             grav_acc=0.0, floor_enabled=0, bond_damp_z=0.0, slow_damp_z=0.0, col_damp_z=0.0,
             vary_temp=0, temp_amplitude=0.0, enable_collision=0)
    p.update(kw)
    return p


@pytest.mark.parametrize("name", ["config0", "robot5x4x3", "beam_bc", "bilinear", "data_model"])
def test_golden_reference_cpu(name):
    """Oracle (CVoxelyze integrator semantics) vs the reference CPU Voxelyze built
    from src/old: positions and orientations every 50 steps for 600 steps."""
    g = np.load(os.path.join(GOLDEN, f"ref_cpu_{name}.npz"))
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import golden_problems

    prob = golden_problems()[name]
    m = O.OracleModel(prob)
    every = int(g["every"])
    for k in range(g["pos"].shape[0]):
        m.step(every, float(g["dt"]), O.VXO_CPU_SEMANTICS)
        s = m.state()
        e_pos = np.max(np.linalg.norm(s["pos"] - g["pos"][k], axis=1) / np.maximum(np.linalg.norm(g["pos"][k], axis=1), 0.01))
        e_q = np.max(np.abs(s["quat"] - g["quat"][k]))
        # ~1e-16 through step 500; a floor-friction switch near step 550 then
        # amplifies the last-ulp difference between glibc's sin and the oracle's
        assert e_pos <= 1e-11 and e_q <= 1e-11, (k, e_pos, e_q)


def test_bilinear_golden_exercises_plasticity():
    """The bilinear fixture's links go past the yield strain: the same robot with
    a linear first material ends elsewhere."""
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import golden_problems

    prob = golden_problems()["bilinear"]
    g = np.load(os.path.join(GOLDEN, "ref_cpu_bilinear.npz"))
    lin = dict(prob, materials=[dict(prob["materials"][0], yield_stress=0.0), prob["materials"][1]])
    m = O.OracleModel(lin)
    m.step(600, float(g["dt"]), O.VXO_CPU_SEMANTICS)
    assert np.max(np.abs(m.state()["pos"] - g["pos"][-1])) > 1e-5


def test_data_model_golden_runs_along_the_curves():
    """The data-model fixture's links pass the first data point: replacing the
    curves by their first segments (linear materials of the same modulus) ends
    elsewhere; its link curves are the series combinations of the data."""
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import golden_problems

    prob = golden_problems()["data_model"]
    g = np.load(os.path.join(GOLDEN, "ref_cpu_data_model.npz"))
    lin = dict(prob, materials=[dict(m, ss_data=None, E=m["ss_data"][1][1] / m["ss_data"][0][1]) for m in prob["materials"]])
    m = O.OracleModel(lin)
    m.step(600, float(g["dt"]), O.VXO_CPU_SEMANTICS)
    assert np.max(np.abs(m.state()["pos"] - g["pos"][-1])) > 1e-5
    d = O.OracleModel(prob)
    curves = {tuple(map(tuple, d.link_curve(l))) for l in range(d.nl)}
    assert len(curves) == 3  # A-A, A-B, B-B
    for c in curves:  # strains rise from (0, 0); stresses follow the softer material
        c = np.array(c)
        assert c[0].tolist() == [0.0, 0.0] and np.all(np.diff(c[:, 0]) > 0) and np.all(np.diff(c[:, 1]) > 0)


@pytest.mark.skipif(not os.path.exists(os.path.join(ROOT, "oracle", "_ref", "vx3_ref_driver")),
                    reason="reference CPU driver not built (make -C oracle ref needs the reference tree)")
def test_live_reference_cpu_config1_subblock():
    """A 10x10x6 two-material block of config 1 (collisions off) vs the live reference build."""
    import sys
    sys.path.insert(0, GOLDEN)
    from make_golden import run_reference

    prob = configs.config1(seed=3, dims=(10, 10, 6))
    prob["enable_collision"] = 0
    dt, rec = run_reference(prob, 300, 100)
    m = O.OracleModel(prob)
    for k in range(rec.shape[0]):
        m.step(100, float(dt), O.VXO_CPU_SEMANTICS)
        e = np.max(np.abs(m.state()["pos"] - rec[k, :, 0:3])) / 0.01
        assert e <= 1e-11, (k, e)


def test_recommended_dt_closed_form():
    m = O.OracleModel(configs.config0())
    E, L, rho = np.float32(1e6), np.float32(0.01), 1000.0
    mass = np.float32(0.01 ** 3 * rho)
    k = np.float32(E * L)
    expect = 1.0 / (np.float32(6.283185) * np.sqrt(float(k) / float(mass)))
    assert m.recommended_dt() == pytest.approx(expect, rel=1e-15)
    assert m.recommended_dt() == pytest.approx(1.0 / (2 * np.pi * np.sqrt(1e6 * 0.01 / 1e-3)), rel=1e-6)
    # two materials: the stiffest link (harmonic mean of 1e7 and 1e7) decides
    m1 = O.OracleModel(configs.config1(seed=0, dims=(8, 8, 4)))
    assert m1.recommended_dt() == pytest.approx(1.0 / (2 * np.pi * np.sqrt(1e7 * 0.01 / 1e-3)), rel=1e-6)


def test_free_fall_exact_sequence():
    p = quiet((1, 1, 1), grav_acc=-9.81)
    m = O.OracleModel(p)
    dt = np.float32(m.step_dt())
    mass = np.float32(1e-3)
    grav_mult = np.float32(-np.float32(-9.81) / 9.80665)
    gforce = float(np.float32(np.float32(-mass * np.float32(9.80665)) * grav_mult))
    mass_inv = float(np.float32(1.0) / mass)  # _massInverse is stored as float
    a = gforce * mass_inv
    z0 = m.state()["pos"][0, 2]
    for n in (1, 10, 1000):
        m2 = O.OracleModel(p)
        m2.step(n, float(dt))
        z = m2.state()["pos"][0, 2]
        expect = z0 + a * float(dt) ** 2 * n * (n + 1) / 2
        assert z == pytest.approx(expect, rel=1e-12, abs=1e-18)


def test_floor_rest_penetration():
    p = quiet((1, 1, 1), grav_acc=-9.81, floor_enabled=1, slow_damp_z=1.0, col_damp_z=0.8)
    m = O.OracleModel(p)
    m.step(3000)
    z = m.state()["pos"][0, 2]
    mass = float(np.float32(1e-3))
    grav = float(np.float32(-9.81)) * -1.0
    expect = mass * 9.80665 * float(np.float32(grav / 9.80665)) / (2 * 1e6 * 0.01)
    assert -z == pytest.approx(expect, rel=1e-5)


def test_two_voxel_axial_oscillation_period():
    p = quiet((2, 1, 1))
    m = O.OracleModel(p)
    s = m.state()
    pos = s["pos"].copy()
    pos[1, 0] += 1e-4
    m.set_state(pos=pos)
    dt = m.step_dt()
    ext = []
    for _ in range(400):
        m.step(1)
        q = m.state()["pos"]
        ext.append(q[1, 0] - q[0, 0] - 0.01)
    ext = np.array(ext)
    crossings = np.where(np.diff(np.sign(ext)) != 0)[0]
    period = 2 * np.mean(np.diff(crossings)) * dt
    k, mass = 1e6 * 0.01, 1e-3
    omega = np.sqrt(2 * k / mass)
    omega_d = 2 / dt * np.arcsin(omega * dt / 2)  # symplectic Euler dispersion
    assert period == pytest.approx(2 * np.pi / omega_d, rel=2e-2)
    assert period == pytest.approx(2 * np.pi / omega, rel=3e-2)
    # no damping: the amplitude neither grows nor decays
    assert np.max(np.abs(ext[-100:])) == pytest.approx(np.max(np.abs(ext[:100])), rel=2e-2)


def test_uniform_temperature_equilibrium_spacing():
    p = quiet((3, 3, 3), temp_amplitude=-20.0, vary_temp=0, slow_damp_z=1.0, bond_damp_z=0.1)
    m = O.OracleModel(p)
    m.step(4000)
    pos = m.state()["pos"].reshape(3, 3, 3, 3)  # z, y, x, xyz
    d = np.linalg.norm(pos[:, :, 1:, :] - pos[:, :, :-1, :], axis=-1)
    assert np.allclose(d, 0.01 * (1 + 0.01 * -20.0), rtol=1e-6)


def test_linear_momentum_conservation():
    rng = np.random.default_rng(5)
    p = configs.random_robot(rng, (4, 4, 3), 0.8, enable_collision=0)
    p.update(grav_acc=0.0, floor_enabled=0, slow_damp_z=0.0)
    p["materials"] = [dict(mm, rho=1000.0) for mm in p["materials"]]
    m = O.OracleModel(p)
    m.step(600)
    lm = m.state()["linmom"]
    assert np.linalg.norm(lm.sum(axis=0)) <= 1e-11 * np.abs(lm).sum()
    assert np.abs(lm).sum() > 0


def _quat_mul(a, b):
    w1, x1, y1, z1 = a
    w2, x2, y2, z2 = b
    return np.array([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2, w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                     w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2, w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2])


def test_link_forces_objective_under_rigid_motion():
    rng = np.random.default_rng(9)
    p = configs.random_robot(rng, (4, 3, 3), 0.9, enable_collision=0)
    a = O.OracleModel(p)
    s = a.state()
    pos = s["pos"] + rng.normal(scale=3e-4, size=s["pos"].shape)
    ang = rng.normal(scale=0.03, size=(a.nv, 3))
    th = np.linalg.norm(ang, axis=1)
    q = np.concatenate([np.cos(th / 2)[:, None], (ang / th[:, None]) * np.sin(th / 2)[:, None]], axis=1)
    a.set_state(pos=pos, quat=q)
    fa = a.update_links()
    # rigid motion: rotation R (quaternion r) and translation t
    r = np.array([np.cos(0.35), *(np.sin(0.35) * np.array([0.3, -0.5, 0.81]) / np.linalg.norm([0.3, -0.5, 0.81]))])
    w, x, y, z = r
    R = np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                  [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                  [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])
    b = O.OracleModel(p)
    b.set_state(pos=pos @ R.T + np.array([0.03, -0.2, 0.5]), quat=np.array([_quat_mul(r, qq) for qq in q]))
    fb = b.update_links()
    assert np.max(np.abs(fa - fb)) <= 1e-9 * np.max(np.abs(fa))


def test_two_sphere_contact_force():
    # two unlinked voxels (lattice x = 0 and x = 2), the second moved to x = 1.1 L
    p = quiet((3, 1, 1), cells=[1, 0, 1], enable_collision=1)
    m = O.OracleModel(p)
    pos = m.state()["pos"]
    pos[1, 0] = 0.011
    m.set_state(pos=pos)
    dt = m.step_dt()
    m.step(1)
    lm = m.state()["linmom"]
    pk = np.float32(2 * np.float32(1e6) * 0.01)  # penetrationStiffness, float
    k = float(np.float32(2.0) / (np.float32(1.0) / pk + np.float32(1.0) / pk))  # VX3_Collision.cu:6
    force = k * (1.25 * 0.01 - 0.011)
    assert lm[0, 0] == pytest.approx(-force * dt, rel=1e-12)
    assert lm[1, 0] == pytest.approx(force * dt, rel=1e-12)
    # a linked pair pushed to the same distance feels only its link, never a contact
    q = quiet((2, 1, 1), enable_collision=1)
    a, b = O.OracleModel(q), O.OracleModel(dict(q, enable_collision=0))
    for mm in (a, b):
        ps = mm.state()["pos"]
        ps[1, 0] = 0.011
        mm.set_state(pos=ps)
        mm.step(1)
    assert np.array_equal(a.state()["linmom"], b.state()["linmom"])


def test_volume_effects_keep_the_transverse_info_of_creation():
    """VX3_Link::updateForces leaves updateTransverseInfo commented out
    (VX3_Link.cu:145-149): with volume effects the transverse strain sum and
    area stay those of the link's creation (zero strain: 0 and L^2).  A bar of
    two voxels (Poisson 0.3) stretched by 2 % then has the axial force
    eHat (1 - nu) strain L^2; with the transverse strains of the stretch
    (-nu strain per voxel and axis, summed over two axes) the stress would be
    lower by eHat nu * 2 nu strain, 26 % of it."""
    nu, E, L = 0.3, 1e6, 0.01
    p = quiet((2, 1, 1), volume_effects=1, materials=[dict(E=E, rho=1000.0, cte=0.0, nu=nu)])
    m = O.OracleModel(p)
    pos = m.state()["pos"]
    pos[1, 0] = L * 1.02
    m.set_state(pos=pos)
    f = m.update_links()[0]
    strain = np.float32(np.float32(0.02 * L) / np.float32(L))
    e_hat = np.float32(E) / ((np.float32(1) - np.float32(2) * np.float32(nu)) * (np.float32(1) + np.float32(nu)))
    frozen = float(e_hat * (np.float32(1) - np.float32(nu)) * strain) * (L * L)
    updated = float(e_hat * ((np.float32(1) - np.float32(nu)) * strain + np.float32(nu) * (-2 * nu * strain))) * (L * L)
    assert f[0] == pytest.approx(frozen, rel=1e-5)  # forceNeg.x pulls the negative end along +x
    assert abs(f[0] - updated) > 0.2 * abs(frozen)
    assert f[3] == -f[0]


def test_cpu_semantics_flag_only_changes_static_friction_rotation():
    """VX3 zeroes angular momentum of voxels held by static floor friction
    (VX3_Voxel.cu:259-264); CVoxelyze does not.  Without a floor they agree."""
    p = dict(configs.config0(), floor_enabled=0)
    a, b = O.OracleModel(p), O.OracleModel(p)
    a.step(300)
    b.step(300, flags=O.VXO_CPU_SEMANTICS)
    # only the float vs double dt*massInverse product differs without a floor
    assert np.max(np.abs(a.state()["pos"] - b.state()["pos"])) <= 1e-12


def test_static_friction_zeroes_angular_momentum():
    """VX3_Voxel.cu:259-264: a voxel held by static floor friction ends every step
    with zero angular momentum.  A 3x3x1 slab resting on the floor, a twisting and
    tipping moment on its centre voxel: under VX3 semantics every voxel with the
    static-friction flag and a non-negative floor penetration has angMom exactly
    0 after each step, so the twisted voxel turns only by the one step of
    rotation each step's moment gives it; under CVoxelyze semantics (no reset)
    the same voxel carries angular momentum and turns faster until the links
    balance the moment."""
    p = quiet((3, 3, 1), grav_acc=-9.81, floor_enabled=1, bond_damp_z=0.1, slow_damp_z=1.0,
              materials=[dict(E=1e6, rho=1000.0, cte=0.0, mu_s=2.0, mu_k=0.5)])
    vx3, cpu = O.OracleModel(p), O.OracleModel(p)
    for m in (vx3, cpu):
        m.set_external(4, dof=0, moment=(2e-6, 0.0, 5e-6))
    held = 0
    for k in range(200):
        vx3.step(1)
        cpu.step(1, flags=O.VXO_CPU_SEMANTICS)
        s, c = vx3.state(), cpu.state()
        vf, _ = vx3.flags()
        static = ((vf & 8) != 0) & (s["pos"][:, 2] <= 0.0)  # FLOOR_STATIC_FRICTION, floorPenetration >= 0
        held += int(static.sum())
        assert np.all(s["angmom"][static] == 0.0), k
        if k < 10:  # spinning up: CVoxelyze keeps the angular momentum, VX3 drops it
            assert static[4] and np.linalg.norm(c["angmom"][4]) > 1e-11, k
            assert k == 0 or c["quat"][4][3] > s["quat"][4][3] * 1.05, k
    assert held >= 9 * 150  # the slab rests in static friction for most of the run


def test_contacts_grid_equals_all_pairs():
    """The oracle's grid-pruned contact search against the literal all-pairs loop
    of updateAttach (VX3_VoxelyzeKernel.cu:448-458): a 10x10x4 block squeezed to
    0.8 of its spacing (every lattice diagonal inside the contact envelope) and
    jostled by 0.06 voxel, so that pairs straddle the grid cells in every
    direction.  Same pairs in the same order, bit-identical states."""
    p = quiet((10, 10, 4), enable_collision=1, col_damp_z=0.8, bond_damp_z=0.1)
    a, b = O.OracleModel(p), O.OracleModel(p)
    rng = np.random.default_rng(7)
    pos = a.state()["pos"]
    pos = 0.8 * (pos - pos.mean(axis=0)) + rng.normal(scale=0.0006, size=pos.shape)
    lm = rng.normal(scale=1e-6, size=pos.shape)
    for m in (a, b):
        m.set_state(pos=pos, linmom=lm)
    h = 1.25 * 0.01 * 1.0001  # the search's cell edge (largest watch distance)
    crossing = total = 0
    for k in range(20):
        a.step(1)
        b.step(1, flags=O.VXO_ALL_PAIRS)
        pa, pb = a.contact_pairs(), b.contact_pairs()
        assert len(pa) > 0 and np.array_equal(pa, pb), k
        total += len(pa)
        cells = np.floor(a.state()["pos"] / h).astype(np.int64)
        crossing += int(np.any(cells[pa[:, 0]] != cells[pa[:, 1]], axis=1).sum())
        sa, sb = a.state(), b.state()
        for key in ("pos", "quat", "linmom", "angmom"):
            assert np.array_equal(sa[key], sb[key]), (k, key)
    assert total > 5000 and crossing > 4000


def test_divergence_stops_the_step():
    """doTimeStep returns false once a link's axial strain exceeds 100
    (VX3_VoxelyzeKernel.cu:271-280): a bar driven at 40x its stable timestep
    blows up and the oracle stops advancing it; at the stable timestep it runs."""
    p = quiet((3, 1, 1))
    m = O.OracleModel(p)
    ps = m.state()["pos"]
    ps[2, 0] += 1e-4
    m.set_state(pos=ps)
    dt = m.step_dt()
    done = m.step(2000, 40 * dt)
    assert done < 2000
    t = m.time
    assert m.step(10, 40 * dt) == 0 and m.time == t  # stays stopped
    ok = O.OracleModel(p)
    ok.set_state(pos=ps)
    assert ok.step(2000, dt) == 2000


def test_force_field_constant_force_and_operand_order():
    """A force field F_x = 3 - 1 (mtSUB: left minus right) on a free body with
    no gravity, floor or damping: every voxel follows the exact discrete
    sequence x_n = x_0 + a dt^2 n(n+1)/2 with a = 2 N / m."""
    p = quiet((2, 2, 2), force_field={"x": ("mtSUB", ("mtCONST", 3.0), ("mtCONST", 1.0))})
    m = O.OracleModel(p)
    dt = m.step_dt()
    x0 = m.state()["pos"][:, 0].copy()
    mass_inv = float(np.float32(1.0) / np.float32(1000.0 * 0.01 ** 3))  # the material's float 1/m
    n = 200
    m.step(n, dt)
    expect = x0 + 2.0 * mass_inv * float(dt) ** 2 * n * (n + 1) / 2
    assert np.allclose(m.state()["pos"][:, 0], expect, rtol=1e-10, atol=0)


def _cilia_problem(dims, base, cilia=1.0, on_after=0.0, remove=None, **kw):
    """quiet() body with cilia on every voxel: BaseCiliaForce `base` (N), the
    material's Cilia multiplier, TurnOnCiliaAfterThisManySeconds; `remove`
    maps cells to a second, removable material (RemoveFromSimulationAfter...)."""
    n = int(np.prod(dims))
    cells = np.ones(n, dtype=np.int8)
    mats = [dict(E=1e6, rho=1000.0, cte=0.01, cilia=cilia, cilia_on_after=on_after)]
    sec = 0
    if remove is not None:
        mask, after = remove
        cells[np.asarray(mask).reshape(-1)] = 2
        mats.append(dict(mats[0], remove_after=after))
        sec = 1
    p = quiet(dims, cells=cells, enable_cilia=1, secondary_experiment=sec,
              cilia_base=np.tile(np.asarray(base, dtype=np.float64), (n, 1)))
    p["materials"] = mats
    p.update(kw)
    return p


def test_cilia_force_turns_with_the_voxel():
    """One free voxel turned 90 degrees about z: its cilia force is
    Cilia * R(q) base, so a base force along x pushes it along +y with the exact
    sequence y_n = y_0 + a dt^2 n(n+1)/2, a = Cilia |base| / m (a force rotated
    the wrong way would push along -y)."""
    p = _cilia_problem((1, 1, 1), (1e-3, 0.0, 0.0), cilia=2.0)
    m = O.OracleModel(p)
    c = np.sqrt(0.5)
    m.set_state(quat=np.array([[c, 0.0, 0.0, c]]))
    dt = m.step_dt()
    y0 = m.state()["pos"][0, 1]
    mass_inv = float(np.float32(1.0) / np.float32(1000.0 * 0.01 ** 3))
    n = 300
    m.step(n, dt)
    s = m.state()
    expect = y0 + 2.0 * 1e-3 * mass_inv * float(dt) ** 2 * n * (n + 1) / 2
    assert s["pos"][0, 1] == pytest.approx(expect, rel=1e-12)
    assert abs(s["pos"][0, 0]) < 1e-12 * abs(expect - y0)


def _steps_before(t_on, dt, n):
    """Steps among the first n whose pre-step time has reached t_on (the
    reference's test `TurnOn > currentTime` fails), with the time accumulated as
    the reference does (double += float dt)."""
    t, k = 0.0, 0
    for _ in range(n):
        if not (t_on > t):
            k += 1
        t += float(np.float32(dt))
    return k


def test_cilia_only_on_surface_voxels_after_turn_on():
    """3x3x3 cube, every voxel with the same cilia force: only the 26 surface
    voxels are pushed (the centre is interior), and only on steps whose start
    time has reached TurnOnCiliaAfterThisManySeconds.  Links conserve momentum,
    so the total momentum is (active steps) x dt x 26 x F."""
    dt = None
    F = 1e-3
    for t_on in (0.0, 0.004):
        p = _cilia_problem((3, 3, 3), (F, 0.0, 0.0), on_after=t_on)
        m = O.OracleModel(p)
        dt = m.step_dt()
        n = 200
        m.step(n, dt)
        px = m.state()["linmom"][:, 0].sum()
        active = _steps_before(t_on, dt, n)
        assert 0 < active <= n
        assert px == pytest.approx(active * float(dt) * 26 * F, rel=1e-6)


def test_removal_step_frozen_voxels_and_free_remainder():
    """A 3-voxel bar falling freely; the middle voxel's material is removed
    after T.  It is removed at the end of the first step whose start time
    exceeds T (both links with it), then stays where it was, and the two end
    voxels keep to the exact free-fall sequence (a link left in place would
    pull them towards the frozen middle)."""
    T = 0.002
    p = _cilia_problem((3, 1, 1), (0.0, 0.0, 0.0), cilia=0.0, remove=([False, True, False], T), grav_acc=-9.81)
    m = O.OracleModel(p)
    dt = m.step_dt()
    t, k = 0.0, 0
    while not (T < t):
        t += float(np.float32(dt))
        k += 1
    k += 1  # the step that starts at t
    grav_mult = np.float32(-np.float32(-9.81) / 9.80665)
    mass = np.float32(1000.0 * 0.01 ** 3)
    gforce = float(np.float32(np.float32(-mass * np.float32(9.80665)) * grav_mult))
    a = gforce * float(np.float32(1.0) / mass)
    z0 = m.state()["pos"][:, 2].copy()
    m.step(k - 1, dt)
    assert m.removed()[0].tolist() == [0, 0, 0]
    m.step(1, dt)
    vr, lr = m.removed()
    assert vr.tolist() == [0, k, 0] and lr.tolist() == [k, k]
    z_k = m.state()["pos"][1, 2]
    assert z_k == pytest.approx(z0[1] + a * float(dt) ** 2 * k * (k + 1) / 2, rel=1e-12)
    n = k + 150
    m.step(n - k, dt)
    z = m.state()["pos"][:, 2]
    assert z[1] == z_k
    for v in (0, 2):
        assert z[v] == pytest.approx(z0[v] + a * float(dt) ** 2 * n * (n + 1) / 2, rel=1e-12)


def test_removal_regenerates_the_surface_for_cilia():
    """3x3x3 cube with cilia everywhere; the x = 0 face (9 voxels, 21 links) is
    removed after T.  The surface list then holds the 18 remaining voxels (the
    former centre included), and from the next step on the remaining voxels
    gain momentum dt x 18 x F per step."""
    T = 0.002
    F = 1e-3
    mask = np.zeros((3, 3, 3), dtype=bool)
    mask[:, :, 0] = True  # cells are x fastest: [z, y, x]
    p = _cilia_problem((3, 3, 3), (F, 0.0, 0.0), remove=(mask, T))
    m = O.OracleModel(p)
    assert m.topology()["surface"].size == 26
    dt = m.step_dt()
    k = 0
    t = 0.0
    while not (T < t):
        t += float(np.float32(dt))
        k += 1
    k += 1
    m.step(k, dt)
    vr, lr = m.removed()
    assert (vr > 0).sum() == 9 and (lr > 0).sum() == 21
    keep = vr == 0
    assert m.topology()["surface"].size == 18
    p1 = m.state()["linmom"][keep, 0].sum()
    m.step(100, dt)
    p2 = m.state()["linmom"][keep, 0].sum()
    assert p2 - p1 == pytest.approx(100 * float(dt) * 18 * F, rel=1e-6)


def _chain(n=10, rho=0.8, delay=0.002, inactive=0.003, period=0.05, **kw):
    """A free n x 1 x 1 bar whose voxel 0 is a pacemaker (no gravity, floor,
    actuation or damping)."""
    cells = np.ones(n, dtype=np.int8)
    cells[0] = 2
    sig = dict(signal_decay=rho, signal_delay=delay, inactive_period=inactive)
    p = quiet((n, 1, 1), cells=cells, enable_signals=1)
    p["materials"] = [dict(E=1e6, rho=1000.0, **sig), dict(E=1e6, rho=1000.0, is_pacemaker=1, pacemaker_period=period, **sig)]
    p.update(kw)
    return p


def test_signal_front_speed_value_and_no_echo():
    """A pacemaker pulse runs down a bar: voxel k first holds 100 rho^k (the
    value decays by signalValueDecay per hop), one signalTimeDelay (rounded up
    to a step) after voxel k - 1; the next own section decays it by 0.9; the
    inactive period keeps the pulse from coming back, so every local signal
    only falls until the next pacemaker pulse."""
    rho, delay = 0.8, 0.002
    m = O.OracleModel(_chain(rho=rho, delay=delay))
    dt = float(np.float32(m.step_dt()))
    n = 10
    hist = []
    for _ in range(600):
        m.step(1)
        hist.append(m.signals().copy())
    hist = np.array(hist)  # step x voxel
    t = np.arange(600) * dt  # start time of each step (exact enough for the bounds below)
    first = [int(np.argmax(hist[:, k] > 0)) for k in range(n)]
    assert first[0] == 0 and all(hist[first[k], k] > 0 for k in range(n))
    for k in range(1, n):
        assert hist[first[k], k] == pytest.approx(100 * rho ** k, rel=1e-12), k
        assert hist[first[k] + 1, k] == pytest.approx(0.9 * 100 * rho ** k, rel=1e-12), k
        if k > 1:
            gap = t[first[k]] - t[first[k - 1]]
            assert delay - 1e-9 <= gap <= delay + dt + 1e-9, (k, gap)
        assert np.all(np.diff(hist[first[k]:, k]) <= 0), k


def test_cilia_force_scales_with_local_signal():
    """A lone pacemaker voxel with a ShiftCiliaForce and no base force: its
    momentum grows by dt Cilia localSignal shift per step (the signal entering
    the step), checked wherever the signal stays constant."""
    p = _chain(n=1, period=1.0, enable_cilia=1, cilia_shift=np.array([[1e-3, 0.0, 0.0]]))
    p["materials"][1].update(cilia=2.0)
    m = O.OracleModel(p)
    dt = float(np.float32(m.step_dt()))
    prev_p, prev_s = m.state()["linmom"][0, 0], m.signals()[0]
    checked = 0
    for _ in range(400):
        m.step(1)
        px, s = m.state()["linmom"][0, 0], m.signals()[0]
        assert px - prev_p == pytest.approx(dt * 2.0 * prev_s * 1e-3, rel=1e-9, abs=1e-18)
        checked += prev_s > 0
        prev_p, prev_s = px, s
    assert checked > 300 and 0 < prev_s < 90


def test_detach_failed_link():
    """Two voxels pulled apart by opposite external forces through a link that
    fails at Fail_Stress: with EnableDetach the link leaves both voxels at the end
    of the first step whose strain passes the failure strain; from then on the
    voxels are free bodies under their constant forces (exact discrete
    sequence).  With collisions on, the now-unlinked neighbours overlap the
    contact envelope (1.25 voxel lengths) and are pushed apart much further."""
    F, fail, E = 0.5, 1e3, 1e6
    sep = {}
    for col in (0, 1):
        p = quiet((2, 1, 1), enable_detach=1, enable_collision=col,
                  externals=[dict(voxel=0, force=(-F, 0, 0)), dict(voxel=1, force=(F, 0, 0))])
        p["materials"] = [dict(E=E, rho=1000.0, fail_stress=fail)]
        m = O.OracleModel(p)
        dt = float(np.float32(m.step_dt()))
        k = 0
        while True:
            m.step(1)
            k += 1
            _, strain = m.link_state()
            if strain[0] > np.float32(fail) / np.float32(E):
                break
            assert m.removed()[1][0] == 0 and k < 5000
        assert m.removed()[1].tolist() == [k]
        s0 = m.state()
        n = 200
        m.step(n)
        s1 = m.state()
        sep[col] = s1["pos"][1, 0] - s1["pos"][0, 0]
        if not col:
            mass_inv = float(np.float32(1.0) / np.float32(1000.0 * 0.01 ** 3))
            for v, sgn in ((0, -1), (1, 1)):
                x = s0["pos"][v, 0] + s0["linmom"][v, 0] * mass_inv * dt * n + sgn * F * mass_inv * dt ** 2 * n * (n + 1) / 2
                assert s1["pos"][v, 0] == pytest.approx(x, rel=1e-12)
    assert sep[1] > sep[0] + 0.001


def test_target_contacts_count_and_signal():
    """A non-target and a target voxel (not linked) pushed together: every step
    that starts with the pair inside the contact envelope (1.25 voxel lengths
    for unheated voxels) adds one to collisionCount; two non-targets add
    nothing.  With signals on, the non-target voxel's first contact step leaves
    its local signal at 0.9 x 100 (received forced, then decayed by its own
    section)."""
    for target in (1, 0):
        p = quiet((3, 1, 1), cells=[1, 0, 2], enable_collision=1, enable_signals=1,
                  externals=[dict(voxel=0, force=(0.05, 0, 0)), dict(voxel=1, force=(-0.05, 0, 0))])
        p["materials"] = [dict(E=1e6, rho=1000.0), dict(E=1e6, rho=1000.0, is_target=target)]
        m = O.OracleModel(p)
        expect, first = 0, None
        for k in range(400):
            pos = m.state()["pos"]
            inside = np.linalg.norm(pos[1] - pos[0]) <= 1.25 * 0.01
            m.step(1)
            expect += int(inside and target)
            if inside and first is None:
                first = k
                if target:
                    assert m.signals()[0] == pytest.approx(90.0, rel=1e-15) and m.signals()[1] == 0
        assert first is not None and m.collision_count() == expect
        assert (expect > 5) == bool(target)
