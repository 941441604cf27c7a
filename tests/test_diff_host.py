"""Differentiable mode, host checks (no GPU).

1. The product's hand-written vector-Jacobian products (vx3mi_diff.h, built for
   the host in double by g++ from tests/diff_vjp_check.cpp) against central
   finite differences of the same forward functions, for every primitive and
   for whole link and voxel updates on random inputs, covering the small- and
   large-angle link branches with and without local damping and the voxel with
   and without the softplus floor.  The score is |g - fd| / (|fd| + rounding
   noise of fd): about 1e-6..1e-4 for a correct derivative, O(1) for a wrong one.
2. The oracle's differentiable mode against closed forms: free fall, the
   softplus floor's rest height, kinetic sliding deceleration, the axial
   oscillation frequency of a two-voxel bar with the harmonic-mean modulus.
"""
import ctypes as C
import math
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from vx3mi import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TOL = 1e-2


@pytest.fixture(scope="module")
def vjp(tmp_path_factory):
    out = tmp_path_factory.mktemp("vjp") / "diff_vjp_check.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "voxcraft-sim_amd", "csrc"),
                    os.path.join(ROOT, "tests", "diff_vjp_check.cpp"), "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    for f in ("check_link", "check_voxel", "check_primitives"):
        getattr(L, f).restype = C.c_double
    L.check_link.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.link_branch.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double]
    L.check_voxel.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
    return L


def test_primitive_adjoints(vjp):
    assert max(vjp.check_primitives(s) for s in range(8)) < TOL


# (flags in, bend, stretch) -> expected (small angle, local velocity valid) after the update
LINK_CASES = [((2 | 1), 0.01, 0.05, (1, 1)), (2, 0.01, 0.05, (1, 0)), (1, 0.3, 0.2, (0, 1)), (0, 0.3, 0.2, (0, 0))]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("flags,bend,stretch,branch", LINK_CASES)
def test_link_adjoint(vjp, axis, flags, bend, stretch, branch):
    for seed in range(6):
        b = vjp.link_branch(seed, axis, flags, bend, stretch)
        assert (b & 1, b >> 1) == branch
        e = vjp.check_link(seed, axis, flags, bend, stretch)
        assert 0 <= e < TOL, (seed, e)


@pytest.mark.parametrize("floor", [0, 1])
@pytest.mark.parametrize("height", [-5e-4, -1e-4, 0.0, 2e-4])
def test_voxel_adjoint(vjp, floor, height):
    assert max(vjp.check_voxel(s, floor, 0, height) for s in range(5)) < TOL


def test_fixed_voxel_adjoint_is_identity(vjp):
    assert vjp.check_voxel(0, 1, 1, 0.0) < TOL


# ---- oracle differentiable mode: closed forms ----------------------------------------
def single(dims=(1, 1, 1), cells=None, **kw):
    n = int(np.prod(dims))
    p = dict(configs.BASE, dims=dims, cells=np.ones(n, np.int8) if cells is None else cells,
             materials=[dict(configs.SOFT, cte=0.0)], enable_collision=0, vary_temp=0)
    p.update(kw)
    return p


def test_free_fall_closed_form():
    dt = 1e-5
    p = single(floor_enabled=0, slow_damp_z=0.0)
    o = O.OracleDiff(p, [0.0], [1e6], dt=dt)
    g = O.OracleModel(p).voxel_material(0)  # mass is entry 6
    n = 400
    o.step(n)
    m = float(g[6])
    grav = float(np.float32(np.float32(-g[6]) * np.float32(9.80665)) * np.float32(-np.float32(-9.81) / 9.80665))
    z = o.state()["pos"][0, 2]
    assert z == pytest.approx(grav / m * dt * dt * n * (n + 1) / 2, rel=1e-12)


def test_softplus_floor_rest_height():
    """At rest k w softplus(pen/w) = m g, so pen = w log(exp(m g / (k w)) - 1).
    (Global damping near critical for the soft rest stiffness, so it settles.)"""
    k, w = 1e5, 1e-4
    p = single(slow_damp_z=0.1)
    o = O.OracleDiff(p, [0.0], [1e6], dt=1e-5, floor_k=k, floor_w=w)
    o.step(50000)
    m = O.OracleModel(p).voxel_material(0)[6]
    weight = float(m) * 9.81
    pen = w * math.log(math.expm1(weight / (k * w)))
    assert -o.state()["pos"][0, 2] == pytest.approx(pen, rel=1e-6)


def test_kinetic_sliding_deceleration():
    """A voxel sliding well above the regularisation speed decelerates at
    muK N / m with N the softplus floor force (= m g at rest), the
    reference's kinetic friction law.  Settled with global damping, then slid
    on an undamped copy placed at the settled state."""
    dt, v0, n = 1e-5, 0.05, 500
    settled = O.OracleDiff(single(slow_damp_z=0.1), [0.0], [1e6], dt=dt)
    settled.step(50000)
    s = settled.state()
    o = O.OracleDiff(single(slow_damp_z=0.0), [0.0], [1e6], dt=dt)
    m = o.masses()[0]
    o.set_state(pos=s["pos"], quat=s["quat"], linmom=np.array([[v0 * m, 0.0, 0.0]]), angmom=s["angmom"])
    o.step(n)
    v = o.state()["linmom"][0, 0] / m
    mu, g, ve = 0.5, 9.81, 1e-3
    # v/sqrt(v^2 + ve^2) stays within 3e-4 of 1 while v > 0.025
    expect = v0 - mu * g * n * dt
    assert v == pytest.approx(expect, rel=1e-3)
    assert abs(o.state()["linmom"][0, 2] / m) < 1e-4  # still resting vertically


def test_two_voxel_bar_frequency_uses_harmonic_modulus():
    """Two free voxels along x, no gravity, floor or damping: the axial mode
    oscillates at sqrt(k (1/m1 + 1/m2)) with k = E_link L and E_link the
    harmonic mean 2 E1 E2 / (E1 + E2) (the arithmetic mean would be 2.5e6)."""
    dt = 1e-6
    p = single(dims=(2, 1, 1), floor_enabled=0, grav_acc=0.0, slow_damp_z=0.0, bond_damp_z=0.0)
    E1, E2 = 1e6, 4e6
    o = O.OracleDiff(p, [0.0, 0.0], [E1, E2], dt=dt)
    pos = o.state()["pos"].copy()
    pos[1, 0] += 1e-5
    o.set_state(pos=pos)
    L, m = 0.01, o.masses()[0]
    k = 2 * E1 * E2 / (E1 + E2) * L
    omega = math.sqrt(2 * k / m)
    crossings, prev = [], None
    for i in range(4000):
        o.step(1)
        d = o.state()["pos"][1, 0] - o.state()["pos"][0, 0] - L
        if prev is not None and prev > 0 >= d:
            crossings.append(i + prev / (prev - d))
        prev = d
    period = (crossings[-1] - crossings[0]) / (len(crossings) - 1) * dt
    assert period == pytest.approx(2 * math.pi / omega, rel=2e-3)


def test_fd_fixture_reproduces_from_oracle():
    """The committed finite-difference gradients (tests/golden/diff_fd_4x4x4.npz)
    are what the oracle gives now, on a sample of parameters."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_diff_fd as FD
    fix = np.load(os.path.join(ROOT, "tests", "golden", "diff_fd_4x4x4.npz"))
    prob, psi, log10E = FD.subset_problem()
    assert np.array_equal(psi, fix["psi"]) and np.array_equal(log10E, fix["log10E"])
    which = [0, 21, 42, 63]
    g_psi, g_l = FD.fd_gradients(prob, psi, log10E, which=which)
    assert np.allclose(g_psi[which], fix["grad_psi"][which], rtol=1e-9, atol=0)
    assert np.allclose(g_l[which], fix["grad_log10E"][which], rtol=1e-9, atol=0)
    assert FD.loss(prob, psi, log10E) == pytest.approx(float(fix["loss"]), rel=1e-12)


@pytest.fixture(scope="module")
def adjoint_host(tmp_path_factory):
    out = tmp_path_factory.mktemp("adj") / "diff_adjoint_host.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "voxcraft-sim_amd", "csrc"),
                    os.path.join(ROOT, "tests", "diff_adjoint_host.cpp"),
                    os.path.join(ROOT, "voxcraft-sim_amd", "csrc", "vx3mi_model.cpp"), "-o", str(out)], check=True)
    L = C.CDLL(str(out))
    dp = C.POINTER(C.c_double)
    L.harness_adjoint.argtypes = [C.c_void_p, dp, dp, C.c_int, C.c_double, C.c_int, dp, dp, dp, dp]
    return L


def host_adjoint(L, prob, phase, E, nsteps, dt, K, gpos):
    import vx3mi
    # the lattice description the product's builder reads (built like Model.from_problem)
    p = dict(vx3mi.PROBLEM_DEFAULTS)
    p.update(prob)
    nx, ny, nz = p["dims"]
    cells = np.ascontiguousarray(p["cells"], dtype=np.int8).reshape(-1)
    mats = [dict(vx3mi.MATERIAL_DEFAULTS, **m) for m in p["materials"]]
    marr = (vx3mi._Material * len(mats))()
    for i, m in enumerate(mats):
        r, g, b, a = m["rgba"]
        marr[i] = vx3mi._Material(m["E"], m["rho"], m["nu"], m["cte"], m["mu_s"], m["mu_k"], -1.0, 0.0, 0.0, 0,
                                  int(m["fixed"]), int(m["is_measured"]), 0, 0.0, r, g, b, a)
    spec = vx3mi._Lattice(nx, ny, nz, cells.ctypes.data_as(C.POINTER(C.c_int8)), None, None, len(mats), marr,
                          p["lattice_dim"], p["dt_frac"], p["bond_damp_z"], p["col_damp_z"], p["slow_damp_z"], 0, 0, 0,
                          p["grav_acc"], int(p["floor_enabled"]), int(p["temp_enabled"]), int(p["vary_temp"]),
                          p["temp_base"], p["temp_amplitude"], p["temp_period"])
    nv = int((cells > 0).sum())
    ph, e = np.ascontiguousarray(phase, np.float64), np.ascontiguousarray(E, np.float64)
    g = np.ascontiguousarray(gpos, np.float64)
    gphase, gE, pos = np.zeros(nv), np.zeros(nv), np.zeros((nv, 3))
    ptr = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    assert L.harness_adjoint(C.byref(spec), ptr(ph), ptr(e), nsteps, dt, K, ptr(g), ptr(gphase), ptr(gE), ptr(pos)) == 0
    return gphase, gE, pos


@pytest.mark.parametrize("K", [200, 7])
def test_adjoint_pipeline_on_host_matches_finite_differences(adjoint_host, K):
    """The GPU kernels' work items, run on the host in double: forward parity
    with the oracle's differentiable mode and gradients against the FP64
    finite differences of the fixture, with one segment and with K = 7
    (checkpoint and recompute)."""
    import sys
    sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))
    import make_diff_fd as FD
    fix = np.load(os.path.join(ROOT, "tests", "golden", "diff_fd_4x4x4.npz"))
    prob, psi, log10E = FD.subset_problem()
    E = 10.0 ** log10E
    nsteps, dt = int(fix["nsteps"]), float(fix["dt"])
    gpos = np.zeros((64, 3))
    gpos[:, 0] = -1.0 / 64
    gphase, gE, pos = host_adjoint(adjoint_host, prob, psi / (2 * math.pi), E, nsteps, dt, K, gpos)
    o = O.OracleDiff(prob, psi / (2 * math.pi), E, dt=dt)
    o.step(nsteps)
    assert np.max(np.abs(pos - o.state()["pos"])) < 1e-14
    g_psi, g_l = gphase / (2 * math.pi), gE * E * math.log(10.0)
    rel_psi = np.linalg.norm(g_psi - fix["grad_psi"]) / np.linalg.norm(fix["grad_psi"])
    rel_l = np.linalg.norm(g_l - fix["grad_log10E"]) / np.linalg.norm(fix["grad_log10E"])
    assert rel_psi < 1e-6 and rel_l < 1e-6, (rel_psi, rel_l)
