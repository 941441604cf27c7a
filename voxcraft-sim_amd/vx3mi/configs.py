"""Seeded synthetic workloads (problem dicts) for the five benchmark configurations.

Each generator returns plain problem dicts (see ``vx3mi.PROBLEM_DEFAULTS``)
that carry exactly what a VXA/VXD pair carries; ``vxa.write_vxa`` turns one
into real VXA text.  Phases are given to the simulator in cycles (the
reference evaluates sin(2*pi*(t/period + phase)), VX3_VoxelyzeKernel.cu:639),
so a phase drawn in radians is divided by 2*pi.
"""
from __future__ import annotations

import numpy as np
from scipy import ndimage

# environment shared by every configuration (config 0 text): gravity -9.81,
# floor on, mu_s 1.0, mu_k 0.5, global damping 1.0, collision damping 0.8,
# 39 degC amplitude at 5 Hz, DtFrac 0.9.  BondDampingZ is written explicitly
# (0.1, the reference ReadXML default) because the reference leaves it
# uninitialised when absent (src/VXA/VX_Sim.cpp:197-212).
BASE = dict(lattice_dim=0.01, dt_frac=0.9, bond_damp_z=0.1, col_damp_z=0.8, slow_damp_z=1.0,
            volume_effects=0, enable_expansion=0, grav_acc=-9.81, floor_enabled=1,
            temp_enabled=1, vary_temp=1, temp_base=25.0, temp_amplitude=39.0, temp_period=0.2)

SOFT = dict(E=1e6, rho=1000.0, nu=0.35, cte=0.01, mu_s=1.0, mu_k=0.5, rgba=(0.2, 0.6, 1.0, 1.0))
STIFF = dict(E=1e7, rho=1000.0, nu=0.35, cte=0.01, mu_s=1.0, mu_k=0.5, rgba=(1.0, 0.3, 0.2, 1.0))


def largest_component(occ: np.ndarray) -> np.ndarray:
    """Largest face-connected (6-neighbour) component of a boolean (x, y, z) grid."""
    lab, n = ndimage.label(occ, structure=ndimage.generate_binary_structure(3, 1))
    if n == 0:
        return occ
    sizes = np.bincount(lab.ravel())
    sizes[0] = 0
    return lab == int(np.argmax(sizes))


def _cells_from_grid(mat_xyz: np.ndarray) -> np.ndarray:
    """(x, y, z) material grid -> flat array with x fastest (VXC layer order)."""
    return np.ascontiguousarray(np.transpose(mat_xyz, (2, 1, 0))).reshape(-1).astype(np.int8)


def random_robot(rng: np.random.Generator, dims, p_occ: float, two_materials: bool = True,
                 **overrides) -> dict:
    dims = tuple(int(d) for d in dims)
    occ = largest_component(rng.random(dims) < p_occ)
    mats = rng.integers(1, 3, size=dims) if two_materials else np.ones(dims, dtype=np.int64)
    grid = np.where(occ, mats, 0)
    cells = _cells_from_grid(grid)
    nv = int((cells > 0).sum())
    phase = rng.uniform(0.0, 2 * np.pi, size=nv) / (2 * np.pi)
    prob = dict(BASE, dims=dims, cells=cells, phase=phase,
                materials=[SOFT, STIFF] if two_materials else [SOFT], enable_collision=1)
    prob.update(overrides)
    return prob


def config0() -> dict:
    """Plain method, FP64, collisions off: 4x4x4 solid cube, one material, phase 0."""
    cells = np.ones(64, dtype=np.int8)
    return dict(BASE, dims=(4, 4, 4), cells=cells, phase=np.zeros(64), materials=[SOFT], enable_collision=0)


def config1(seed: int = 0, dims=(64, 64, 16)) -> dict:
    """Single large body, collisions on: Bernoulli(0.7), largest component, 2 materials."""
    rng = np.random.default_rng(seed)
    return random_robot(rng, dims, 0.7, enable_collision=1)


def config2(seed: int = 0, n: int = 1024, lo: int = 4, hi: int = 12, collisions: int = 0) -> list:
    """Batched: n robots, dims U{lo..hi} per axis, Bernoulli(0.6), largest component."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dims = rng.integers(lo, hi + 1, size=3)
        out.append(random_robot(rng, dims, 0.6, enable_collision=collisions))
    return out


def config3(seed: int = 0, dims=(8, 8, 8)) -> dict:
    """Differentiable robot: solid block, per-voxel phase and log10 E in [6, 7]."""
    rng = np.random.default_rng(seed)
    n = int(np.prod(dims))
    phase = rng.uniform(0.0, 2 * np.pi, size=n) / (2 * np.pi)
    logE = rng.uniform(6.0, 7.0, size=n)
    return dict(BASE, dims=tuple(dims), cells=np.ones(n, dtype=np.int8), phase=phase,
                voxel_E=(10.0 ** logE).astype(np.float32), materials=[SOFT], enable_collision=0)


def config4(seed: int = 0, n: int = 256) -> list:
    """Batched adjoint: n robots, dims U{6..16}, p=0.6, per-voxel phase and E."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        dims = rng.integers(6, 17, size=3)
        prob = random_robot(rng, dims, 0.6, two_materials=False, enable_collision=0)
        nv = int((prob["cells"] > 0).sum())
        prob["voxel_E"] = (10.0 ** rng.uniform(6.0, 7.0, size=nv)).astype(np.float32)
        out.append(prob)
    return out


def counts(problems) -> tuple:
    """(voxels, links) of a problem or list of problems, from the lattice alone."""
    if isinstance(problems, dict):
        problems = [problems]
    nv = nl = 0
    for p in problems:
        nx, ny, nz = p["dims"]
        g = (np.asarray(p["cells"]).reshape(nz, ny, nx) > 0)
        nv += int(g.sum())
        nl += int((g[:, :, 1:] & g[:, :, :-1]).sum() + (g[:, 1:, :] & g[:, :-1, :]).sum()
                  + (g[1:, :, :] & g[:-1, :, :]).sum())
    return nv, nl
