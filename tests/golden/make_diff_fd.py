"""Writes the finite-difference gradient fixtures of the differentiable mode.

diff_fd_4x4x4.npz: the 4x4x4 corner subset of the seeded 8x8x8 differentiable
robot (configuration 3).  diff_fd_c4_robot0.npz: the 4x4x4 box of robot 0 of
the seeded batched-adjoint workload (configuration 4: dims 6..16, occupancy
0.6, largest component) that holds the most voxels of one face-connected
component, that component kept.  Both: per-voxel phase psi in [0, 2 pi) and
log10 E in [6, 7] from the seed, softplus floor (k 1e5 N/m, w 1e-4 m),
collisions off, 200 steps at the timestep of E = 1e7.  Loss = -(x
displacement of the mass-weighted centre of mass).  Gradients with respect to
psi (radians) and log10 E are FP64 central differences of the oracle's
differentiable mode with eps = 1e-6.

    python tests/golden/make_diff_fd.py
"""
import math
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "voxcraft-sim_amd"), os.path.join(ROOT, "oracle")]

from vx3mi import configs  # noqa: E402
import oracle as O  # noqa: E402

NSTEPS = 200
EPS = 1e-6
DT = float(np.float32(0.9 / (2 * math.pi * math.sqrt(1e7 * 0.01 / 1e-3))))
OUT = os.path.join(HERE, "diff_fd_4x4x4.npz")
OUT_C4 = os.path.join(HERE, "diff_fd_c4_robot0.npz")


def subset_problem():
    """4x4x4 corner of config 3's seeded parameters (x fastest)."""
    full = configs.config3()
    nx = ny = nz = 8
    idx = [x + nx * (y + ny * z) for z in range(4) for y in range(4) for x in range(4)]
    psi = np.asarray(full["phase"])[idx] * 2 * math.pi
    log10E = np.log10(np.asarray(full["voxel_E"], dtype=np.float64)[idx])
    prob = dict(configs.BASE, dims=(4, 4, 4), cells=np.ones(64, np.int8), materials=[configs.SOFT],
                enable_collision=0)
    return prob, psi, log10E


def c4_subset_problem(robot=0, box=4):
    """The box x box x box part of a configuration-4 robot holding the largest
    face-connected component (ties: the first box in x-fastest order), with the
    robot's per-voxel phase and modulus; cells in the subset's x-fastest order."""
    full = configs.config4()[robot]
    nx, ny, nz = full["dims"]
    occ = np.transpose(np.asarray(full["cells"]).reshape(nz, ny, nx) > 0, (2, 1, 0))  # (x, y, z)
    order = np.transpose(occ, (2, 1, 0)).reshape(-1)  # x fastest
    ids = -np.ones(order.size, np.int64)
    ids[order] = np.arange(int(order.sum()))
    vid = np.transpose(ids.reshape(nz, ny, nx), (2, 1, 0))
    best, where = None, None
    for z in range(max(1, nz - box + 1)):
        for y in range(max(1, ny - box + 1)):
            for x in range(max(1, nx - box + 1)):
                sub = configs.largest_component(occ[x:x + box, y:y + box, z:z + box])
                if best is None or sub.sum() > best.sum():
                    best, where = sub, (x, y, z)
    x, y, z = where
    ids_sub = vid[x:x + box, y:y + box, z:z + box]
    cells = np.transpose(best, (2, 1, 0)).reshape(-1).astype(np.int8)
    keep = np.transpose(ids_sub, (2, 1, 0)).reshape(-1)[cells > 0]
    psi = np.asarray(full["phase"])[keep] * 2 * math.pi
    log10E = np.log10(np.asarray(full["voxel_E"], dtype=np.float64)[keep])
    prob = dict(configs.BASE, dims=best.shape, cells=cells, materials=[configs.SOFT], enable_collision=0)
    return prob, psi, log10E


def loss(prob, psi, log10E):
    o = O.OracleDiff(prob, psi / (2 * math.pi), 10.0 ** log10E, dt=DT)
    x0 = o.state()["pos"][:, 0].mean()
    o.step(NSTEPS)
    m = o.masses()
    return -((o.state()["pos"][:, 0] * m).sum() / m.sum() - x0)


def fd_gradients(prob, psi, log10E, which=None):
    n = psi.size
    which = range(n) if which is None else which
    g_psi, g_l = np.full(n, np.nan), np.full(n, np.nan)
    for i in which:
        for arr, g in ((psi, g_psi), (log10E, g_l)):
            a = arr.copy()
            a[i] = arr[i] + EPS
            lp = loss(prob, *((a, log10E) if arr is psi else (psi, a)))
            a[i] = arr[i] - EPS
            lm = loss(prob, *((a, log10E) if arr is psi else (psi, a)))
            g[i] = (lp - lm) / (2 * EPS)
    return g_psi, g_l


def write(out, prob, psi, log10E):
    g_psi, g_l = fd_gradients(prob, psi, log10E)
    np.savez_compressed(out, psi=psi, log10E=log10E, dt=DT, nsteps=NSTEPS, eps=EPS, loss=loss(prob, psi, log10E),
                        grad_psi=g_psi, grad_log10E=g_l, cells=np.asarray(prob["cells"]))
    print(f"wrote {out}: {psi.size} voxels, loss {loss(prob, psi, log10E):.6e}, |g_psi| {np.linalg.norm(g_psi):.3e}, "
          f"|g_log10E| {np.linalg.norm(g_l):.3e}")


def main():
    which = sys.argv[1:] or ["c3", "c4"]
    if "c3" in which:
        write(OUT, *subset_problem())
    if "c4" in which:
        write(OUT_C4, *c4_subset_problem())


if __name__ == "__main__":
    main()
