"""Writes golden trajectories of the reference's own CPU Voxelyze.

Runs oracle/_ref/vx3_ref_driver (compiled by `make -C oracle ref` from the
reference sources src/old/Voxelyze.cpp, VX_Voxel.cpp, VX_Link.cpp, ...) on two
small robots and stores positions and orientations every `every` steps as
compressed .npz fixtures next to this script.  The fixtures are data (inputs and
outputs); the reference itself is not needed to read them.

    python tests/golden/make_golden.py [name ...]
"""
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "voxcraft-sim_amd"), os.path.join(ROOT, "oracle")]

from vx3mi import configs  # noqa: E402
import oracle as O  # noqa: E402

DRIVER = os.path.join(ROOT, "oracle", "_ref", "vx3_ref_driver")


def driver_input(prob, dt, nsteps, every):
    nx, ny, nz = prob["dims"]
    cells = np.asarray(prob["cells"]).reshape(-1)
    nv = int((cells > 0).sum())
    phase = np.zeros(nv) if prob.get("phase") is None else np.asarray(prob["phase"])
    mats = prob["materials"]
    lines = [f"{nx} {ny} {nz} {prob['lattice_dim']!r} {prob['bond_damp_z']!r} {prob['col_damp_z']!r} "
             f"{prob['slow_damp_z']!r} {prob['grav_acc']!r} {int(prob['floor_enabled'])} "
             f"{prob['temp_amplitude']!r} {prob['temp_period']!r} {int(prob['vary_temp'])} "
             f"{int(prob['enable_expansion'])} 0 {len(mats)}"]
    for m0 in mats:
        m = dict(O.MATERIAL_DEFAULTS, **m0)
        line = (f"{m['E']!r} {m['rho']!r} {m['cte']!r} {m['mu_s']!r} {m['mu_k']!r} {O.material_model(m)} "
                f"{m['plastic_mod']!r} {m['yield_stress']!r} {m['fail_stress']!r}")
        if m.get("ss_data") is not None:
            st, ss = m["ss_data"]
            line += f" {len(st)} " + " ".join(f"{float(np.float32(a))!r} {float(np.float32(b))!r}" for a, b in zip(st, ss))
        lines.append(line)
    lines.append(" ".join(str(int(c)) for c in cells))
    lines.append(" ".join(repr(float(p)) for p in phase))
    lines.append(f"{float(np.float32(dt)):.9g} {nsteps} {every}")
    ext = prob.get("externals") or []
    lines.append(str(len(ext)))
    for e in ext:
        vals = [e["voxel"], e.get("dof", 0)]
        vals += [repr(float(x)) for x in e.get("translation", (0, 0, 0))]
        vals += [repr(float(x)) for x in e.get("rotation", (0, 0, 0))]
        vals += [repr(float(np.float32(x))) for x in e.get("force", (0, 0, 0))]
        vals += [repr(float(np.float32(x))) for x in e.get("moment", (0, 0, 0))]
        lines.append(" ".join(str(v) for v in vals))
    return "\n".join(lines).encode()


def run_reference(prob, nsteps, every):
    dt = np.float32(O.OracleModel(prob).step_dt())
    r = subprocess.run([DRIVER], input=driver_input(prob, dt, nsteps, every), capture_output=True, check=True)
    nv = int((np.asarray(prob["cells"]) > 0).sum())
    rec = np.frombuffer(r.stdout, dtype=np.float64).reshape(-1, nv, 13)
    return dt, rec


def beam_with_externals():
    """A 6x2x3 cantilever: the x=0 face fully fixed (one voxel displaced, one
    rotated), loads and a moment at the tip, a partial rotation fix and a
    prescribed x translation in the span; gravity on, floor off."""
    p = dict(configs.config0(), dims=(6, 2, 3), cells=np.ones(36, dtype=np.int8), phase=None, floor_enabled=0,
             vary_temp=0, temp_amplitude=0.0, enable_collision=0, bond_damp_z=0.1, slow_damp_z=0.01)
    vid = lambda x, y, z: x + 6 * (y + 2 * z)  # noqa: E731  voxel order of a full lattice
    ext = []
    for y in range(2):
        for z in range(3):
            e = dict(voxel=vid(0, y, z), dof=63)
            if (y, z) == (1, 2):
                e["translation"] = (0.0, 0.0, 5e-4)
            if (y, z) == (0, 1):
                e["rotation"] = (0.0, 0.02, 0.0)
            ext.append(e)
    for y in range(2):
        for z in range(3):
            ext.append(dict(voxel=vid(5, y, z), force=(0.0, 0.002, -0.01)))
    ext[-1]["moment"] = (1e-4, 0.0, 2e-5)
    ext.append(dict(voxel=vid(3, 0, 0), dof=8 | 32))
    ext.append(dict(voxel=vid(2, 1, 1), dof=1, translation=(2e-4, 0.0, 0.0), force=(0.0, 0.0, 0.003)))
    p["externals"] = ext
    return p


def bilinear_robot():
    """A 5x4x3 robot of a bilinear material (yield strain 0.002, plastic modulus
    E/5) and a linear one, actuated so that links yield, unload and reload;
    floor off (the reference CPU path has no static-friction angMom reset)."""
    rng = np.random.default_rng(11)
    p = configs.random_robot(rng, (5, 4, 3), 0.9, enable_collision=0)
    p.update(floor_enabled=0, grav_acc=0.0, temp_amplitude=39.0, vary_temp=1, temp_period=0.004,
             phase=rng.uniform(0, 1, int((np.asarray(p["cells"]) > 0).sum())),
             materials=[dict(E=1e6, rho=1000.0, cte=0.001, yield_stress=2000.0, plastic_mod=2e5),
                        dict(E=2e6, rho=1000.0, cte=0.001)])
    return p


def data_model_robot():
    """A 5x4x3 robot of two MatModel 3 materials (stress-strain data: one
    with a leading (0, 0) point and four segments of falling slope, one of
    three points), actuated as the bilinear robot so links run along their
    curves; floor off."""
    rng = np.random.default_rng(12)
    p = configs.random_robot(rng, (5, 4, 3), 0.9, enable_collision=0)
    p.update(floor_enabled=0, grav_acc=0.0, temp_amplitude=39.0, vary_temp=1, temp_period=0.004,
             phase=rng.uniform(0, 1, int((np.asarray(p["cells"]) > 0).sum())),
             materials=[dict(rho=1000.0, cte=0.001,
                             ss_data=([0.0, 0.002, 0.006, 0.015, 0.05], [0.0, 2000.0, 4400.0, 7000.0, 9000.0])),
                        dict(rho=1000.0, cte=0.001, ss_data=([0.001, 0.004, 0.03], [2000.0, 5000.0, 8000.0]))])
    return p


def golden_problems():
    rng = np.random.default_rng(7)
    robot = configs.random_robot(rng, (5, 4, 3), 0.8, enable_collision=0)
    return {"config0": configs.config0(), "robot5x4x3": robot, "beam_bc": beam_with_externals(),
            "bilinear": bilinear_robot(), "data_model": data_model_robot()}


def main():
    only = sys.argv[1:]
    for name, prob in golden_problems().items():
        if only and name not in only:
            continue
        dt, rec = run_reference(prob, 600, 50)
        np.savez_compressed(os.path.join(HERE, f"ref_cpu_{name}.npz"), dt=dt, every=50,
                            pos=rec[:, :, 0:3], quat=rec[:, :, 3:7])
        print(name, rec.shape)


if __name__ == "__main__":
    main()
