#!/usr/bin/env python3
"""Benchmark of the VX3 timestep on one MI355X.

Workload (the batched configuration the headline metric is quoted on): 1024
independent robots, lattice dims U{4..12} per axis, Bernoulli(0.6) occupancy,
largest face-connected component, two materials, random actuation phases,
FP32 state, collisions off, seed 0.  A *step* is one timestep of every robot
of the batch (VX3_VoxelyzeKernel::doTimeStep for all 1024 simulations).

Printed: one JSON line with voxel-steps/s of the autotuned batched kernel
(value), the same workload on the plain path (one robot per launch sequence,
as the reference launches per simulation) and the ratio, the relative error
of the FP32 run against the FP64 oracle on a sample of robots, the roofline of
the dominant kernel (its binding roof, VALU issue, from counted instructions
and the timed launch; its measured HBM bytes beside it), the same workload
with collisions on, per-configuration parity and config-1 HBM traffic, the
adjoint's cost, and the reference CPU Voxelyze timed on a bounded sample of
the same robots on this host.  Counter passes run this script as a child
under rocprofv3 (--no-counters skips them).
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "voxcraft-sim_amd"))

import numpy as np  # noqa: E402

import vx3mi  # noqa: E402
from vx3mi import configs  # noqa: E402

HBM_PEAK_GBS = 8000.0
PERSISTENT_CHUNK = 1000  # steps per persistent launch


def algorithmic_bytes_per_voxel_step(nv: int, nl: int, fp64: bool = False) -> float:
    """Compulsory bytes of a step that keeps no state on chip across steps:
    per voxel 13 state words r/w + flags r/w + phase and material + 6 link
    indices; per link 11 state words r/w + endpoints, axis and material."""
    if fp64:
        return 252.0 + 180.0 * nl / nv
    return 144.0 + 100.0 * nl / nv


# VALU issue roof: every compute unit has four SIMDs, each issues one wave64
# VALU instruction per four cycles (16 lanes wide; packed FP32 is not used);
# MI355X: 256 compute units at a peak engine clock of 2.4 GHz
VALU_ISSUE_PEAK_G = 256 * 4 * 2.4 / 4  # G wave-instructions per second
COPY_PROBE_MB = 256


def rocprof_counters(counters, child_args, timeout=120.0):
    """Runs this script as a child under `rocprofv3 --pmc <counters>` (one pass)
    and returns {kernel name: [per-dispatch {counter: value}]} from its CSV, or
    None when rocprofv3 is missing or the run fails."""
    import csv
    import shutil
    import tempfile

    exe = shutil.which("rocprofv3") or ("/opt/rocm/bin/rocprofv3" if os.path.exists("/opt/rocm/bin/rocprofv3") else None)
    if exe is None:
        return None
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        cmd = [exe, "--pmc", *counters, "--output-format", "csv", "-d", d, "-o", "run", "--",
               sys.executable, os.path.abspath(__file__), "--counters-child", *child_args]
        try:
            r = subprocess.run(cmd, capture_output=True, timeout=timeout, cwd=d)
        except subprocess.TimeoutExpired:
            log(f"bench: rocprofv3 --pmc {' '.join(counters)} timed out")
            return None
        files = [os.path.join(dp, f) for dp, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        if r.returncode != 0 or not files:
            log(f"bench: rocprofv3 --pmc {' '.join(counters)} failed (rc {r.returncode})")
            return None
        out = {}
        disp = {}
        with open(files[0]) as f:
            for row in csv.DictReader(f):
                key = (row["Dispatch_Id"], row["Kernel_Name"])
                disp.setdefault(key, {})
                disp[key][row["Counter_Name"]] = disp[key].get(row["Counter_Name"], 0.0) + float(row["Counter_Value"])
        for (did, name), vals in sorted(disp.items(), key=lambda kv: int(kv[0][0])):
            out.setdefault(name, []).append(vals)
        return out


def kernel_rows(rows, prefix):
    return [v for k, vs in rows.items() if k.split("(")[0].replace("void ", "").startswith(prefix) for v in vs]


def copy_calibration(rows, counter):
    """Bytes per counted KiB of `counter`, from the largest copy dispatch (the
    COPY_PROBE_MB device-to-device copy at the child's end)."""
    copies = [v[counter] for k, vs in rows.items() if "copyBuffer" in k for v in vs]
    if not copies or max(copies) <= 0:
        return None
    return COPY_PROBE_MB * 2**20 / max(copies)


def counter_run(kernel_prefix, child_args):
    """HBM bytes (FETCH_SIZE + WRITE_SIZE, calibrated on the copy) and VALU wave
    instructions per dispatch of the kernels whose names start with kernel_prefix,
    from three counter passes; values are summed over those dispatches and
    divided by their number."""
    res = {}
    for counters in (["FETCH_SIZE"], ["WRITE_SIZE"], ["SQ_INSTS_VALU", "SQ_WAVES"]):
        rows = rocprof_counters(counters, child_args)
        if rows is None:
            return None
        ks = kernel_rows(rows, kernel_prefix)
        if not ks:
            return None
        for c in counters:
            res[c] = sum(v[c] for v in ks) / len(ks)
            if c in ("FETCH_SIZE", "WRITE_SIZE"):
                f = copy_calibration(rows, c)
                if f is None:
                    return None
                res[c + "_bytes"] = res[c] * f
                res[c + "_calibration"] = f
        res["dispatches"] = len(ks)
    res["traffic"] = res["FETCH_SIZE_bytes"] + res["WRITE_SIZE_bytes"]
    return res


def copy_probe(mb: int = COPY_PROBE_MB) -> None:
    """One device-to-device hipMemcpy (the runtime's copy kernel) of a known size:
    the reference dispatch that calibrates the byte counters."""
    import ctypes as C

    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(mb * 2**20)
    src, dst = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(src), nbytes) == 0 and hip.hipMalloc(C.byref(dst), nbytes) == 0
    assert hip.hipMemset(src, 1, nbytes) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(dst, src, nbytes, 3) == 0  # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(src)
    hip.hipFree(dst)


def counters_child(args) -> int:
    """The process the counter passes profile: the batched workload (or config 1)
    in the given launch configuration, one launch of args.steps steps after one
    warm-up launch, then the calibration copy."""
    if args.workload == "config1":
        b = vx3mi.Batch([vx3mi.Model.from_problem(configs.config1())], vx3mi.FP32)
        b.step(args.warmup, mode=vx3mi.MODE_FLAT)
        b.step(args.steps, mode=vx3mi.MODE_FLAT)
    else:
        probs = configs.config2(seed=0, n=args.robots, collisions=args.collisions)
        b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
        b.configure(args.child_mode, args.child_tile, args.child_threads)
        b.step(args.steps, mode=vx3mi.MODE_AUTO)  # warm-up launch (code load)
        b.step(args.steps, mode=vx3mi.MODE_AUTO)
    b.sync()
    copy_probe()
    b.close()
    return 0


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), or "unknown"."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(problems, nsteps: int, omp: bool = False):
    """Reference CPU Voxelyze (oracle/_ref, built from the reference sources) on
    a sample of robots; falls back to the FP64 oracle port when absent.  omp:
    the reference's own OpenMP build (Voxelyze.cpp:260-279) on every thread
    OMP_NUM_THREADS allows."""
    ref_exe = os.path.join(ROOT, "oracle", "_ref", "vx3_ref_driver_omp" if omp else "vx3_ref_driver")
    if omp and not os.path.exists(ref_exe):
        return None, None
    if os.path.exists(ref_exe):
        secs = vs = 0.0
        for p in problems:
            nx, ny, nz = p["dims"]
            cells = np.asarray(p["cells"]).reshape(-1)
            nv = int((cells > 0).sum())
            m0 = p["materials"]
            lines = [f"{nx} {ny} {nz} {p['lattice_dim']!r} {p['bond_damp_z']!r} {p['col_damp_z']!r} "
                     f"{p['slow_damp_z']!r} {p['grav_acc']!r} {int(p['floor_enabled'])} {p['temp_amplitude']!r} "
                     f"{p['temp_period']!r} {int(p['vary_temp'])} {int(p['enable_expansion'])} 0 {len(m0)}"]
            for mm in m0:
                m = dict(vx3mi.MATERIAL_DEFAULTS, **mm)
                lines.append(f"{m['E']!r} {m['rho']!r} {m['cte']!r} {m['mu_s']!r} {m['mu_k']!r} "
                             f"{vx3mi.material_model(m)} {m['plastic_mod']!r} {m['yield_stress']!r} "
                             f"{m['fail_stress']!r}")
            lines.append(" ".join(str(int(c)) for c in cells))
            lines.append(" ".join(repr(float(x)) for x in p["phase"]))
            dt = vx3mi.Model.from_problem(p).recommended_dt() * p["dt_frac"]
            lines.append(f"{float(np.float32(dt)):.9g} {nsteps} 0")
            lines.append("0")  # no externals
            r = subprocess.run([ref_exe], input="\n".join(lines).encode(), capture_output=True, timeout=600)
            if r.returncode != 0:
                raise RuntimeError(r.stderr.decode())
            tail = r.stderr.decode().split()
            secs += float(tail[tail.index("seconds") + 1])
            vs += nv * nsteps
        kind = "reference"
    else:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O  # the checker, timed as a port of the reference CPU path

        secs = vs = 0.0
        for p in problems:
            m = O.OracleModel(p)
            t0 = time.perf_counter()
            m.step(nsteps)
            secs += time.perf_counter() - t0
            vs += m.nv * nsteps
        kind = "port"
    return vs / secs, kind


def fp32_error_vs_oracle(problems, nsteps: int) -> float:
    """max over robots of e_pos = max_i |p_i - p_i*| / max(|p_i*|, L) after nsteps."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in problems], vx3mi.FP32)
    b.step(nsteps, mode=vx3mi.MODE_AUTO)
    worst = 0.0
    for r, p in enumerate(problems):
        ref = O.OracleModel(p)
        ref.step(nsteps)
        s, o = b.state(r)["pos"], ref.state()["pos"]
        e = np.max(np.linalg.norm(s - o, axis=1) / np.maximum(np.linalg.norm(o, axis=1), p["lattice_dim"]))
        worst = max(worst, float(e))
    b.close()
    return worst


def adjoint_cost(problems, nsteps: int, checkpoint_every: int, budget: float, precision=None, reverse_threads=0):
    """Forward (with checkpoints) and backward (recompute + reverse sweep) of the
    differentiable mode on the given robots; loss = -(x displacement of each
    robot's centre of mass).  Returns timings and checkpoint sizes."""
    precision = vx3mi.FP32 if precision is None else precision
    models = [vx3mi.Model.from_problem(p) for p in problems]
    d = vx3mi.Diff(models, precision, checkpoint_every=checkpoint_every, hbm_budget=budget,
                   reverse_threads=reverse_threads)
    phase = np.concatenate([np.asarray(p["phase"], dtype=np.float64) for p in problems])
    E = np.concatenate([np.asarray(p["voxel_E"], dtype=np.float64) for p in problems])
    d.set_params(phase, E)
    d.forward(min(nsteps, 200))  # warm-up (code load, buffers)
    d.backward(grad_pos=np.zeros((d.nv, 3)))
    d.forward(nsteps)
    g = np.zeros((d.nv, 3))
    off = d.voxel_offsets
    for r in range(d.nr):
        g[off[r]:off[r + 1], 0] = -1.0 / (off[r + 1] - off[r])
    d.backward(grad_pos=g)
    info = d.info()
    d.close()
    return {"robots": len(problems), "voxels": int(off[-1]), "steps": nsteps,
            "forward_ms": info["forward_ms"], "backward_ms": info["backward_ms"],
            "adjoint_over_forward": info["backward_ms"] / info["forward_ms"],
            "recompute_ms": info["recompute_ms"], "reverse_ms": info["reverse_ms"],
            "checkpoint_every": info["checkpoint_every"], "checkpoint_bytes": info["checkpoint_bytes"],
            "reverse_threads": reverse_threads or 512,  # -1: the flat sweep (one item per thread, all robots)
            "tape_bytes": info["tape_bytes"], "hbm_budget_bytes": budget,
            "forward_voxel_steps_per_s": off[-1] * nsteps / (info["forward_ms"] * 1e-3)}


def e_pos(p, ref, L=0.01):
    return float(np.max(np.linalg.norm(p - ref, axis=1) / np.maximum(np.linalg.norm(ref, axis=1), L)))


def e_quat(q, ref):
    return float(np.max(np.minimum(np.linalg.norm(q - ref, axis=1), np.linalg.norm(q + ref, axis=1))))


def com_rel(c, c0, ref, ref0):
    d, dr = np.asarray(c) - np.asarray(c0), np.asarray(ref) - np.asarray(ref0)
    return float(np.linalg.norm(d - dr) / max(np.linalg.norm(dr), 1e-300))


def parity_report(c1_steps: int = 20000, c1_window: int = 1000, every: int = 50):
    """Relative errors per configuration (position e_pos, orientation e_q, centre-of-mass
    displacement), against the FP64 oracle or the FP64 device run (bit-identical to the
    oracle on every configuration the parity tests cover)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    out = {}
    # config 0: FP64 device vs oracle over 10,000 steps
    p0 = configs.config0()
    b = vx3mi.Batch([vx3mi.Model.from_problem(p0)], vx3mi.FP64)
    o = O.OracleModel(p0)
    c00, r00 = b.com()[0], o.com()
    b.step(10000)
    o.step(10000)
    s, r = b.state(), o.state()
    out["c0_fp64_vs_oracle_10000"] = {"e_pos": e_pos(s["pos"], r["pos"]), "e_q": e_quat(s["quat"], r["quat"]),
                                     "com": com_rel(b.com()[0], c00, o.com(), r00)}
    b.close()
    # config 1: FP32 vs FP64 device, collisions on
    p1 = configs.config1()
    m1 = vx3mi.Model.from_problem(p1)
    b32, b64 = vx3mi.Batch([m1], vx3mi.FP32), vx3mi.Batch([m1], vx3mi.FP64)
    c32, c64 = b32.com()[0], b64.com()[0]
    ep = eq = ep999 = 0.0
    over = 0
    for k in range(0, c1_window, every):
        b32.step(every)
        b64.step(every)
        s32, s64 = b32.state(), b64.state()
        dp = np.linalg.norm(s32["pos"] - s64["pos"], axis=1) / np.maximum(np.linalg.norm(s64["pos"], axis=1), 0.01)
        dq = np.minimum(np.linalg.norm(s32["quat"] - s64["quat"], axis=1),
                        np.linalg.norm(s32["quat"] + s64["quat"], axis=1))
        ep, eq = max(ep, float(dp.max())), max(eq, float(dq.max()))
        ep999 = max(ep999, float(np.quantile(dp, 0.999)))
        over = max(over, int(np.sum((dp > 1e-5) | (dq > 1e-5))))
    rest = c1_steps - c1_window
    b32.step(rest)
    ms32 = b32.last_step_ms()  # HIP events on the batch stream around the steps
    b64.step(rest)
    ms64 = b64.last_step_ms()
    nv1 = int(len(dp))
    out["c1_fp32_vs_fp64"] = {"e_pos_first_1000": ep, "e_q_first_1000": eq,
                              "e_pos_p999_first_1000": ep999,
                              "most_voxels_over_1e-5_at_a_sample": over, "voxels": int(len(dp)),
                              f"com_after_{c1_steps}": com_rel(b32.com()[0], c32, b64.com()[0], c64),
                              "diverged": [int(b32.diverged()[0]), int(b64.diverged()[0])],
                              "voxel_steps_per_s": {"fp32": nv1 * rest / (ms32 * 1e-3), "fp64": nv1 * rest / (ms64 * 1e-3),
                                                    "mode": "flat (hipGraph chunks), collisions on",
                                                    "steps": rest}}
    b32.close()
    b64.close()
    # config 3: differentiable FP32 vs the oracle's differentiable mode, 2000 steps
    p3 = configs.config3()
    d = vx3mi.Diff(vx3mi.Model.from_problem(p3), vx3mi.FP32)
    ph, E = np.asarray(p3["phase"]), np.asarray(p3["voxel_E"], dtype=np.float64)
    d.set_params(ph, E)
    d.forward(2000)
    od = O.OracleDiff(p3, ph, E, dt=d.info()["dt"])
    x0 = od.state()["pos"]
    od.step(2000)
    s, r = d.state(), od.state()
    out["c3_diff_fp32_vs_oracle_2000"] = {"e_pos": e_pos(s["pos"], r["pos"]), "e_q": e_quat(s["quat"], r["quat"]),
                                          "com": com_rel(s["pos"].mean(0), x0.mean(0), r["pos"].mean(0), x0.mean(0))}
    d.close()
    return out


def config1_only(args) -> int:
    """The 64x64x16 robot (45,800 voxels) with collisions, FP32, flat mode: warmup, then
    args.steps timed steps with HIP events on the batch stream."""
    m = vx3mi.Model.from_problem(configs.config1())
    b = vx3mi.Batch([m], vx3mi.FP32)
    b.step(args.warmup, mode=vx3mi.MODE_FLAT)
    b.sync()
    b.step(args.steps, mode=vx3mi.MODE_FLAT)
    b.sync()
    ms = b.last_step_ms()
    nv = b.nv
    out = {"metric": "voxel-steps/s", "value": nv * args.steps / (ms * 1e-3), "unit": "voxel-steps/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms / args.steps,
           "higher_is_better": True, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": "one 64x64x16 robot, collisions on, flat mode (hipGraph chunks)", "voxels": nv},
           "diverged": int(b.diverged()[0])}
    b.close()
    print(json.dumps(out), flush=True)
    return 0


def launch_roofline(mode, tile, threads, nv, nl, steps_per_launch, launch_s, args, collisions):
    """The dominant kernel's roofline.  The tiled kernel keeps the batch state in
    registers and LDS for a whole launch, so HBM does not bound it: its binding
    roof is VALU issue (achieved = VALU wave instructions per launch, counted,
    / launch duration, timed; peak VALU_ISSUE_PEAK_G).  Beside it: the bytes a
    step that keeps nothing on chip would move (algorithmic, per voxel 144 B +
    100 B per link) and the HBM bytes the launch moves (FETCH_SIZE +
    WRITE_SIZE, calibrated on a copy of known size)."""
    kernel = {vx3mi.MODE_TILED: "vx3mi::k_tile", vx3mi.MODE_PERSISTENT: "vx3mi::k_persistent"}.get(mode)
    bpv = algorithmic_bytes_per_voxel_step(nv, nl)
    alg = bpv * nv * steps_per_launch / launch_s / 1e9
    out = {"kernel": kernel, "steps_per_launch": steps_per_launch, "launch_ms": launch_s * 1e3,
           "bytes_per_voxel_step": bpv, "algorithmic_gbs": alg}
    cnt = None
    if kernel and not args.no_counters:
        child = ["--steps", str(int(steps_per_launch)), "--robots", str(args.robots), "--collisions", str(collisions),
                 "--child-mode", str(mode), "--child-tile", str(tile), "--child-threads", str(threads)]
        cnt = counter_run(kernel, child)
    if cnt is None:
        out.update(bound="valu-issue", achieved=None, peak=VALU_ISSUE_PEAK_G, unit="G VALU wave-instructions/s",
                   frac=None, traffic=None, note="counters not collected")
        return out
    rate = cnt["SQ_INSTS_VALU"] / launch_s / 1e9
    out.update(bound="valu-issue", achieved=rate, peak=VALU_ISSUE_PEAK_G, unit="G VALU wave-instructions/s",
               frac=rate / VALU_ISSUE_PEAK_G, traffic=cnt["traffic"],
               hbm={"achieved": cnt["traffic"] / launch_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": cnt["traffic"] / launch_s / 1e9 / HBM_PEAK_GBS},
               valu_per_wave_step=cnt["SQ_INSTS_VALU"] / cnt["SQ_WAVES"] / steps_per_launch,
               counters={k: cnt[k] for k in ("FETCH_SIZE", "WRITE_SIZE", "FETCH_SIZE_calibration",
                                             "WRITE_SIZE_calibration", "SQ_INSTS_VALU", "SQ_WAVES", "dispatches")},
               note="achieved: counted VALU wave instructions per launch / timed launch duration; peak: 256 CUs "
                    "x 4 SIMDs x 2.4 GHz / 4 cycles per wave64 instruction; traffic: HBM bytes per launch "
                    "(FETCH_SIZE + WRITE_SIZE x their copy calibration); algorithmic_gbs: the bytes of a step "
                    "that keeps no state on chip, which the tiled kernel beats")
    return out


def config1_traffic(step_s: float, steps: int = 100):
    """Config 1 (one 64x64x16 robot, collisions on, flat mode): HBM bytes per step
    of all its kernels (FETCH_SIZE + WRITE_SIZE over `steps` steps, calibrated on
    the copy), against the HBM peak at the timed step duration step_s."""
    child = ["--workload", "config1", "--steps", str(steps), "--warmup", "0"]
    per = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        rows = rocprof_counters([c], child)
        if rows is None:
            return None
        ks = kernel_rows(rows, "vx3mi::k_")
        f = copy_calibration(rows, c)
        if not ks or f is None:
            return None
        per[c] = sum(v[c] for v in ks) * f / steps
    traffic = per["FETCH_SIZE"] + per["WRITE_SIZE"]
    nv, nl = configs.counts(configs.config1())
    alg = algorithmic_bytes_per_voxel_step(nv, nl) * nv
    return {"bound": "hbm", "achieved": traffic / step_s / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": traffic / step_s / 1e9 / HBM_PEAK_GBS, "traffic_per_step": traffic,
            "read_per_step": per["FETCH_SIZE"], "written_per_step": per["WRITE_SIZE"],
            "algorithmic_per_step": alg, "algorithmic_gbs": alg / step_s / 1e9, "step_us": step_s * 1e6,
            "note": "achieved: measured HBM bytes per step (all kernels of the step) / timed step duration"}


def collisions_line(args, steps: int = 2000):
    """The batched workload with collisions on (the reference's default,
    VX3_SimulationManager.cu:328): autotuned launch configuration, timed steps,
    candidate-list rebuilds, and the dominant kernel's roofline."""
    probs = configs.config2(seed=0, n=args.robots, collisions=1)
    b = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
    tune = b.autotune(args.tune_steps)
    b.step(args.warmup, mode=vx3mi.MODE_AUTO)
    b.sync()
    l0 = b.listings().copy()
    b.step(steps, mode=vx3mi.MODE_AUTO)
    b.sync()
    ms, launches = b.last_step_ms(), b.last_step_launches()
    lst = b.listings() - l0
    out = {"value": b.nv * steps / (ms * 1e-3), "unit": "voxel-steps/s", "steps": steps,
           "ms_per_step": ms / steps, "voxels": b.nv, "autotune": tune,
           "listings_per_robot_per_1000_steps": float(lst.mean()) * 1000 / steps,
           "diverged": int(b.diverged().sum()),
           "workload": "the batched workload with collisions on (EnableCollision, the reference default)"}
    mode = tune["mode"]
    per = steps / max(1, launches) if mode in (vx3mi.MODE_TILED, vx3mi.MODE_PERSISTENT) else 1
    out["roofline"] = launch_roofline(mode, tune["tile"], tune["threads"], b.nv, b.nl, per,
                                      ms * 1e-3 / max(1, launches), args, collisions=1)
    b.close()
    return out


def adjoint_report(args) -> dict:
    """The adjoint's cost on the 8x8x8 differentiable robot (2000 steps) and on
    256 batched robots (5000 steps), against the differentiable forward and
    against the product timestep on the same robots."""
    # the reverse sweep (512 or 1024 threads per robot, or flat: every robot's
    # reverse step at once, one item per thread) is chosen by timing each on
    # each workload
    best = {}
    for name, probs, steps, K in (("c3", [configs.config3()], 2000, 100), ("c4", configs.config4(), 5000, 0)):
        for rt in (512, 1024, -1):
            c = adjoint_cost(probs, steps, K, args.adjoint_budget, reverse_threads=rt)
            log(f"bench: adjoint {name} reverse threads {rt}: forward {c['forward_ms']:.1f} ms, backward "
                f"{c['backward_ms']:.1f} ms, ratio {c['adjoint_over_forward']:.2f}")
            if name not in best or c["backward_ms"] < best[name]["backward_ms"]:
                best[name] = c
    c3, c4 = best["c3"], best["c4"]
    # the 256 robots again with a budget sized for the MI355X's 288 GB: the whole
    # run's tape (164 GB) stays in HBM, so the backward is the reverse sweep alone
    try:
        if args.adjoint_tape_budget <= 0:
            raise vx3mi.VX3MIError("--adjoint-tape-budget 0")
        c4w = adjoint_cost(configs.config4(), 5000, 0, args.adjoint_tape_budget, reverse_threads=-1)
        log(f"bench: adjoint c4 whole tape ({c4w['tape_bytes'] / 1e9:.0f} GB, budget "
            f"{args.adjoint_tape_budget / 1e9:.0f} GB): forward {c4w['forward_ms']:.1f} ms, backward "
            f"{c4w['backward_ms']:.1f} ms, ratio {c4w['adjoint_over_forward']:.2f}")
    except vx3mi.VX3MIError as e:  # not enough free device memory for the whole tape
        log(f"bench: adjoint c4 whole tape skipped: {e}")
        c4w = None
    # the same robots through the product timestep (autotuned, FP32, rigid floor):
    # the adjoint's cost per voxel-step against the fastest forward of the library
    for c, probs, steps in ((c3, [configs.config3()], 2000), (c4, configs.config4(), 5000)):  # (c4w: below)
        pb = vx3mi.Batch([vx3mi.Model.from_problem(p) for p in probs], vx3mi.FP32)
        pt = pb.autotune(50)
        pb.step(min(steps, 1000))
        pb.sync()
        pb.step(steps)
        pb.sync()
        fwd_ms = pb.last_step_ms()
        pb.close()
        c["product_forward_ms"] = fwd_ms
        c["product_forward_mode"] = pt["mode"]
        c["adjoint_over_product_forward"] = c["backward_ms"] / fwd_ms
        c["diff_forward_over_product_forward"] = c["forward_ms"] / fwd_ms
    log(f"bench: adjoint 8x8x8 2000 steps K=100: forward {c3['forward_ms']:.1f} ms, backward "
        f"{c3['backward_ms']:.1f} ms, ratio {c3['adjoint_over_forward']:.2f}")
    log(f"bench: batched adjoint 256 robots 5000 steps K={c4['checkpoint_every']}: forward {c4['forward_ms']:.1f} ms, "
        f"backward {c4['backward_ms']:.1f} ms, ratio {c4['adjoint_over_forward']:.2f}")
    for name, c in (("8x8x8", c3), ("256 robots", c4)):
        log(f"bench: adjoint {name}: product forward {c['product_forward_ms']:.1f} ms (mode "
            f"{c['product_forward_mode']}), backward / product forward {c['adjoint_over_product_forward']:.2f}")
    if c4w is not None:
        c4w["product_forward_ms"] = c4["product_forward_ms"]
        c4w["adjoint_over_product_forward"] = c4w["backward_ms"] / c4["product_forward_ms"]
    return {"single_8x8x8": c3, "batched_256": c4, "batched_256_whole_tape": c4w,
            "note": "backward = recompute from checkpoints + reverse sweep; FP32; softplus floor; collisions off; "
                    "adjoint_over_forward: against the differentiable forward (checkpoints, softplus floor, "
                    "per-voxel moduli); adjoint_over_product_forward: against the product timestep on the "
                    "same robots (autotuned launch configuration)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=1000)
    ap.add_argument("--robots", type=int, default=1024)
    ap.add_argument("--plain-steps", type=int, default=20)
    ap.add_argument("--cpu-robots", type=int, default=48)
    ap.add_argument("--cpu-steps", type=int, default=1000)
    ap.add_argument("--check-robots", type=int, default=16)
    ap.add_argument("--check-steps", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--mode", choices=["auto", "tiled", "persistent", "flat"], default="auto",
                    help="auto: the batch autotuner times every launch configuration and keeps the fastest")
    ap.add_argument("--tune-steps", type=int, default=200)
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-parity", action="store_true")
    ap.add_argument("--no-counters", action="store_true",
                    help="skip the rocprofv3 counter passes (HBM bytes, VALU instructions of the dominant kernel)")
    ap.add_argument("--no-collisions-line", action="store_true")
    ap.add_argument("--counters-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--collisions", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--child-mode", type=int, default=vx3mi.MODE_TILED, help=argparse.SUPPRESS)
    ap.add_argument("--child-tile", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--child-threads", type=int, default=0, help=argparse.SUPPRESS)
    ap.add_argument("--adjoint-budget", type=float, default=16.0 * 2**30)
    ap.add_argument("--adjoint-tape-budget", type=float, default=192e9,
                    help="HBM budget of the whole-tape adjoint line (0: skip it)")
    ap.add_argument("--workload", choices=["batched", "config1", "adjoint"], default="batched",
                    help="config1: only the 64x64x16 robot with collisions (flat mode), for profiling the "
                         "collision kernels; adjoint: only the adjoint cost lines; each prints its own line")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); robots shard across GPUs "
              "as independent batches, which this build does not launch.", file=sys.stderr)
        return 2
    if args.counters_child:
        return counters_child(args)
    if args.workload == "config1":
        return config1_only(args)
    if args.workload == "adjoint":
        print(json.dumps({"workload": "adjoint", "adjoint": adjoint_report(args)}), flush=True)
        return 0

    mode = {"auto": vx3mi.MODE_AUTO, "tiled": vx3mi.MODE_TILED, "persistent": vx3mi.MODE_PERSISTENT,
            "flat": vx3mi.MODE_FLAT}[args.mode]
    problems = configs.config2(seed=0, n=args.robots)
    t0 = time.perf_counter()
    models = [vx3mi.Model.from_problem(p) for p in problems]
    batch = vx3mi.Batch(models, vx3mi.FP32)
    nv, nl = batch.nv, batch.nl
    log(f"bench: {args.robots} robots, {nv} voxels, {nl} links, setup {time.perf_counter() - t0:.2f}s")

    tune = None
    if mode == vx3mi.MODE_AUTO:
        tune = batch.autotune(args.tune_steps)
        chosen = tune["mode"]
        log(f"bench: autotune ({args.tune_steps} steps each): {tune['ms']} -> mode {chosen} tile {tune['tile']} "
            f"threads {tune['threads']}")
    else:
        chosen = mode
    # warmup
    batch.step(args.warmup, mode=mode)
    batch.sync()
    # timed region: inputs resident in HBM, device events around the launches
    t0 = time.perf_counter()
    batch.step(args.steps, mode=mode)
    batch.sync()
    wall = time.perf_counter() - t0
    dev_ms = batch.last_step_ms()
    launches = batch.last_step_launches()
    div = int(batch.diverged().sum())
    value = nv * args.steps / (dev_ms * 1e-3)
    log(f"bench: batched {args.steps} steps: device {dev_ms:.1f} ms, wall {wall * 1e3:.1f} ms, "
        f"{launches} launches, {value:.3e} voxel-steps/s, diverged robots {div}")

    # plain path: one robot per launch sequence
    plain = vx3mi.Batch(models, vx3mi.FP32)
    plain.step(2, mode=vx3mi.MODE_PLAIN)
    plain.sync()
    plain.step(args.plain_steps, mode=vx3mi.MODE_PLAIN)
    plain.sync()
    plain_ms = plain.last_step_ms()
    plain_value = nv * args.plain_steps / (plain_ms * 1e-3)
    plain.close()
    log(f"bench: plain {args.plain_steps} steps: {plain_ms:.1f} ms, {plain_value:.3e} voxel-steps/s")

    # roofline of the dominant kernel: its launch duration from the HIP events
    # above; its VALU instructions and HBM bytes per launch from counter passes
    # over launches of the same shape in a child process
    tile = tune["tile"] if tune else 0
    threads = tune["threads"] if tune else 256
    steps_per_launch = args.steps / max(1, launches)
    avg_launch_s = dev_ms * 1e-3 / max(1, launches)
    roofline = launch_roofline(chosen, tile, threads, nv, nl, steps_per_launch, avg_launch_s, args, collisions=0)
    log(f"bench: roofline {roofline}")

    collisions_on = None
    if not args.no_collisions_line:
        collisions_on = collisions_line(args)
        log(f"bench: batched with collisions on: {collisions_on}")

    err = None
    if args.check_robots > 0:
        err = fp32_error_vs_oracle(problems[: args.check_robots], args.check_steps)
        log(f"bench: FP32 vs FP64 oracle, {args.check_robots} robots x {args.check_steps} steps: e_pos {err:.3e}")

    cpu = None
    if not args.no_cpu:
        sample = problems[: args.cpu_robots]
        cv, kind = cpu_baseline(sample, args.cpu_steps)
        svox = sum(int((np.asarray(p["cells"]) > 0).sum()) for p in sample)
        cpu = {"value": cv, "unit": "voxel-steps/s", "cores": 1, "kind": kind,
               "sample": f"first {len(sample)} robots of the workload ({svox} voxels) x {args.cpu_steps} steps, "
                         "single thread", "cpu_model": cpu_model()}
        log(f"bench: CPU baseline ({kind}) {cv:.3e} voxel-steps/s")
        osteps = max(1, args.cpu_steps // 5)
        ov, _ = cpu_baseline(sample, osteps, omp=True)
        if ov is not None:
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            cpu["all_threads"] = {"value": ov, "threads": threads,
                                  "note": f"reference built with -DUSE_OMP -fopenmp, same robots x {osteps} steps"}
            log(f"bench: CPU baseline (reference, OpenMP, {threads} threads) {ov:.3e} voxel-steps/s")

    parity = None
    if not args.no_parity:
        t0 = time.perf_counter()
        parity = parity_report()
        log(f"bench: parity per config ({time.perf_counter() - t0:.1f}s): {parity}")
        c1 = parity["c1_fp32_vs_fp64"]
        if not args.no_counters:
            step_s = c1["voxels"] / c1["voxel_steps_per_s"]["fp32"]
            c1["roofline"] = config1_traffic(step_s)
            log(f"bench: config 1 HBM traffic {c1['roofline']}")

    adjoint = None if args.no_adjoint else adjoint_report(args)

    out = {
        "metric": "voxel-steps/s batched vs plain @<=1e-5 rel err to FP64; adjoint/forward cost",
        "value": value,
        "unit": "voxel-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dev_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded random robots)",
        "config": {"workload": f"batched {args.robots} robots, dims U{{4..12}}^3, p=0.6, largest component, "
                               "2 materials, random phases, FP32, collisions off",
                   "robots": args.robots, "voxels": nv, "links": nl, "parallelism": "single GPU"},
        "plain_value": plain_value,
        "batched_over_plain": value / plain_value,
        "fp32_rel_err_vs_fp64": err,
        "roofline": roofline,
        "collisions_on": collisions_on,
        "cpu_baseline": cpu,
        "adjoint": adjoint,
        "autotune": tune,
        "parity": parity,
    }
    print(json.dumps(out), flush=True)
    batch.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
