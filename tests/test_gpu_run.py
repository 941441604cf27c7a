"""The simulation driver (CUDA_Simulation loop) and the command-line drop-in on
the GPU: stop formula, centre-of-mass and fitness at the stop, .history frames,
robots stopping at different steps inside one batch, and the report XML.

FP64 batches reproduce the oracle bit for bit, so final centres of mass are
compared with the oracle run for the same number of steps at 1e-12.
"""
import math
import os
import re
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

import oracle as O
import vx3mi
from vx3mi import configs
from vx3mi import vxa as W

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def log(msg):
    print(msg, flush=True)


def steps_until(dt, t_stop):
    """Iterations the reference runs with stop formula t - t_stop (> 0 stops)."""
    t, n = 0.0, 0
    while not (t - t_stop > 0):
        t += float(np.float32(dt))
        n += 1
    return n, t


def small_robot(seed, dims=(3, 3, 2)):
    return configs.random_robot(np.random.default_rng(seed), dims, 0.8, enable_collision=0)


def test_stop_formula_fitness_and_history():
    prob = small_robot(1)
    t_stop = 0.004
    text = W.write_vxa(prob, stop=("mtSUB", ("mtVAR", "t"), ("mtCONST", t_stop)),
                       fitness=("mtVAR", "x"), record=(5, 1, 1))
    m = vx3mi.Model.from_xml(text)
    b = vx3mi.Batch([m], vx3mi.FP64)
    log("runner: single robot FP64")
    res, hist, finals = vx3mi.run(b, [m])
    r = res[0]
    dt = m.settings()["dt_frac"] * m.recommended_dt()
    n, t = steps_until(dt, t_stop)
    assert r["steps"] == n and r["current_time"] == t and not r["diverged"]
    o = O.OracleModel(prob)
    c0 = o.com()
    o.step(n)
    assert np.allclose(r["initial_com"], c0, rtol=0, atol=1e-15)
    assert np.allclose(r["current_com"], o.com(), rtol=1e-12, atol=0)
    assert r["fitness_score"] == pytest.approx(o.com()[0] - c0[0], rel=1e-12)
    assert np.allclose(finals[0], o.state()["pos"], rtol=1e-12, atol=0)
    # frames every real_stepsize iterations, positions in mm with one decimal
    real = int(5 / (10000 * m.recommended_dt() * m.settings()["dt_frac"])) + 1
    assert r["real_stepsize"] == real
    frames = re.findall(r"<<<Step(\d+) Time:([0-9.]+)>>>(.*?)<<<>>>", hist[0])
    assert [int(f[0]) for f in frames] == list(range(0, n, real))
    assert hist[0].startswith("\n{{{setting}}}<rescale>0.001</rescale>\n{{{setting}}}<matcolor><id>1</id>")
    # the last frame's first surface voxel against the oracle state after that iteration
    last = int(frames[-1][0])
    o2 = O.OracleModel(prob)
    o2.step(last + 1)
    v = int(o2.topology()["surface"][0])
    p = o2.state()["pos"][v] * 1000
    first = frames[-1][2].split(";")[0].split(",")
    assert first[:3] == ["%.1f" % p[0], "%.1f" % p[1], "%.1f" % p[2]]
    assert "|[[[" in hist[0] and "[[[]]]" in hist[0]


def render_frame(prob, ref, j, t, L=0.01):
    """A .history frame (VX3_SimulationManager.cu:61-111) rendered from the oracle's
    state with the reference's printf formats and its cornerOffset
    (VX3_Voxel.cu:147-159: per axis (1 + axialStrain(posLink)) of the link on
    that side, float; half the thermally scaled size), for a robot without
    failing links, signals or removal."""
    vs = 1 / 0.001
    st = ref.state()
    topo = ref.topology()
    temp, _ = ref.temperature()
    _, strain = ref.link_state()
    pal = np.asarray(prob["cells"]).reshape(-1)
    pal = pal[pal > 0]
    E = np.array([m.get("E", 1e6) for m in prob["materials"]], dtype=np.float32)
    cte = np.array([m.get("cte", 0.0) for m in prob["materials"]], dtype=np.float32)
    slot = {}
    for l, (n, p_, a) in enumerate(zip(topo["link_neg"], topo["link_pos"], topo["link_axis"])):
        slot[(int(n), 2 * int(a))] = l
        slot[(int(p_), 2 * int(a) + 1)] = l
    out = ["<<<Step%d Time:%f>>>" % (j, t)]
    for v in topo["surface"]:
        v = int(v)
        p = st["pos"][v]
        q = st["quat"][v]
        out.append("%.1f,%.1f,%.1f," % (p[0] * vs, p[1] * vs, p[2] * vs))
        w = q[0]
        out.append("%.1f,%.2f,%.2f,%.2f," % (2.0 * math.acos(1 if w > 1 else w) * 57.29577951308232, q[1], q[2], q[3]))
        m = pal[v] - 1
        factor = np.float32(np.float32(1) + np.float32(temp[v]) * cte[m])
        half = 0.5 * (L * float(factor))
        offs = []
        for pos_link in (False, True):
            for i in range(3):
                l = slot.get((v, 2 * i + (0 if pos_link else 1)))
                if l is None:
                    s_ = 1.0 if pos_link else -1.0
                else:
                    ratio = np.float32(E[pal[topo["link_pos"][l]] - 1] / E[pal[topo["link_neg"][l]] - 1])
                    two = np.float32(2.0)
                    a = (two * strain[l] * ratio / (np.float32(1) + ratio)) if pos_link else (two * strain[l] / (np.float32(1) + ratio))
                    s_ = float(np.float32(np.float32(1) + np.float32(a))) * (1 if pos_link else -1)
                offs.append(float(np.float32(half * s_)))
        out.append("%.1f,%.1f,%.1f,%.1f,%.1f,%.1f," % tuple(o * vs for o in offs))
        out.append("%d," % pal[v])
        out.append("%.1f," % 0.0)
        out.append(";")
    out.append("<<<>>>")
    out.append("|[[[%d]]]" % j)
    for n, p_ in zip(topo["link_neg"], topo["link_pos"]):
        a, b = st["pos"][p_], st["pos"][n]
        out.append("%.4f,%.4f,%.4f," % tuple(a) + "%.4f,%.4f,%.4f," % tuple(b) + ";")
    out.append("[[[]]]")
    return "".join(out)


def test_history_frames_byte_for_byte():
    """Every frame of an FP64 run (voxels and links) equals the frame rendered
    from the oracle at that iteration with the reference's formats."""
    prob = dict(configs.BASE, dims=(3, 3, 2), cells=np.array([1, 2, 1, 2, 1, 2, 1, 2, 1, 1, 1, 2, 2, 1, 2, 1, 2, 1],
                                                          dtype=np.int8), phase=None, enable_collision=0,
                materials=[dict(configs.SOFT, E=1e6, cte=0.01), dict(configs.SOFT, E=1e7, cte=0.01)])
    m0 = vx3mi.Model.from_problem(prob)
    dt = float(np.float32(m0.settings()["dt_frac"] * m0.recommended_dt()))
    text = W.write_vxa(prob, stop=("mtSUB", ("mtVAR", "t"), ("mtCONST", 40.5 * dt)), record=(2, 1, 1))
    m = vx3mi.Model.from_xml(text)
    b = vx3mi.Batch([m], vx3mi.FP64)
    log("runner: history bytes")
    res, hist, _ = vx3mi.run(b, [m])
    real = res[0]["real_stepsize"]
    lines = [x for x in hist[0].split("\n") if x.startswith("<<<Step")]
    assert len(lines) >= 3
    ref = O.OracleModel(prob)
    done, t = 0, 0.0
    for line in lines:
        j = int(line[len("<<<Step"):line.index(" ")])
        assert j % real == 0
        while done < j + 1:
            ref.step(1)
            done += 1
        assert line == render_frame(prob, ref, j, ref.time), f"frame {j}"
    b.close()


def test_robots_stop_independently_in_one_batch():
    """Three robots with different stop times run in one batch; each one's
    results equal those of the robot run alone."""
    probs = [small_robot(s) for s in (2, 3, 4)]
    stops = [0.002, 0.0035, 0.0011]
    models = [vx3mi.Model.from_xml(W.write_vxa(p, stop=("mtSUB", ("mtVAR", "t"), ("mtCONST", ts)),
                                               fitness=("mtVAR", "x")))
              for p, ts in zip(probs, stops)]
    log("runner: three robots in one batch")
    res, _, fin = vx3mi.run(vx3mi.Batch(models, vx3mi.FP64), models, history=False)
    for k, mdl in enumerate(models):
        alone, _, fa = vx3mi.run(vx3mi.Batch([mdl], vx3mi.FP64), [mdl], history=False)
        assert res[k]["steps"] == alone[0]["steps"]
        assert res[k]["current_com"] == alone[0]["current_com"]
        assert res[k]["fitness_score"] == alone[0]["fitness_score"]
        assert np.array_equal(fin[k], fa[0])


def test_max_steps_bound_without_stop_formula():
    prob = small_robot(5)
    m = vx3mi.Model.from_xml(W.write_vxa(prob))
    res, hist, _ = vx3mi.run(vx3mi.Batch([m], vx3mi.FP32), [m], max_steps=37)
    assert res[0]["steps"] == 37 and hist[0] == ""


def test_cli_report(tmp_path):
    exe = os.path.join(ROOT, "voxcraft-sim_amd", "lib", "vx3mi_run")
    if not os.path.exists(exe):
        pytest.fail("vx3mi_run not built (make -C voxcraft-sim_amd cli)")
    d = tmp_path / "gen"
    d.mkdir()
    probs = [small_robot(s) for s in (6, 7)]
    base, _ = W.write_pair(probs[0], stop=("mtSUB", ("mtVAR", "t"), ("mtCONST", 0.002)), fitness=("mtVAR", "x"),
                           record=(5, 1, 0))
    (d / "base.vxa").write_text(base)
    for k, p in enumerate(probs):
        (d / f"robot{k}.vxd").write_text(W.write_pair(p, record=(5, 1, 0))[1])
    out = tmp_path / "report.xml"
    log("runner: command line")
    r = subprocess.run([exe, "-i", str(d), "-o", str(out)], capture_output=True, timeout=120)
    assert r.returncode == 0, r.stdout.decode() + r.stderr.decode()
    stdout = r.stdout.decode()
    assert stdout.count("<<<Step0 Time:") == 2 and "real_stepsize:" in stdout
    rep = ET.parse(out).getroot()
    assert rep.tag == "report" and rep.find("inputdir").text == "gen"
    details = list(rep.find("detail"))
    assert sorted(e.tag for e in details) == ["robot0", "robot1"]
    fits = [float(e.find("fitness_score").text) for e in details]
    assert fits == sorted(fits, reverse=True)
    assert rep.find("bestfit/fitness_score").text == details[0].find("fitness_score").text
    for e in details:
        c0 = [float(e.find(f"initialCenterOfMass/{a}").text) for a in "xyz"]
        c1 = [float(e.find(f"currentCenterOfMass/{a}").text) for a in "xyz"]
        assert float(e.find("fitness_score").text) == pytest.approx(c1[0] - c0[0], rel=1e-6, abs=1e-12)
        assert float(e.find("currentTime").text) > 0.002
    # the same generation through a vx3_node_worker task file (Boost ptree XML
    # with several top-level keys), robots in the file's order
    vxt = tmp_path / "job.vxt"
    vxt.write_text('<?xml version="1.0" encoding="utf-8"?>\n'
                   f'<vxa>{d / "base.vxa"}</vxa><input_dir>{d}</input_dir><vxd><f>robot1.vxd</f><f>robot0.vxd</f></vxd>')
    out2 = tmp_path / "report2.xml"
    r2 = subprocess.run([exe, "-i", str(vxt), "-o", str(out2)], capture_output=True, timeout=120)
    assert r2.returncode == 0, r2.stdout.decode() + r2.stderr.decode()
    assert "Simulation 0 runs: robot1.vxd" in r2.stdout.decode()
    assert ET.tostring(ET.parse(out2).getroot()) == ET.tostring(rep)


def test_history_frames_follow_removal():
    """.history with SecondaryExperiment removal: the frame of the step that
    removes the x = 0 face of a 3x3x3 cube lists the old surface minus the
    removed voxels; later frames list the regenerated surface (the former
    centre included); removed links are left out from that frame on."""
    n = 27
    cells = np.ones(n, dtype=np.int8)
    cells.reshape(3, 3, 3)[:, :, 0] = 2
    prob = dict(configs.BASE, dims=(3, 3, 3), cells=cells, phase=None, enable_collision=0, vary_temp=0,
                secondary_experiment=1, materials=[configs.SOFT, dict(configs.SOFT, remove_after=0.0)])
    m0 = vx3mi.Model.from_problem(prob)
    dt = float(np.float32(m0.settings()["dt_frac"] * m0.recommended_dt()))
    rs = int(1 / (10000 * m0.recommended_dt() * m0.settings()["dt_frac"])) + 1  # real_stepsize of RecordStepSize 1
    k = 4 * rs + 1  # removal at the end of step k (1-based): frame index k - 1 shows it
    t = [0.0]
    for _ in range(k):
        t.append(t[-1] + dt)
    prob["materials"][1]["remove_after"] = 0.5 * (t[k - 2] + t[k - 1])
    text = W.write_vxa(prob, stop=("mtSUB", ("mtVAR", "t"), ("mtCONST", t[k] + 6 * rs * dt)), record=(1, 1, 1))
    m = vx3mi.Model.from_xml(text)
    b = vx3mi.Batch([m], vx3mi.FP64)
    log("runner: history with removal")
    res, hist, _ = vx3mi.run(b, [m])
    assert res[0]["real_stepsize"] == rs
    vr, lr = b.removed(0)
    assert (vr == k).sum() == 9 and (lr == k).sum() == 21
    frames = re.findall(r"<<<Step(\d+) Time:[^>]*>>>(.*?)<<<>>>\|\[\[\[\d+\]\]\](.*?)\[\[\[\]\]\]", hist[0])
    assert len(frames) > 6
    seen = set()
    for j, vox, links in frames:
        j = int(j)
        nvox, nlink = vox.count(";"), links.count(";")
        if j < k - 1:
            want = (26, 54)
        elif j == k - 1:
            want = (17, 33)
        else:
            want = (18, 33)
        assert (nvox, nlink) == want, j
        seen.add(want)
    assert seen == {(26, 54), (17, 33), (18, 33)}
    b.close()


def test_history_frames_list_attached_links():
    """.history with attach: the links attach makes join the frame's link list
    (after the lattice links) from the frame of the step that makes them."""
    from test_attach_oracle import run_until_links, stacked_cubes
    prob = stacked_cubes()
    ref = O.OracleModel(prob)
    k, _ = run_until_links(ref, 3000)  # the links come in step k + 1 (1-based): frame index k
    m0 = vx3mi.Model.from_problem(prob)
    dt = float(np.float32(m0.settings()["dt_frac"] * m0.recommended_dt()))
    text = W.write_vxa(prob, stop=("mtSUB", ("mtVAR", "t"), ("mtCONST", (k + 60) * dt)), record=(1, 1, 1))
    m = vx3mi.Model.from_xml(text)
    assert m.settings()["enable_attach"] == 1
    b = vx3mi.Batch([m], vx3mi.FP64)
    log("runner: history with attach")
    res, hist, _ = vx3mi.run(b, [m])
    frames = re.findall(r"<<<Step(\d+) Time:[^>]*>>>(.*?)<<<>>>\|\[\[\[\d+\]\]\](.*?)\[\[\[\]\]\]", hist[0])
    assert len(frames) > 3
    counts = {(int(j) >= k, vox.count(";"), links.count(";")) for j, vox, links in frames}
    assert counts == {(False, 16, 24), (True, 16, 28)}
    assert len(b.attach_state(0)["links"]) == 4
    b.close()


def test_runner_hits_and_target_closeness():
    """Stop formula on hit (collisionCount): the run stops at the first
    iteration whose count exceeds 20; fitness targetCloseness and the report's
    numClosePairs come from the last CycleStep sample (sum of 1 / distance over
    target pairs, pairs closer than MaxDist voxel lengths), as the oracle's
    positions at that step give them."""
    prob = dict(configs.BASE, dims=(4, 1, 1), cells=np.array([1, 0, 2, 2], dtype=np.int8), phase=None,
                enable_collision=1, vary_temp=0, temp_period=0.002, floor_enabled=0, grav_acc=0.0,
                max_dist_pairs=1.5, materials=[dict(configs.SOFT), dict(configs.SOFT, is_target=1)],
                externals=[dict(voxel=0, force=(0.2, 0, 0)), dict(voxel=1, force=(-0.2, 0, 0))], slow_damp_z=0.0)
    m = vx3mi.Model.from_problem(prob)
    text = W.write_vxa(prob, stop=("mtSUB", ("mtVAR", "hit"), ("mtCONST", 20.0)), fitness=("mtVAR", "targetCloseness"))
    mx = vx3mi.Model.from_xml(text)
    for mm in (m, mx):
        for e in prob["externals"]:
            mm.set_external(**e)
    b = vx3mi.Batch([mx], vx3mi.FP64)
    log("runner: hits and target closeness")
    res, _, _ = vx3mi.run(b, [mx])
    r = res[0]
    o = O.OracleModel(prob)
    dt = float(np.float32(mx.settings()["dt_frac"] * mx.recommended_dt()))
    cycle = int(0.002 / dt)
    n, last = 0, None
    while o.collision_count() <= 20:
        o.step(1)
        n += 1
        if n % cycle == 0:
            p = o.state()["pos"]
            d = np.linalg.norm(p[1] - p[2])
            last = (1 / d, int(d < 1.5 * 0.01))
        assert n < 5000
    assert r["steps"] == n and not r["diverged"]
    assert last is not None
    assert r["fitness_score"] == last[0] and r["num_close_pairs"] == last[1]
    b.close()
