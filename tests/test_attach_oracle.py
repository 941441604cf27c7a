"""Attach in the FP64 oracle (handle_collision_attachment, VX3_VoxelyzeKernel.cu:729-831).

The checks here do not restate the oracle's formulas: the step at which links
appear is tied to the contact envelope measured from positions, the depth-5
neighbour rule to a breadth-first search over the oracle's own topology, the
new-link force scaling to the same link's forces without it, and the pruned
pair search to the literal all-pairs loop.
"""
from collections import deque

import numpy as np
import pytest

import oracle as O

L0 = 0.01
ENVELOPE = 2 * 0.625 * L0  # (baseSizeAverage1 + baseSizeAverage2) * COLLISION_ENVELOPE_RADIUS at rest size


def stacked_cubes(**kw):
    """Two 2x2x2 cubes one above the other with a one-voxel gap; the upper one
    falls onto the lower (no global damping)."""
    cells = np.zeros((5, 2, 2), dtype=np.int8)  # z, y, x
    cells[0:2] = 1
    cells[3:5] = 1
    mat = dict(E=1e6, rho=1000, sticky=1)
    mat.update(kw.pop("mat", {}))
    p = dict(dims=(2, 2, 5), cells=cells.reshape(-1), materials=[mat], enable_attach=1, lattice_dim=L0,
             slow_damp_z=0.0)
    p.update(kw)
    return p


def u_chain(w, h):
    """A one-voxel-thick U in the xy plane: columns x = 0 and x = w (y = 0..h)
    joined by the row y = 0.  The tips are (0, h) and (w, h), h + w + h links
    apart."""
    cells = np.zeros((1, h + 1, w + 1), dtype=np.int8)
    cells[0, :, 0] = 1
    cells[0, :, w] = 1
    cells[0, 0, :] = 1
    return dict(dims=(w + 1, h + 1, 1), cells=cells.reshape(-1), materials=[dict(E=1e6, rho=1000, sticky=1)],
                enable_attach=1, lattice_dim=L0, floor_enabled=0, grav_acc=0.0, slow_damp_z=0.0)


def new_links(m, nl0):
    t = m.topology()
    return np.stack([t["link_neg"][nl0:], t["link_pos"][nl0:], t["link_axis"][nl0:]], axis=1)


def check_slots(m):
    """Every voxel slot holds at most one link: pVNeg takes the POS slot of the
    axis, pVPos the NEG slot."""
    t = m.topology()
    removed = m.removed()[1]
    used = set()
    for l, (n, p, a) in enumerate(zip(t["link_neg"], t["link_pos"], t["link_axis"])):
        if removed[l]:
            continue
        for key in ((int(n), 2 * int(a)), (int(p), 2 * int(a) + 1)):
            assert key not in used, f"slot {key} taken twice"
            used.add(key)


def graph_distance(m, a, b):
    t = m.topology()
    removed = m.removed()[1]
    adj = [[] for _ in range(m.nv)]
    for l, (n, p) in enumerate(zip(t["link_neg"], t["link_pos"])):
        if not removed[l]:
            adj[n].append(p)
            adj[p].append(n)
    dist = {a: 0}
    q = deque([a])
    while q:
        x = q.popleft()
        for y in adj[x]:
            if y not in dist:
                dist[y] = dist[x] + 1
                q.append(y)
    return dist.get(b, np.inf)


def run_until_links(m, max_steps, flags=0):
    """Steps one at a time; returns (step index, positions entering that step)
    of the first step that made links, or (None, None)."""
    nl0 = m.nl
    for k in range(max_steps):
        pos = m.state()["pos"].copy()
        m.step(1, flags=flags)
        if m.nl != nl0:
            return k, pos
    return None, None


def test_falling_cube_attaches_face_to_face_when_the_envelopes_touch():
    m = O.OracleModel(stacked_cubes())
    nl0 = m.nl
    k, pos = run_until_links(m, 3000)
    assert k is not None
    links = new_links(m, nl0)
    # the four voxels of the lower cube's top face and those of the upper cube's
    # bottom face, vertically aligned: Z links, the lower voxel the negative end
    lat = m.topology()["lattice"]
    assert len(links) == 4
    for n, p, a in links:
        assert a == 2
        assert lat[n][2] == 1 and lat[p][2] == 3
        assert lat[n][0] == lat[p][0] and lat[n][1] == lat[p][1]
    # made in the step whose entering positions first bring the pairs inside the envelope
    gaps = [np.linalg.norm(pos[p] - pos[n]) for n, p, _ in links]
    assert max(gaps) <= ENVELOPE
    m2 = O.OracleModel(stacked_cubes())
    m2.step(k)
    before = m2.state()["pos"]
    assert all(np.linalg.norm(before[p] - before[n]) == np.linalg.norm(pos[p] - pos[n]) for n, p, _ in links)
    m3 = O.OracleModel(stacked_cubes())
    m3.step(max(k - 1, 0))
    earlier = m3.state()["pos"]
    assert min(np.linalg.norm(earlier[p] - earlier[n]) for n, p, _ in links) > ENVELOPE
    # each new link starts with SafetyGuard updates at 0.01 force
    flags, new_left = m.attach_state()
    assert list(new_left[nl0:]) == [500] * 4
    check_slots(m)


def test_attached_pairs_give_their_contact_force_back():
    """In the step that makes the links the interface pairs' contact forces are
    added and given back, and the new links carry no force yet: the step moves
    the voxels exactly as one without any contact does, unlike one where the
    pairs stay in contact."""
    m = O.OracleModel(stacked_cubes())
    k, _ = run_until_links(m, 3000)
    no_contact = O.OracleModel(stacked_cubes(enable_collision=0, mat=dict(sticky=0)))
    no_contact.step(k + 1)
    contact = O.OracleModel(stacked_cubes(mat=dict(sticky=0)))
    contact.step(k + 1)
    s, s0, s1 = m.state(), no_contact.state(), contact.state()
    for key in ("pos", "linmom"):
        np.testing.assert_array_equal(s[key], s0[key])
    assert np.abs(s1["linmom"] - s0["linmom"]).max() > 0


@pytest.mark.parametrize("case", ["attach_off", "not_sticky", "fixed", "condition_false"])
def test_attach_needs_the_switch_a_sticky_free_material_and_its_conditions(case):
    kw = dict(attach_off=dict(enable_attach=0), not_sticky=dict(mat=dict(sticky=0)),
              fixed=dict(mat=dict(fixed=1)),
              condition_false=dict(attach_conditions=[None, ("mtGREATERTHAN", ("mtVAR", "z"), ("mtCONST", 1.0))]))[case]
    m = O.OracleModel(stacked_cubes(**kw))
    nl0 = m.nl
    m.step(1500)
    assert m.nl == nl0


def test_attach_condition_on_time_delays_the_links():
    # Condition_0: t > t_on.  A voxel's flag is evaluated after its update with
    # the step's start time, so links can come at the earliest in the step after
    # the first one that starts beyond t_on
    free = O.OracleModel(stacked_cubes())
    k_free, _ = run_until_links(free, 3000)
    dt = free.step_dt()
    t_on = (k_free + 200) * dt
    m = O.OracleModel(stacked_cubes(attach_conditions=[("mtGREATERTHAN", ("mtVAR", "t"), ("mtCONST", t_on))]))
    k, _ = run_until_links(m, 4000)
    assert k is not None
    starts = np.arange(k + 1) * np.float64(dt)
    first_true = int(np.argmax(starts > t_on))
    assert k >= first_true + 1
    flags, _ = m.attach_state()
    assert flags.all()


def test_safety_guard_scales_a_new_links_first_updates_by_a_hundredth():
    a = O.OracleModel(stacked_cubes(safety_guard=500))
    b = O.OracleModel(stacked_cubes(safety_guard=0))
    ka, _ = run_until_links(a, 3000)
    kb, _ = run_until_links(b, 3000)
    assert ka == kb
    nl0 = a.nl - 4
    # up to here both runs are the same (a new link carries zero force in the
    # step that makes it); the next link update differs by the factor only
    np.testing.assert_array_equal(a.state()["pos"], b.state()["pos"])
    fa, fb = a.update_links(), b.update_links()
    np.testing.assert_array_equal(fa[:nl0], fb[:nl0])
    np.testing.assert_array_equal(fa[nl0:], fb[nl0:] * 0.01)
    assert np.abs(fb[nl0:]).max() > 0
    assert list(a.attach_state()[1][nl0:]) == [499] * 4
    assert list(b.attach_state()[1][nl0:]) == [0] * 4


@pytest.mark.parametrize("w,h", [(2, 1), (3, 1), (2, 2), (3, 2)])
def test_touching_tips_attach_only_beyond_five_links(w, h):
    """is_neighbor(voxel1, voxel2, NULL, 5) (:651-680): the tips of a U touch
    after the right one is moved next to the left one; they attach exactly when
    more than five links separate them."""
    m = O.OracleModel(u_chain(w, h))
    lat = m.topology()["lattice"]
    tip1 = int(np.flatnonzero((lat[:, 0] == 0) & (lat[:, 1] == h))[0])
    tip2 = int(np.flatnonzero((lat[:, 0] == w) & (lat[:, 1] == h))[0])
    s = m.state()
    pos = s["pos"].copy()
    pos[tip2] = pos[tip1] + np.array([L0, 0.0, 0.0])
    m.set_state(pos=pos)
    d = graph_distance(m, tip1, tip2)
    assert d == w + 2 * h
    nl0 = m.nl
    m.step(1)
    links = new_links(m, nl0)
    if d > 5:
        # voxel1 = tip2 (the larger index) sees tip1 at -x: its X_NEG slot, so it
        # is the positive end
        assert links.tolist() == [[tip1, tip2, 0]]
    else:
        assert len(links) == 0
    check_slots(m)


@pytest.mark.parametrize("problem", ["cubes", "u"])
def test_pruned_pair_search_equals_the_all_pairs_loop_with_attach(problem):
    def make():
        if problem == "cubes":
            return O.OracleModel(stacked_cubes())
        m = O.OracleModel(u_chain(3, 2))
        lat = m.topology()["lattice"]
        t1 = int(np.flatnonzero((lat[:, 0] == 0) & (lat[:, 1] == 2))[0])
        t2 = int(np.flatnonzero((lat[:, 0] == 3) & (lat[:, 1] == 2))[0])
        pos = m.state()["pos"].copy()
        pos[t2] = pos[t1] + np.array([L0, 0.0, 0.0])
        m.set_state(pos=pos)
        return m

    a, b = make(), make()
    n = 1100 if problem == "cubes" else 300
    a.step(n)
    b.step(n, flags=O.VXO_ALL_PAIRS)
    assert a.nl == b.nl and a.nl > (24 if problem == "cubes" else 0)
    np.testing.assert_array_equal(a.state()["pos"], b.state()["pos"])
    ta, tb = a.topology(), b.topology()
    for key in ("link_neg", "link_pos", "link_axis"):
        np.testing.assert_array_equal(ta[key], tb[key])
