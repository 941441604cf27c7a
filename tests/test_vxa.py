"""VXA/VXD input (host, no GPU): the product's XML reader and merge rule.

A model read from VXA text must equal the model built from the same problem
through the lattice entry point (itself pinned against the oracle's builder in
test_model.py): same topology, bit-equal material constants, settings and
phases.  Formulas are pinned by an independent recursive evaluation of random
expression trees (the product evaluates a FIFO over the reversed breadth-first
token order, as VX3_MathTree::eval does).
"""
import math
import os

import numpy as np
import pytest

import vx3mi
from vx3mi import configs
from vx3mi import vxa as W

PROBLEMS = {
    "config0": configs.config0,
    "config1_block": lambda: configs.config1(seed=2, dims=(9, 7, 5)),
    "config2_robot": lambda: configs.config2(seed=4, n=1)[0],
}

SETTING_KEYS = ("nx", "ny", "nz", "lattice_dim", "dt_frac", "floor_enabled", "enable_collision", "enable_expansion",
                "vary_temp", "temp_base", "temp_amplitude", "temp_period", "initial_temp")


def assert_same_model(a: vx3mi.Model, b: vx3mi.Model):
    assert (a.nv, a.nl, a.ns) == (b.nv, b.nl, b.ns)
    ta, tb = a.topology(), b.topology()
    for k in ta:
        assert np.array_equal(ta[k], tb[k]), k
    for v in range(a.nv):
        assert np.array_equal(a.voxel_constants(v), b.voxel_constants(v)), v
    for l in range(a.nl):
        assert np.array_equal(a.link_constants(l), b.link_constants(l)), l
    sa, sb = a.settings(), b.settings()
    for k in SETTING_KEYS:
        assert sa[k] == sb[k], k
    assert np.array_equal(a.phase(), b.phase())
    assert a.recommended_dt() == b.recommended_dt()


@pytest.mark.parametrize("compression", W.COMPRESSIONS)
@pytest.mark.parametrize("name", list(PROBLEMS))
def test_vxa_equals_lattice_model(name, compression):
    prob = PROBLEMS[name]()
    ref = vx3mi.Model.from_problem(prob)
    assert_same_model(vx3mi.Model.from_xml(W.write_vxa(prob, compression)), ref)


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_base_plus_vxd_merge(name, tmp_path):
    prob = PROBLEMS[name]()
    base, vxd = W.write_pair(prob, "BASE64")
    # the base alone holds an empty 1x1x1 structure
    assert vx3mi.Model.from_xml(base).nv == 0
    assert_same_model(vx3mi.Model.from_xml(base, vxd), vx3mi.Model.from_problem(prob))
    (tmp_path / "base.vxa").write_text(base)
    (tmp_path / "robot.vxd").write_text(vxd)
    m = vx3mi.Model.from_vxa(str(tmp_path / "base.vxa"), str(tmp_path / "robot.vxd"))
    assert_same_model(m, vx3mi.Model.from_problem(prob))


def test_vxd_replaces_nested_paths_and_creates_parents():
    prob = configs.config0()
    base = W.write_vxa(dict(prob, bond_damp_z=0.3, dt_frac=0.5))
    vxd = ('<VXD><Damping replace="VXA.Simulator.Damping"><BondDampingZ>0.1</BondDampingZ>'
           '<ColDampingZ>0.8</ColDampingZ><SlowDampingZ>1.0</SlowDampingZ></Damping>'
           '<DtFrac replace="VXA.Simulator.Integration.DtFrac">0.9</DtFrac>'
           '<RecordHistory replace="VXA.Simulator.RecordHistory"><RecordStepSize>7</RecordStepSize></RecordHistory>'
           '<Ignored>no replace attribute</Ignored></VXD>')
    m = vx3mi.Model.from_xml(base, vxd)
    assert_same_model(m, vx3mi.Model.from_problem(prob))
    s = m.settings()
    assert (s["record_step_size"], s["record_voxel"], s["record_link"]) == (7, 1, 0)


DEMO_LIKE = """<?xml version="1.0"?>
<!-- one voxel resting one cell above the floor -->
<VXA Version="1.1">
<Simulator>
  <RecordHistory><RecordStepSize>100</RecordStepSize></RecordHistory>
  <StopCondition><StopConditionFormula>
    <mtSUB><mtVAR>t</mtVAR><mtCONST>1</mtCONST></mtSUB>
  </StopConditionFormula></StopCondition>
</Simulator>
<VXC Version="0.94">
  <Lattice><Lattice_Dim>0.01</Lattice_Dim></Lattice>
  <Palette><Material><Name>A &amp; B</Name>
    <Display><Red>1</Red><Green>0</Green><Blue>0</Blue><Alpha>1</Alpha></Display>
    <Mechanical><MatModel>0</MatModel><Elastic_Mod>1e+05</Elastic_Mod><Density>1e+3</Density>
      <Poissons_Ratio>0.35</Poissons_Ratio><CTE>0</CTE><uStatic>1</uStatic><uDynamic>0.8</uDynamic>
    </Mechanical></Material></Palette>
  <Structure Compression="ASCII_READABLE">
    <X_Voxels>1</X_Voxels><Y_Voxels>1</Y_Voxels><Z_Voxels>2</Z_Voxels>
    <Data><Layer><![CDATA[0]]></Layer><Layer><![CDATA[1]]></Layer></Data>
  </Structure>
</VXC>
</VXA>
"""


def test_reference_defaults_without_environment():
    """No <Environment>: CVX_Environment constructor defaults (gravity, floor and
    temperature on, TempBase 25, amplitude 0, period 0.1); no <Integration>:
    DtFrac 0.9; EnableCollision defaults to true; RecordVoxel to 1."""
    m = vx3mi.Model.from_xml(DEMO_LIKE, "<VXD>\n</VXD>")
    assert (m.nv, m.nl) == (1, 0)
    assert m.topology()["lattice"].tolist() == [[0, 0, 1]]
    s = m.settings()
    assert s["dt_frac"] == 0.9 and s["floor_enabled"] == 1 and s["enable_collision"] == 1
    assert (s["temp_base"], s["temp_amplitude"], s["temp_period"], s["vary_temp"]) == (25.0, 0.0, 0.1, 0)
    assert (s["record_step_size"], s["record_voxel"], s["record_link"]) == (100, 1, 0)
    assert s["initial_temp"] == 0.0
    c = m.voxel_constants(0)
    f = np.float32
    mass = f(f(1000.0) * f(0.01) ** 3)
    assert c[0] == f(1000.0 * 0.01 ** 3) or c[0] == mass
    assert c[7] == f(f(-c[0]) * f(9.80665)) * f(-f(-9.81) / 9.80665)  # gravity force, GravAcc -9.81
    assert c[9] == f(1.0) and c[10] == f(0.8)
    # no links: dt from the voxel fallback E*L/m
    assert m.recommended_dt() == pytest.approx(1.0 / (6.283185 * math.sqrt(1e5 * 0.01 / float(c[0]))), rel=1e-6)
    assert m.eval_formula(vx3mi.FORMULA_STOP, t=0.25) == -0.75


def test_environment_missing_key_defaults_inside_blocks():
    """Inside <Gravity> / <Thermal> the missing-key defaults differ from the
    constructor's (VX_Environment.cpp:150-171): FloorEnabled and TempEnabled
    read as off; the legacy TempAmp key is an absolute temperature."""
    prob = configs.config0()
    text = W.write_vxa(prob).replace("<FloorEnabled>1</FloorEnabled>", "")
    text = text.replace("<TempAmplitude>39.0</TempAmplitude>", "<TempAmp>64.0</TempAmp>")
    s = vx3mi.Model.from_xml(text).settings()
    assert s["floor_enabled"] == 0
    assert s["temp_amplitude"] == 64.0 - 25.0
    t2 = W.write_vxa(prob).replace("<TempEnabled>1</TempEnabled>", "").replace("<VaryTempEnabled>1</VaryTempEnabled>", "")
    s2 = vx3mi.Model.from_xml(t2).settings()
    assert s2["vary_temp"] == 0 and s2["initial_temp"] == 0.0


def test_static_temperature_sets_initial_temp():
    """TempEnabled without VaryTemp: every voxel starts at TempBase + amplitude
    relative to TempBase (UpdateMatTemps)."""
    prob = dict(configs.config0(), vary_temp=0, temp_amplitude=5.0)
    s = vx3mi.Model.from_xml(W.write_vxa(prob)).settings()
    assert s["initial_temp"] == np.float32(5.0)


def test_phase_offset_layers_follow_voxel_order():
    prob = configs.config2(seed=9, n=1)[0]
    m = vx3mi.Model.from_xml(W.write_vxa(prob))
    assert np.array_equal(m.phase(), np.asarray(prob["phase"]))


# ---- formulas -----------------------------------------------------------------
BIN = {"mtADD": lambda a, b: a + b, "mtSUB": lambda a, b: a - b, "mtMUL": lambda a, b: a * b,
       "mtDIV": lambda a, b: a / b if b != 0 else math.copysign(math.inf, a) * math.copysign(1, b) if a else math.nan,
       "mtGREATERTHAN": lambda a, b: float(a > b), "mtLESSTHAN": lambda a, b: float(a < b),
       "mtAND": lambda a, b: float(bool(a) and bool(b)), "mtOR": lambda a, b: float(bool(a) or bool(b))}
UN = {"mtABS": abs, "mtSIN": math.sin, "mtCOS": math.cos, "mtATAN": math.atan, "mtNOT": lambda a: float(a == 0),
      "mtINT": lambda a: float(round(a)), "mtNORMALCDF": lambda a: 0.5 * math.erfc(-a / math.sqrt(2))}
VARS = vx3mi.FORMULA_VARS


def random_tree(rng, depth):
    if depth == 0 or rng.random() < 0.25:
        r = rng.random()
        if r < 0.45:
            return ("mtVAR", VARS[rng.integers(len(VARS))])
        if r < 0.9:
            return ("mtCONST", float(np.round(rng.normal() * 3, 3)))
        return ("mtPI",) if r < 0.95 else ("mtE",)
    if rng.random() < 0.7:
        op = list(BIN)[rng.integers(len(BIN))]
        return (op, random_tree(rng, depth - 1), random_tree(rng, depth - 1))
    op = list(UN)[rng.integers(len(UN))]
    return (op, random_tree(rng, depth - 1))


def eval_tree(e, env):
    op = e[0]
    if op == "mtVAR":
        return env[e[1]]
    if op == "mtCONST":
        return e[1]
    if op == "mtPI":
        return math.pi
    if op == "mtE":
        return math.e
    if op in BIN:
        return BIN[op](eval_tree(e[1], env), eval_tree(e[2], env))
    return UN[op](eval_tree(e[1], env))


def test_formulas_match_recursive_evaluation():
    rng = np.random.default_rng(5)
    base = configs.config0()
    checked = 0
    for _ in range(120):
        tree = random_tree(rng, 5)
        if tree[0] in ("mtVAR", "mtCONST", "mtPI", "mtE"):
            tree = ("mtADD", tree, ("mtCONST", 0.5))
        m = vx3mi.Model.from_xml(W.write_vxa(base, stop=tree, fitness=("mtMUL", tree, ("mtCONST", 2.0))))
        env = {k: float(np.round(rng.normal() * 2, 2)) for k in VARS}
        want = eval_tree(tree, env)
        got = m.eval_formula(vx3mi.FORMULA_STOP, **env)
        if math.isnan(want):
            assert math.isnan(got)
            continue
        assert got == pytest.approx(want, rel=1e-12, abs=1e-12), tree
        fit = m.eval_formula(vx3mi.FORMULA_FITNESS, **env)
        assert fit == pytest.approx(2.0 * want, rel=1e-12, abs=1e-12) or (math.isinf(want) and math.isinf(fit))
        checked += 1
    assert checked > 80


def test_operand_order_of_noncommutative_operators():
    m = vx3mi.Model.from_xml(W.write_vxa(configs.config0(), stop=("mtDIV", ("mtPOW", ("mtVAR", "x"), ("mtCONST", 3)),
                                                                    ("mtSUB", ("mtVAR", "t"), ("mtVAR", "z")))))
    assert m.eval_formula(vx3mi.FORMULA_STOP, x=2.0, t=5.0, z=1.0) == 8.0 / 4.0


# ---- errors -------------------------------------------------------------------
BAD = {
    "unterminated": "<VXA><Simulator></VXA>",
    "no_vxa": "<Foo/>",
    "layer_len": DEMO_LIKE.replace("<![CDATA[0]]>", "<![CDATA[00]]>"),
    "zlib": DEMO_LIKE.replace('Compression="ASCII_READABLE"', 'Compression="ZLIB"'),
    "few_layers": DEMO_LIKE.replace("<Z_Voxels>2</Z_Voxels>", "<Z_Voxels>3</Z_Voxels>"),
    "attach_bad_var": DEMO_LIKE.replace("<Simulator>", "<Simulator><AttachDetach><EnableAttach>true</EnableAttach>"
                                                       "<AttachCondition><Condition_2><mtVAR>speed</mtVAR></Condition_2>"
                                                       "</AttachCondition></AttachDetach>"),
    "bc_count": DEMO_LIKE.replace("<VXC ", "<Environment><Boundary_Conditions><NumBCs>1</NumBCs></Boundary_Conditions>"
                                          "</Environment><VXC "),
    "bc_mesh_index": DEMO_LIKE.replace("<VXC ", "<Environment><Boundary_Conditions><NumBCs>1</NumBCs><FRegion><PrimType>3"
                                               "</PrimType><CMesh><Vertices><Vertex><Vx>1</Vx></Vertex></Vertices>"
                                               "<Facets><Facet><V0>0</V0><V1>1</V1><V2>0</V2></Facet></Facets></CMesh>"
                                               "</FRegion></Boundary_Conditions></Environment><VXC "),
    "bc_no_prim": DEMO_LIKE.replace("<VXC ", "<Environment><Boundary_Conditions><NumBCs>1</NumBCs><FRegion><X>0</X>"
                                            "</FRegion></Boundary_Conditions></Environment><VXC "),
    "bad_token": DEMO_LIKE.replace("<mtSUB>", "<mtFOO>").replace("</mtSUB>", "</mtFOO>"),
    "bad_var": DEMO_LIKE.replace("<mtVAR>t</mtVAR>", "<mtVAR>speed</mtVAR>"),
    "missing_operand": DEMO_LIKE.replace("<mtCONST>1</mtCONST>", ""),
}


@pytest.mark.parametrize("case", list(BAD))
def test_malformed_or_unsupported_input_is_a_parse_error(case):
    with pytest.raises(vx3mi.VX3MIError, match="error -2"):
        vx3mi.Model.from_xml(BAD[case])


def test_attach_settings_sticky_and_conditions_are_read():
    # VX3_SimulationManager.cu:329-343 (EnableAttach, SafetyGuard, AttachCondition)
    # and VX_Object.cpp:1436 (Sticky)
    text = DEMO_LIKE.replace("<Simulator>", "<Simulator><AttachDetach><EnableAttach>1</EnableAttach>"
                                            "<SafetyGuard>37</SafetyGuard><AttachCondition><Condition_0>"
                                            "<mtGREATERTHAN><mtVAR>t</mtVAR><mtCONST>0.5</mtCONST></mtGREATERTHAN>"
                                            "</Condition_0></AttachCondition></AttachDetach>")
    text = text.replace("<MatModel>0</MatModel>", "<MatModel>0</MatModel><Sticky>1</Sticky>", 1)
    m = vx3mi.Model.from_xml(text)
    st = m.settings()
    assert st["enable_attach"] == 1 and st["safety_guard"] == 37
    assert m.material(1)["sticky"] == 1
    # without the keys: attach off, the reference's SafetyGuard default
    st0 = vx3mi.Model.from_xml(DEMO_LIKE).settings()
    assert st0["enable_attach"] == 0 and st0["safety_guard"] == 500


def test_attach_model_limits():
    base = dict(dims=(2, 1, 1), cells=[1, 1], materials=[dict(sticky=1)], enable_attach=1)
    with pytest.raises(vx3mi.VX3MIError, match="SafetyGuard"):
        vx3mi.Model.from_problem(dict(base, safety_guard=70000))
    with pytest.raises(vx3mi.VX3MIError, match="per-voxel"):
        vx3mi.Model.from_problem(dict(base, voxel_E=[1e6, 2e6]))
    with pytest.raises(vx3mi.VX3MIError, match="Poisson"):
        vx3mi.Model.from_problem(dict(base, volume_effects=1, materials=[dict(sticky=1, nu=0.3)]))
    m = vx3mi.Model.from_problem(base)
    m.set_attach_condition(1, ("mtVAR", "angle"))
    m.set_attach_condition(4, ("mtLESSTHAN", ("mtVAR", "z"), ("mtCONST", 0.05)))


def test_sampled_variables_are_accepted_in_force_fields_and_conditions():
    # angle, targetCloseness and numClosePairs are the CycleStep samples
    # (VX3_VoxelyzeKernel.cu:314-334, 545-562) that every force-field evaluation reads
    for var in ("angle", "targetCloseness", "numClosePairs"):
        text = DEMO_LIKE.replace("<Simulator>", f"<Simulator><ForceField><x_forcefield><mtVAR>{var}</mtVAR>"
                                                "</x_forcefield></ForceField>")
        vx3mi.Model.from_xml(text)


def test_zero_force_field_and_collision_false_are_accepted():
    text = DEMO_LIKE.replace("<Simulator>", "<Simulator><ForceField><x_forcefield><mtCONST>0</mtCONST></x_forcefield>"
                                            "</ForceField><AttachDetach><EnableCollision>false</EnableCollision>"
                                            "</AttachDetach>")
    assert vx3mi.Model.from_xml(text).settings()["enable_collision"] == 0


def test_missing_file_is_a_parse_error(tmp_path):
    with pytest.raises(vx3mi.VX3MIError, match="cannot read"):
        vx3mi.Model.from_vxa(str(tmp_path / "nope.vxa"))


# ---- boundary conditions (Environment/Boundary_Conditions) --------------------
def _externals(m):
    return {v: m.external(v) for v in range(m.nv) if m.external(v) is not None}


def _cube_with_regions(regions, **kw):
    return vx3mi.Model.from_xml(W.write_vxa(dict(configs.config0(), bc_regions=regions, **kw)))


def test_box_region_fixes_a_face_and_shares_its_force():
    """A box over the first fifth of x holds the 16 voxels of the x = 0 face
    (DofFixed 63) and splits ForceZ over them (CVX_Sim::Import divides by the
    number of voxels the region touches)."""
    m = _cube_with_regions([dict(PrimType=0, X=0.0, Y=0.0, Z=0.0, dX=0.2, dY=1.0, dZ=1.0, DofFixed=63,
                                 ForceZ=-1.6)])
    ext = _externals(m)
    lat = m.topology()["lattice"]
    assert sorted(ext) == [v for v in range(m.nv) if lat[v, 0] == 0] and len(ext) == 16
    for e in ext.values():
        assert e["dof"] == 63 and e["force"][2] == float(np.float32(-0.1)) and not e["force"][:2].any()


def test_sphere_and_cylinder_regions_select_by_distance():
    """Sphere of radius 0.1 workspace at the centre: the 8 central voxels (their
    centres lie within radius + half a voxel diagonal); a z cylinder of radius 0.2
    through the centre: the 2x2 central columns, 16 voxels."""
    sph = _externals(_cube_with_regions([dict(PrimType=2, X=0.5, Y=0.5, Z=0.5, Radius=0.1, DofFixed=1)]))
    cyl = _externals(_cube_with_regions([dict(PrimType=1, X=0.5, Y=0.5, Z=0.0, dZ=1.0, Radius=0.2, DofFixed=2)]))
    lat = vx3mi.Model.from_problem(configs.config0()).topology()["lattice"]
    centre = lambda v: all(c in (1, 2) for c in lat[v])  # noqa: E731
    column = lambda v: lat[v, 0] in (1, 2) and lat[v, 1] in (1, 2)  # noqa: E731
    assert sorted(sph) == [v for v in range(64) if centre(v)] and all(e["dof"] == 1 for e in sph.values())
    assert sorted(cyl) == [v for v in range(64) if column(v)] and all(e["dof"] == 2 for e in cyl.values())


def test_displacements_and_overlapping_regions():
    """Prescribed displacement / rotation of fixed degrees of freedom, and two
    overlapping regions: the reference adds external()->force() + F/n to the
    force already there, so the first region's share counts twice."""
    r1 = dict(PrimType=0, X=0.0, Y=0.0, Z=0.0, dX=0.2, dY=1.0, dZ=1.0, DofFixed=4 | 8, DisplaceZ=0.001,
              AngDisplaceX=0.1, ForceX=0.8)
    r2 = dict(PrimType=0, X=0.0, Y=0.0, Z=0.0, dX=1.0, dY=0.2, dZ=1.0, ForceX=0.4, TorqueY=0.2)
    ext = _externals(_cube_with_regions([r1, r2]))
    lat = vx3mi.Model.from_problem(configs.config0()).topology()["lattice"]
    f1, f2 = np.float32(0.8 / 16), np.float32(0.4 / 16)
    for v, e in ext.items():
        in1, in2 = lat[v, 0] == 0, lat[v, 1] == 0
        assert in1 or in2
        if in1:
            assert e["dof"] == 12 and e["translation"][2] == 0.001 and e["rotation"][0] == 0.1
        else:
            assert e["dof"] == 0 and not e["translation"].any()
        expect = f1 if in1 else np.float32(0)
        if in2:
            expect = np.float32(expect + np.float32(np.float64(expect) + 0.4 / 16))
        assert e["force"][0] == float(expect), (v, e["force"][0], expect)
        assert e["moment"][1] == (float(np.float32(0.2 / 16)) if in2 else 0.0)
    assert len(ext) == 16 + 16 - 4


CUBE_MESH = (
    [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)],
    [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (3, 7, 6), (3, 6, 2), (0, 4, 7), (0, 7, 3),
     (1, 2, 6), (1, 6, 5)])
TETRA_MESH = ([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)], [(0, 2, 1), (0, 1, 3), (0, 3, 2), (1, 2, 3)])


def _solid(n):
    return dict(configs.config0(), dims=(n, n, n), cells=np.ones(n ** 3, dtype=np.int8), phase=None)


def _mesh_points(n, X, dX, size=1.0):
    """Voxel centres of an n^3 workspace mapped into mesh coordinates as
    CP_Mesh::IsIn maps them (region origin X, uniform scale min(dX * Env / size))."""
    idx = np.stack(np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij"), -1).reshape(-1, 3)[:, ::-1]
    centre = (idx + 0.5) / n  # workspace fractions, x fastest like the voxel order
    scale = min(d / size for d in dX)
    return (centre - np.asarray(X)) / scale


def test_mesh_regions_select_the_voxels_whose_centres_lie_inside():
    """Mesh regions (PrimType 3, CP_Mesh::IsIn, src/old/VX_FRegion.cpp:668-686):
    the mesh is scaled uniformly into the region's box and a voxel is in when its
    centre is inside the closed surface (x-ray parity, CMesh::IsInside).  Region
    offsets keep the rays off the facets' shared edges (a ray through an edge
    meets both facets, as in the reference).  A unit cube mesh: the voxels whose
    mapped centres lie in the open unit cube; a corner tetrahedron: those with
    every coordinate positive and x + y + z < 1."""
    X, dX = (0.25, 0.2, 0.3), (0.5, 0.5, 0.5)
    cube = vx3mi.Model.from_xml(W.write_vxa(dict(configs.config0(), bc_regions=[
        dict(PrimType=3, X=X[0], Y=X[1], Z=X[2], dX=dX[0], dY=dX[1], dZ=dX[2], DofFixed=7, mesh=CUBE_MESH)])))
    ext = _externals(cube)
    pts = _mesh_points(4, X, dX)
    assert np.abs(pts[:, 1] - pts[:, 2]).min() > 1e-6  # the x faces are split along y = z
    want = [v for v in range(64) if ((pts[v] > 0) & (pts[v] < 1)).all()]
    assert len(want) == 8 and sorted(ext) == want
    assert all(e["dof"] == 7 for e in ext.values())
    X, dX = (0.037, 0.021, 0.013), (0.93, 0.91, 0.95)
    tet = vx3mi.Model.from_xml(W.write_vxa(dict(_solid(6), bc_regions=[
        dict(PrimType=3, X=X[0], Y=X[1], Z=X[2], dX=dX[0], dY=dX[1], dZ=dX[2], DofFixed=1, ForceZ=-3.0,
             mesh=TETRA_MESH)])))
    ext = _externals(tet)
    pts = _mesh_points(6, X, dX)
    assert np.abs(pts[:, 1] + pts[:, 2] - 1).min() > 1e-6
    want = [v for v in range(216) if (pts[v] > 0).all() and pts[v].sum() < 1]
    assert len(want) > 20 and sorted(ext) == want
    for e in ext.values():  # the force is shared by the voxels the region touches
        assert e["force"][2] == float(np.float32(-3.0 / len(want)))


def test_mesh_region_scales_by_its_smallest_ratio():
    """The mesh keeps its aspect: a unit cube mesh in a region of 0.5 x 1 x 1
    workspace fractions is scaled by the smallest ratio (0.5) and covers the
    first half of the workspace along every axis from the region's origin."""
    m = vx3mi.Model.from_xml(W.write_vxa(dict(configs.config0(), bc_regions=[
        dict(PrimType=3, X=0.01, Y=0.02, Z=0.03, dX=0.5, dY=1.0, dZ=1.0, DofFixed=63, mesh=CUBE_MESH)])))
    ext = _externals(m)
    lat = m.topology()["lattice"]
    assert sorted(ext) == [v for v in range(64) if all(c in (0, 1) for c in lat[v])]


def test_legacy_fixed_regions_and_binary_fixed_tag():
    text = W.write_vxa(configs.config0()).replace(
        "<Environment>", "<Environment><Fixed_Regions><NumFixed>1</NumFixed><FRegion><PrimType>0</PrimType>"
                         "<X>0</X><Y>0</Y><Z>0</Z><dX>1</dX><dY>1</dY><dZ>0.2</dZ><Fixed>1</Fixed></FRegion>"
                         "</Fixed_Regions>")
    ext = _externals(vx3mi.Model.from_xml(text))
    lat = vx3mi.Model.from_problem(configs.config0()).topology()["lattice"]
    assert sorted(ext) == [v for v in range(64) if lat[v, 2] == 0] and all(e["dof"] == 63 for e in ext.values())


def test_vxa_regions_equal_explicit_externals():
    """The same externals through a VXA region and through set_external give the
    same model (same external per voxel)."""
    reg = dict(PrimType=0, X=0.0, Y=0.0, Z=0.0, dX=0.2, dY=1.0, dZ=1.0, DofFixed=63, DisplaceZ=5e-4)
    a = _cube_with_regions([reg])
    b = vx3mi.Model.from_problem(configs.config0())
    lat = b.topology()["lattice"]
    for v in range(b.nv):
        if lat[v, 0] == 0:
            b.set_external(v, 63, translation=(0, 0, 5e-4))
    ea, eb = _externals(a), _externals(b)
    assert sorted(ea) == sorted(eb)
    for v in ea:
        for k in ("dof", "translation", "rotation", "force", "moment"):
            assert np.array_equal(ea[v][k], eb[v][k]), (v, k)


# ---- material models -------------------------------------------------------------
def _bilinear_problem():
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_golden import bilinear_robot
    return bilinear_robot()


def test_bilinear_material_from_vxa_and_oracle_curves():
    """MatModel 2 read from a VXA gives the same link materials and stress-strain
    curves as the lattice description, and the curves equal the oracle's (which
    is pinned to the reference CPU Voxelyze by the bilinear golden)."""
    import oracle as O
    prob = _bilinear_problem()
    text = W.write_vxa(prob)
    assert "<MatModel>2</MatModel>" in text and "<Yield_Stress>2000.0</Yield_Stress>" in text
    a, b, o = vx3mi.Model.from_problem(prob), vx3mi.Model.from_xml(text), O.OracleModel(prob)
    n_nonlinear = 0
    for l in range(a.nl):
        ca = a.link_curve(l)
        assert np.array_equal(ca, b.link_curve(l)) and np.array_equal(ca, o.link_curve(l)), l
        assert np.array_equal(a.link_constants(l), b.link_constants(l))
        n_nonlinear += len(ca) > 0
        if len(ca):
            assert ca[0, 0] == 0 and np.all(np.diff(ca[:, 0]) > 0) and np.all(np.diff(ca[:, 1]) > 0)
    assert 0 < n_nonlinear < a.nl  # links between two linear voxels stay linear


def test_rejected_material_models_keep_the_default():
    """CVX_Sim::CopyMat: a material setter that rejects its arguments leaves
    addMaterial's default (linear, 1e6 Pa): MatModel 1 with a zero failure
    stress, MatModel 2 without a failure stress above the yield stress."""
    base = configs.config0()
    soft = dict(configs.SOFT, E=5e6)
    penK = lambda m: m.voxel_constants(0)[8]  # noqa: E731  penetrationStiffness = 2 E L
    good = vx3mi.Model.from_xml(W.write_vxa(dict(base, materials=[soft])))
    assert penK(good) == np.float32(2 * 5e6 * 0.01)
    t1 = W.write_vxa(dict(base, materials=[dict(soft, fail_stress=1e4)])).replace(
        "<Fail_Stress>10000.0</Fail_Stress>", "<Fail_Stress>0</Fail_Stress>")
    t2 = W.write_vxa(dict(base, materials=[dict(soft, yield_stress=2e3, plastic_mod=1e5)])).replace(
        "<Fail_Stress>-1.0</Fail_Stress>", "")
    for t in (t1, t2):
        assert penK(vx3mi.Model.from_xml(t)) == np.float32(2 * 1e6 * 0.01)
    ok = W.write_vxa(dict(base, materials=[dict(soft, yield_stress=2e3, plastic_mod=1e5)]))
    m = vx3mi.Model.from_xml(ok)
    assert penK(m) == np.float32(2 * 5e6 * 0.01) and len(m.link_curve(0)) == 3


# ---- force fields ----------------------------------------------------------------
def test_force_fields_are_read_per_axis():
    """Simulator/ForceField/{x,y,z}_forcefield: a non-zero tree is kept (the
    reference adds it to each voxel's force), a constant-zero one is the default;
    a tree deeper than the device evaluator's stack is rejected."""
    ff = {"x": ("mtMUL", ("mtCONST", 0.01), ("mtSIN", ("mtVAR", "t"))), "z": ("mtCONST", 0.0)}
    m = vx3mi.Model.from_xml(W.write_vxa(dict(configs.config0(), force_field=ff)))
    assert m.settings()["nx"] == 4
    deep = ("mtVAR", "x")
    for _ in range(20):  # right-leaning chain: each level keeps one more value on the stack
        deep = ("mtADD", ("mtVAR", "x"), deep)
    with pytest.raises(vx3mi.VX3MIError, match="error -2"):
        vx3mi.Model.from_xml(W.write_vxa(dict(configs.config0(), force_field={"y": deep})))


def _cilia_problem(seed=0):
    p = configs.config1(seed=seed, dims=(5, 4, 3))
    rng = np.random.default_rng(seed)
    nv = int((np.asarray(p["cells"]) > 0).sum())
    mats = [dict(m) for m in p["materials"]]
    mats[0].update(cilia=2.5, cilia_on_after=0.01)
    mats[1].update(remove_after=0.02)
    return dict(p, materials=mats, enable_cilia=1, secondary_experiment=1,
                cilia_base=rng.normal(size=(nv, 3)) * 1e-3, cilia_shift=rng.normal(size=(nv, 3)) * 1e-4)


def test_cilia_and_removal_keys_are_read():
    """Simulator/EnableCilia and SecondaryExperiment, the materials' Cilia,
    TurnOnCiliaAfterThisManySeconds and RemoveFromSimulationAfterThisManySeconds,
    and the BaseCiliaForce / ShiftCiliaForce layers (three values per cell) give
    the model the lattice entry point builds from the same problem."""
    p = _cilia_problem()
    a = vx3mi.Model.from_xml(W.write_vxa(p))
    b = vx3mi.Model.from_problem(p)
    assert_same_model(a, b)
    for m in (a, b):
        st = m.settings()
        assert st["enable_cilia"] == 1 and st["secondary_experiment"] == 1
        assert m.material(1)["cilia"] == 2.5 and m.material(1)["cilia_on_after"] == 0.01
        assert m.material(2)["remove_after"] == 0.02 and m.material(2)["cilia"] == 0.0
        base, shift = m.cilia()
        assert np.array_equal(base, p["cilia_base"]) and np.array_equal(shift, p["cilia_shift"])


def test_short_cilia_layers_pad_zeros():
    """Cells past the end of a BaseCiliaForce / ShiftCiliaForce layer get zeros
    (VX_Object.cpp:1906-1908 pads them).  Values missing from a cell's partial
    triple are read as zero (the reference reads past its data there)."""
    text = DEMO_LIKE.replace("</Data>", "</Data><BaseCiliaForce><Layer><![CDATA[]]></Layer>"
                                        "<Layer><![CDATA[0.5,-1,2]]></Layer></BaseCiliaForce>"
                                        "<ShiftCiliaForce><Layer><![CDATA[]]></Layer>"
                                        "<Layer><![CDATA[3]]></Layer></ShiftCiliaForce>")
    text = text.replace("<Simulator>", "<Simulator><EnableCilia>1</EnableCilia>")
    m = vx3mi.Model.from_xml(text)
    base, shift = m.cilia()
    assert base.tolist() == [[0.5, -1.0, 2.0]]
    assert shift.tolist() == [[3.0, 0.0, 0.0]]
    assert m.settings()["enable_cilia"] == 1


def test_data_material_model_from_vxa_and_oracle_curves():
    """MatModel 3 (SSData) read from a VXA gives the lattice description's link
    materials; the link curves (series combinations of the data, up to 2 x 32 + 1
    points) and link constants equal the oracle's, which the data-model golden
    pins to the reference CPU Voxelyze."""
    import sys
    import oracle as O
    sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden"))
    from make_golden import golden_problems
    prob = golden_problems()["data_model"]
    text = W.write_vxa(prob)
    assert "<MatModel>3</MatModel>" in text and "<NumDataPts>5</NumDataPts>" in text
    a, b, o = vx3mi.Model.from_problem(prob), vx3mi.Model.from_xml(text), O.OracleModel(prob)
    assert_same_model(a, b)
    for l in range(a.nl):
        ca = a.link_curve(l)
        assert len(ca) > 0 and np.array_equal(ca, b.link_curve(l)) and np.array_equal(ca, o.link_curve(l)), l
        assert np.array_equal(a.link_constants(l), o.link_material(l)), l
    # 32 points per material: the longest combination fits
    st = np.cumsum(np.linspace(1e-4, 2e-4, 32))
    ss = np.cumsum(np.linspace(200.0, 20.0, 32))
    st2, ss2 = st * 1.37, ss * 1.5
    big = dict(prob, materials=[dict(rho=1000.0, ss_data=(st, ss)), dict(rho=1000.0, ss_data=(st2, ss2))])
    mb, ob = vx3mi.Model.from_xml(W.write_vxa(big)), O.OracleModel(big)
    lens = {len(mb.link_curve(l)) for l in range(mb.nl)}
    assert max(lens) > 33
    for l in range(mb.nl):
        assert np.array_equal(mb.link_curve(l), ob.link_curve(l)), l


DATA_BAD = {
    "no_points": ("<SSData><NumDataPts>0</NumDataPts></SSData>", "error -2"),
    "short": ("<SSData><NumDataPts>2</NumDataPts><StrainData><Strain>0.1</Strain></StrainData>"
              "<StressData><Stress>1</Stress><Stress>2</Stress></StressData></SSData>", "error -2"),
    "too_many": ("<SSData><NumDataPts>33</NumDataPts></SSData>", "error -5"),
}


@pytest.mark.parametrize("case", list(DATA_BAD))
def test_data_material_input_errors(case):
    """MatModel 3 without points (the reference then reads an empty vector),
    with fewer values than NumDataPts, or with more than 32 points is an error."""
    body, code = DATA_BAD[case]
    text = DEMO_LIKE.replace("<MatModel>0</MatModel>", "<MatModel>3</MatModel>" + body)
    with pytest.raises(vx3mi.VX3MIError, match=code):
        vx3mi.Model.from_xml(text)


def test_rejected_data_keeps_the_default():
    """CVX_Material::setModel rejects strains that do not increase and a first
    point at or below zero; the material keeps addMaterial's default (linear,
    1e6 Pa), as for the other models.  One point is a linear material of that
    slope; decreasing stresses are accepted."""
    base = configs.config0()
    penK = lambda m: m.voxel_constants(0)[8]  # noqa: E731  2 E L
    for st, ss in (([0.001, 0.001], [1.0, 2.0]), ([-0.001, 0.002], [1.0, 2.0]), ([0.002, 0.001], [1.0, 2.0])):
        m = vx3mi.Model.from_xml(W.write_vxa(dict(base, materials=[dict(rho=1000.0, ss_data=(st, ss))])))
        assert penK(m) == np.float32(2 * 1e6 * 0.01) and len(m.link_curve(0)) == 0
    one = vx3mi.Model.from_xml(W.write_vxa(dict(base, materials=[dict(rho=1000.0, ss_data=([0.002], [6000.0]))])))
    E1 = np.float32(np.float32(6000.0) / np.float32(0.002))
    assert penK(one) == np.float32(float(np.float32(2) * E1) * 0.01)  # (float)(2 E L), 2 E in float
    assert len(one.link_curve(0)) == 0
    dec = vx3mi.Model.from_xml(W.write_vxa(dict(base, materials=[dict(rho=1000.0, ss_data=([0.001, 0.002], [2000.0, 1500.0]))])))
    assert len(dec.link_curve(0)) == 3


def test_signal_keys_are_read_with_reference_defaults():
    """Simulator/EnableSignals and the materials' isPaceMaker, PaceMakerPeriod,
    signalValueDecay, signalTimeDelay and inactivePeriod; absent keys take
    ReadXML's defaults (0.9, 0.03, 0.03; VX_Object.cpp:1420-1424)."""
    p = configs.config0()
    p = dict(p, enable_signals=1, materials=[dict(configs.SOFT, is_pacemaker=1, pacemaker_period=0.25,
                                                  signal_decay=0.7, signal_delay=0.004, inactive_period=0.01)])
    m = vx3mi.Model.from_xml(W.write_vxa(p))
    assert m.settings()["enable_signals"] == 1
    mat = m.material(1)
    assert (mat["is_pacemaker"], mat["pacemaker_period"], mat["signal_value_decay"], mat["signal_time_delay"],
            mat["inactive_period"]) == (1, 0.25, 0.7, 0.004, 0.01)
    d = vx3mi.Model.from_xml(DEMO_LIKE).material(1)
    assert (d["is_pacemaker"], d["signal_value_decay"], d["signal_time_delay"], d["inactive_period"]) == (0, 0.9, 0.03, 0.03)


def test_detach_key_is_read():
    """Simulator/AttachDetach/EnableDetach (a bool, VX3_SimulationManager.cu:330)."""
    m = vx3mi.Model.from_xml(W.write_vxa(dict(configs.config0(), enable_detach=1)))
    assert m.settings()["enable_detach"] == 1
    text = DEMO_LIKE.replace("<Simulator>", "<Simulator><AttachDetach><EnableDetach>true</EnableDetach></AttachDetach>")
    assert vx3mi.Model.from_xml(text).settings()["enable_detach"] == 1


def test_target_and_pair_keys_are_read():
    """isTarget and Simulator/MaxDistInVoxelLengthsToCountAsPair."""
    p = dict(configs.config0(), max_dist_pairs=1.5, materials=[dict(configs.SOFT, is_target=1)])
    m = vx3mi.Model.from_xml(W.write_vxa(p))
    assert m.material(1)["is_target"] == 1 and m.settings()["max_dist_pairs"] == 1.5
