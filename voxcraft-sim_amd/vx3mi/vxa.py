"""VXA/VXD writer: turns a problem dict into the XML the reference reads.

The element names and nesting follow what the reference's readers look up:
simulator keys (src/VXA/VX_Sim.cpp:185-294, VX3_SimulationManager.cu:323-376),
environment keys (src/old/VX_Environment.cpp:101-171), palette and structure
(src/VXA/VX_Object.cpp:1373-1475, 1760-1958).  A base VXA plus a VXD whose
children carry ``replace="VXA.a.b"`` (ctool::ptree_merge) is the reference's
batch input; :func:`write_pair` produces such a pair, :func:`write_vxa` one
self-contained VXA.

Numbers are written with ``repr`` so every double reads back exactly.
"""
from __future__ import annotations

import base64
from typing import Sequence

import numpy as np

from . import MATERIAL_DEFAULTS, PROBLEM_DEFAULTS, material_model

COMPRESSIONS = ("ASCII_READABLE", "BASE64")


def _num(x) -> str:
    return repr(float(x))


def _el(tag: str, body: str, attrs: str = "") -> str:
    return f"<{tag}{attrs}>{body}</{tag}>"


def formula_xml(expr) -> str:
    """MathTree XML of a nested tuple expression: ("mtSUB", ("mtVAR", "t"), ("mtCONST", 1.0))."""
    op = expr[0]
    if op == "mtVAR":
        return _el(op, expr[1])
    if op == "mtCONST":
        return _el(op, _num(expr[1]))
    if op in ("mtE", "mtPI"):
        return _el(op, "")
    return _el(op, "".join(formula_xml(e) for e in expr[1:]))


def _attach(p: dict) -> str:
    """EnableAttach, SafetyGuard and AttachCondition/Condition_k (problem key
    "attach_conditions": up to five formulas, None for none)."""
    if not p.get("enable_attach"):
        return ""
    conds = p.get("attach_conditions") or []
    body = "".join(_el(f"Condition_{k}", formula_xml(c)) for k, c in enumerate(conds) if c is not None)
    return (_el("EnableAttach", "1") + _el("SafetyGuard", str(int(p.get("safety_guard", 500))))
            + (_el("AttachCondition", body) if body else ""))


def _simulator(p: dict, stop=None, fitness=None, record=None) -> str:
    parts = [
        _el("Integration", _el("DtFrac", _num(p["dt_frac"]))),
        _el("Damping", _el("BondDampingZ", _num(p["bond_damp_z"])) + _el("ColDampingZ", _num(p["col_damp_z"]))
            + _el("SlowDampingZ", _num(p["slow_damp_z"]))),
        _el("Features", _el("VolumeEffectsEnabled", str(int(p["volume_effects"])))),
        _el("AttachDetach", _el("EnableCollision", str(int(p["enable_collision"])))
            + (_el("EnableDetach", "1") if p.get("enable_detach") else "") + _attach(p)),
        _el("EnableExpansion", str(int(p["enable_expansion"]))),
    ]
    if p.get("enable_cilia"):
        parts.append(_el("EnableCilia", "1"))
    if p.get("secondary_experiment"):
        parts.append(_el("SecondaryExperiment", "1"))
    if p.get("enable_signals"):
        parts.append(_el("EnableSignals", "1"))
    if p.get("max_dist_pairs"):
        parts.append(_el("MaxDistInVoxelLengthsToCountAsPair", _num(p["max_dist_pairs"])))
    if record is not None:
        step, vox, link = record
        parts.append(_el("RecordHistory", _el("RecordStepSize", str(int(step))) + _el("RecordVoxel", str(int(vox)))
                         + _el("RecordLink", str(int(link)))))
    if stop is not None:
        parts.append(_el("StopCondition", _el("StopConditionFormula", formula_xml(stop))))
    if fitness is not None:
        parts.append(_el("FitnessFunction", formula_xml(fitness)))
    ff = p.get("force_field") or {}
    if ff:  # Simulator/ForceField/{x,y,z}_forcefield MathTrees (problem key "force_field": {"x": expr, ...})
        parts.append(_el("ForceField", "".join(_el(f"{a}_forcefield", formula_xml(ff[a])) for a in "xyz" if a in ff)))
    return _el("Simulator", "".join(parts))


def _environment(p: dict) -> str:
    grav = _el("GravEnabled", "1") + _el("GravAcc", _num(p["grav_acc"])) + _el("FloorEnabled", str(int(p["floor_enabled"])))
    therm = (_el("TempEnabled", str(int(p["temp_enabled"]))) + _el("TempBase", _num(p["temp_base"]))
             + _el("TempAmplitude", _num(p["temp_amplitude"])) + _el("VaryTempEnabled", str(int(p["vary_temp"])))
             + _el("TempPeriod", _num(p["temp_period"])))
    return _el("Environment", boundary_conditions_xml(p.get("bc_regions") or []) + _el("Gravity", grav)
               + _el("Thermal", therm))


REGION_KEYS = ("PrimType", "X", "Y", "Z", "dX", "dY", "dZ", "Radius", "DofFixed", "ForceX", "ForceY", "ForceZ",
               "TorqueX", "TorqueY", "TorqueZ", "DisplaceX", "DisplaceY", "DisplaceZ", "AngDisplaceX",
               "AngDisplaceY", "AngDisplaceZ")


def boundary_conditions_xml(regions: Sequence[dict]) -> str:
    """<Boundary_Conditions> with one <FRegion> per dict (keys as in
    CVX_FRegion::ReadXML: PrimType 0 box / 1 cylinder / 2 sphere / 3 mesh, X..dZ
    and Radius in workspace fractions, DofFixed bits, Force*, Torque*, Displace*,
    AngDisplace*; a mesh region's "mesh": (vertices, facets) as a <CMesh>); an
    empty list writes nothing."""
    if not regions:
        return ""
    body = _el("NumBCs", str(len(regions)))
    for r in regions:
        inner = "".join(_el(k, _num(r[k]) if isinstance(r[k], float) else str(r[k])) for k in REGION_KEYS if k in r)
        if r.get("mesh") is not None:
            verts, facets = r["mesh"]
            inner += _el("CMesh", _el("Vertices", "".join(
                _el("Vertex", _el("Vx", _num(float(x))) + _el("Vy", _num(float(y))) + _el("Vz", _num(float(z))))
                for x, y, z in verts)) + _el("Facets", "".join(
                    _el("Facet", _el("V0", str(int(a))) + _el("V1", str(int(b))) + _el("V2", str(int(c))))
                    for a, b, c in facets)))
        body += _el("FRegion", inner)
    return _el("Boundary_Conditions", body)


def _palette(mats: Sequence[dict]) -> str:
    out = []
    for i, m0 in enumerate(mats):
        m = dict(MATERIAL_DEFAULTS, **m0)
        r, g, b, a = m["rgba"]
        fail, model = float(m["fail_stress"]), material_model(m)
        mech = (_el("MatModel", str(model)) + _el("Elastic_Mod", _num(m["E"]))
                + _el("Fail_Stress", _num(fail if model == 2 else max(fail, 0.0))) + _el("Density", _num(m["rho"]))
                + _el("Poissons_Ratio", _num(m["nu"])) + _el("CTE", _num(m["cte"]))
                + _el("uStatic", _num(m["mu_s"])) + _el("uDynamic", _num(m["mu_k"]))
                + _el("Fixed", str(int(m["fixed"]))) + _el("isMeasured", str(int(m["is_measured"]))))
        if model == 2:
            mech += _el("Plastic_Mod", _num(m["plastic_mod"])) + _el("Yield_Stress", _num(m["yield_stress"]))
        if model == 3:  # SSData (VX_Object.cpp:1397-1418)
            strain, stress = m["ss_data"]
            mech += _el("SSData", _el("NumDataPts", str(len(strain)))
                        + _el("StrainData", "".join(_el("Strain", _num(x)) for x in strain))
                        + _el("StressData", "".join(_el("Stress", _num(x)) for x in stress)))
        if m["thermal_on_after"]:
            mech += _el("TurnOnThermalExpansionAfterThisManySeconds", _num(m["thermal_on_after"]))
        if m["cilia"]:
            mech += _el("Cilia", _num(m["cilia"]))
        if m["is_target"]:
            mech += _el("isTarget", "1")
        if m["sticky"]:
            mech += _el("Sticky", "1")
        if m["cilia_on_after"]:
            mech += _el("TurnOnCiliaAfterThisManySeconds", _num(m["cilia_on_after"]))
        if m["remove_after"]:
            mech += _el("RemoveFromSimulationAfterThisManySeconds", _num(m["remove_after"]))
        if m["is_pacemaker"]:
            mech += _el("isPaceMaker", "1") + _el("PaceMakerPeriod", _num(m["pacemaker_period"]))
        mech += (_el("signalValueDecay", _num(m["signal_decay"])) + _el("signalTimeDelay", _num(m["signal_delay"]))
                 + _el("inactivePeriod", _num(m["inactive_period"])))
        disp = _el("Red", _num(r)) + _el("Green", _num(g)) + _el("Blue", _num(b)) + _el("Alpha", _num(a))
        out.append(_el("Material", _el("Name", f"material{i + 1}") + _el("Display", disp) + _el("Mechanical", mech),
                       f' ID="{i + 1}"'))
    return _el("Palette", "".join(out))


def _layer_text(values: np.ndarray, compression: str) -> str:
    if compression == "ASCII_READABLE":
        if values.min() < 0 or values.max() > 78:
            raise ValueError("ASCII_READABLE layers hold palette indices 0..78")
        return "".join(chr(48 + int(v)) for v in values)
    if compression == "BASE64":
        return base64.b64encode(values.astype(np.int8).tobytes()).decode()
    raise ValueError(f"unsupported compression {compression}")


def _structure(p: dict, compression: str, attrs: str = "") -> str:
    nx, ny, nz = (int(d) for d in p["dims"])
    cells = np.asarray(p["cells"], dtype=np.int64).reshape(nz, ny * nx)
    layers = "".join(_el("Layer", f"<![CDATA[{_layer_text(cells[z], compression)}]]>") for z in range(nz))
    body = _el("X_Voxels", str(nx)) + _el("Y_Voxels", str(ny)) + _el("Z_Voxels", str(nz)) + _el("Data", layers)
    if p.get("phase") is not None:
        phase = np.zeros(nx * ny * nz)
        flat = cells.reshape(-1)
        phase[np.flatnonzero(flat > 0)] = np.asarray(p["phase"], dtype=np.float64)
        phase = phase.reshape(nz, ny * nx)
        body += _el("PhaseOffset", "".join(
            _el("Layer", f"<![CDATA[{','.join(repr(float(v)) for v in phase[z])}]]>") for z in range(nz)))
    for key, tag in (("cilia_base", "BaseCiliaForce"), ("cilia_shift", "ShiftCiliaForce")):
        if p.get(key) is None:
            continue
        f = np.zeros((nx * ny * nz, 3))
        f[np.flatnonzero(cells.reshape(-1) > 0)] = np.asarray(p[key], dtype=np.float64).reshape(-1, 3)
        f = f.reshape(nz, ny * nx * 3)
        body += _el(tag, "".join(
            _el("Layer", f"<![CDATA[{','.join(repr(float(v)) for v in f[z])}]]>") for z in range(nz)))
    return _el("Structure", body, f' Compression="{compression}"{attrs}')


def _full(problem: dict) -> dict:
    p = dict(PROBLEM_DEFAULTS)
    p.update(problem)
    if p.get("voxel_E") is not None:
        raise ValueError("per-voxel Young's modulus has no VXA representation")
    return p


def write_vxa(problem: dict, compression: str = "ASCII_READABLE", stop=None, fitness=None, record=None) -> str:
    """One self-contained VXA describing ``problem``."""
    p = _full(problem)
    vxc = _el("VXC", _el("Lattice", _el("Lattice_Dim", _num(p["lattice_dim"]))) + _palette(p["materials"])
              + _structure(p, compression), ' Version="0.94"')
    return ('<?xml version="1.0" encoding="ISO-8859-1"?>\n'
            + _el("VXA", _simulator(p, stop, fitness, record) + _environment(p) + vxc, ' Version="1.1"'))


def write_pair(problem: dict, compression: str = "ASCII_READABLE", stop=None, fitness=None, record=None):
    """(base VXA, VXD): the base carries the simulator, environment and palette;
    the VXD replaces the structure (the robot) through ``replace=``."""
    p = _full(problem)
    empty = dict(p, dims=(1, 1, 1), cells=[0], phase=None)
    vxc = _el("VXC", _el("Lattice", _el("Lattice_Dim", _num(p["lattice_dim"]))) + _palette(p["materials"])
              + _structure(empty, compression), ' Version="0.94"')
    base = _el("VXA", _simulator(p, stop, fitness, record) + _environment(p) + vxc, ' Version="1.1"')
    vxd = _el("VXD", _structure(p, compression, ' replace="VXA.VXC.Structure"'))
    return base, vxd
