"""ctypes front end of the FP64 oracle (vx3_oracle.c).

TEST INFRASTRUCTURE ONLY.  The oracle is the checker for the HIP path; only
tests/, ``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py``
may import this module.  It never runs inside the product path.

A problem is a plain dict (see ``SPEC_KEYS``) holding the lattice, palette and
simulator/environment settings, i.e. the information a VXA file carries.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")

VXO_CPU_SEMANTICS = 1
VXO_NO_TEMPERATURE = 2
VXO_ALL_PAIRS = 4


class _Material(C.Structure):
    _fields_ = [("E", C.c_float), ("rho", C.c_float), ("nu", C.c_float), ("cte", C.c_float),
                ("mu_s", C.c_float), ("mu_k", C.c_float), ("fail_stress", C.c_float),
                ("plastic_mod", C.c_float), ("yield_stress", C.c_float), ("model", C.c_int),
                ("fixed", C.c_int), ("is_measured", C.c_int), ("thermal_on_after", C.c_double),
                ("cilia", C.c_double), ("cilia_on_after", C.c_double), ("remove_after", C.c_double),
                ("ss_points", C.c_int), ("ss_strain", C.c_float * 32), ("ss_stress", C.c_float * 32),
                ("is_pacemaker", C.c_int), ("pacemaker_period", C.c_double), ("signal_decay", C.c_double),
                ("signal_delay", C.c_double), ("inactive_period", C.c_double), ("is_target", C.c_int),
                ("sticky", C.c_int)]


# MathTree token codes (vx3_oracle.h VXO_MT_*) and variable order
MT_CODES = {name: i for i, name in enumerate(
    ["mtEND", "mtCONST", "mtE", "mtPI", "mtVAR", "mtADD", "mtSUB", "mtMUL", "mtDIV", "mtPOW", "mtSQRT", "mtSIN",
     "mtCOS", "mtTAN", "mtATAN", "mtLOG", "mtINT", "mtABS", "mtNOT", "mtGREATERTHAN", "mtLESSTHAN", "mtAND", "mtOR",
     "mtNORMALCDF"])}
MT_VARS = ["x", "y", "z", "hit", "t", "angle", "targetCloseness", "numClosePairs", "num_voxel"]


def material_model(m: dict) -> int:
    """MatModel of a material dict: 3 (stress-strain data) with "ss_data"
    (strains, stresses), 2 (bilinear) with a yield stress, 1 (linear with
    failure) with a positive failure stress, else 0 (linear)."""
    if m.get("ss_data") is not None:
        return 3
    if float(m.get("yield_stress", 0.0)) > 0:
        return 2
    return 1 if float(m.get("fail_stress", -1.0)) > 0 else 0


class _Spec(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("cells", C.POINTER(C.c_int8)), ("phase", C.POINTER(C.c_double)),
                ("voxel_E", C.POINTER(C.c_float)), ("nmat", C.c_int), ("mats", C.POINTER(_Material)),
                ("lattice_dim", C.c_double), ("dt_frac", C.c_double),
                ("bond_damp_z", C.c_float), ("col_damp_z", C.c_float), ("slow_damp_z", C.c_float),
                ("volume_effects", C.c_int), ("enable_collision", C.c_int), ("enable_expansion", C.c_int),
                ("grav_acc", C.c_float), ("floor_enabled", C.c_int),
                ("temp_enabled", C.c_int), ("vary_temp", C.c_int),
                ("temp_base", C.c_double), ("temp_amplitude", C.c_double), ("temp_period", C.c_double),
                ("enable_cilia", C.c_int), ("secondary_experiment", C.c_int),
                ("cilia_base", C.POINTER(C.c_double)), ("cilia_shift", C.POINTER(C.c_double)),
                ("enable_signals", C.c_int), ("enable_detach", C.c_int),
                ("enable_attach", C.c_int), ("safety_guard", C.c_int), ("max_dist_pairs", C.c_double)]


class _DiffOpts(C.Structure):
    _fields_ = [("floor_k", C.c_double), ("floor_w", C.c_double), ("fric_v", C.c_double),
                ("floor_enabled", C.c_int), ("vary_temp", C.c_int), ("enable_expansion", C.c_int),
                ("temp_amplitude", C.c_double), ("temp_period", C.c_double), ("dt", C.c_double)]


SPEC_DEFAULTS = dict(lattice_dim=0.01, dt_frac=0.9, bond_damp_z=0.1, col_damp_z=1.0, slow_damp_z=1.0,
                     volume_effects=0, enable_collision=1, enable_expansion=0, grav_acc=-9.81,
                     floor_enabled=1, temp_enabled=1, vary_temp=0, temp_base=25.0,
                     temp_amplitude=0.0, temp_period=0.1, phase=None, voxel_E=None,
                     enable_cilia=0, secondary_experiment=0, cilia_base=None, cilia_shift=None, enable_signals=0,
                     enable_detach=0, enable_attach=0, safety_guard=500, max_dist_pairs=0.0)
MATERIAL_DEFAULTS = dict(E=1e6, rho=1000.0, nu=0.0, cte=0.0, mu_s=0.0, mu_k=0.0, fail_stress=-1.0,
                         plastic_mod=0.0, yield_stress=0.0,
                         fixed=0, is_measured=1, thermal_on_after=0.0,
                         cilia=0.0, cilia_on_after=0.0, remove_after=0.0, ss_data=None,
                         is_pacemaker=0, pacemaker_period=0.0, signal_decay=0.9, signal_delay=0.03,
                         inactive_period=0.03, is_target=0, sticky=0)


def build_library(force: bool = False) -> str:
    """Compile build/liboracle.so with make (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)
    return _LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        build_library()
        L = C.CDLL(_LIB_PATH)
        vp, dp, fp, ip = C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_float), C.POINTER(C.c_int)
        L.vxo_build.restype = vp
        L.vxo_build.argtypes = [C.POINTER(_Spec)]
        L.vxo_free.argtypes = [vp]
        for f in ("vxo_num_voxels", "vxo_num_links", "vxo_num_surface"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [vp]
        L.vxo_time.restype = C.c_double
        L.vxo_time.argtypes = [vp]
        L.vxo_recommended_dt.restype = C.c_double
        L.vxo_recommended_dt.argtypes = [vp]
        L.vxo_step_dt.restype = C.c_float
        L.vxo_step_dt.argtypes = [vp]
        L.vxo_step.restype = C.c_int
        L.vxo_step.argtypes = [vp, C.c_float, C.c_int, C.c_int]
        L.vxo_update_links.argtypes = [vp, dp]
        L.vxo_get_voxels.argtypes = [vp, dp, dp, dp, dp]
        L.vxo_set_voxels.argtypes = [vp, dp, dp, dp, dp]
        L.vxo_get_flags.argtypes = [vp, ip, ip]
        L.vxo_get_removed.argtypes = [vp, ip, ip]
        L.vxo_get_signals.argtypes = [vp, dp]
        L.vxo_collision_count.restype = C.c_long
        L.vxo_collision_count.argtypes = [vp]
        L.vxo_contact_pairs.restype = C.c_int
        L.vxo_contact_pairs.argtypes = [vp, ip, C.c_int]
        L.vxo_get_link_state.argtypes = [vp, dp, fp]
        L.vxo_get_temperature.argtypes = [vp, fp, dp]
        L.vxo_set_temperature.argtypes = [vp, fp]
        L.vxo_get_topology.argtypes = [vp, ip, ip, ip, ip, ip]
        L.vxo_center_of_mass.argtypes = [vp, dp]
        L.vxo_get_contact.argtypes = [vp, dp]
        L.vxo_set_external.argtypes = [vp, C.c_int, C.c_int, dp, dp, fp, fp]
        L.vxo_set_force_field.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), dp]
        L.vxo_set_attach_condition.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), dp]
        L.vxo_get_attach.argtypes = [vp, ip, ip]
        L.vxo_get_samples.argtypes = [vp, dp]
        L.vxo_get_link_curve.restype = C.c_int
        L.vxo_get_link_curve.argtypes = [vp, C.c_int, fp, fp]
        L.vxo_get_voxel_material.argtypes = [vp, C.c_int, fp]
        L.vxo_get_link_material.argtypes = [vp, C.c_int, fp]
        L.vxo_diff_new.restype = vp
        L.vxo_diff_new.argtypes = [vp, dp, dp, C.POINTER(_DiffOpts)]
        L.vxo_diff_free.argtypes = [vp]
        L.vxo_diff_step.argtypes = [vp, C.c_int]
        L.vxo_diff_get.argtypes = [vp, dp, dp, dp, dp]
        L.vxo_diff_mass.restype = C.c_double
        L.vxo_diff_mass.argtypes = [vp, C.c_int]
        L.vxo_diff_set.argtypes = [vp, dp, dp, dp, dp]
        for f in ("vxo_sin", "vxo_cos", "vxo_acos"):
            getattr(L, f).restype = C.c_double
            getattr(L, f).argtypes = [C.c_double]
        _lib = L
    return _lib


def _ptr(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class OracleModel:
    """One model advanced by the FP64 oracle."""

    def __init__(self, problem: dict):
        p = dict(SPEC_DEFAULTS)
        p.update(problem)
        nx, ny, nz = p["dims"]
        self._cells = np.ascontiguousarray(p["cells"], dtype=np.int8).reshape(-1)
        assert self._cells.size == nx * ny * nz
        mats = [dict(MATERIAL_DEFAULTS, **m) for m in p["materials"]]
        self._mats = (_Material * max(1, len(mats)))()
        for i, m in enumerate(mats):
            model = material_model(m)
            ss = m["ss_data"] if m["ss_data"] is not None else ((), ())
            npts = len(ss[0])
            self._mats[i] = _Material(m["E"], m["rho"], m["nu"], m["cte"], m["mu_s"], m["mu_k"],
                                      m["fail_stress"] if model else -1.0, m["plastic_mod"], m["yield_stress"],
                                      model, int(m["fixed"]), int(m["is_measured"]), m["thermal_on_after"],
                                      m["cilia"], m["cilia_on_after"], m["remove_after"], npts,
                                      (C.c_float * 32)(*ss[0]), (C.c_float * 32)(*ss[1]), int(m["is_pacemaker"]),
                                      m["pacemaker_period"], m["signal_decay"], m["signal_delay"], m["inactive_period"],
                                      int(m["is_target"]), int(m["sticky"]))
        self._phase = None if p["phase"] is None else np.ascontiguousarray(p["phase"], dtype=np.float64)
        self._E = None if p["voxel_E"] is None else np.ascontiguousarray(p["voxel_E"], dtype=np.float32)
        self._cb, self._cs = [None if p[k] is None else np.ascontiguousarray(p[k], dtype=np.float64).reshape(-1)
                              for k in ("cilia_base", "cilia_shift")]
        spec = _Spec(nx, ny, nz, _ptr(self._cells, C.c_int8), _ptr(self._phase, C.c_double),
                     _ptr(self._E, C.c_float), len(mats), self._mats, p["lattice_dim"], p["dt_frac"],
                     p["bond_damp_z"], p["col_damp_z"], p["slow_damp_z"], int(p["volume_effects"]),
                     int(p["enable_collision"]), int(p["enable_expansion"]), p["grav_acc"],
                     int(p["floor_enabled"]), int(p["temp_enabled"]), int(p["vary_temp"]),
                     p["temp_base"], p["temp_amplitude"], p["temp_period"],
                     int(p["enable_cilia"]), int(p["secondary_experiment"]),
                     _ptr(self._cb, C.c_double), _ptr(self._cs, C.c_double), int(p["enable_signals"]),
                     int(p["enable_detach"]), int(p["enable_attach"]), int(p["safety_guard"]),
                     float(p["max_dist_pairs"]))
        self._L = lib()
        self._h = self._L.vxo_build(C.byref(spec))
        self.nv = self._L.vxo_num_voxels(self._h)
        self.ns = self._L.vxo_num_surface(self._h)
        for e in p.get("externals") or []:
            self.set_external(**e)
        for axis, expr in (p.get("force_field") or {}).items():
            self.set_force_field("xyz".index(axis), expr)
        for k, expr in enumerate(p.get("attach_conditions") or []):
            if expr is not None:
                self.set_attach_condition(k, expr)

    @property
    def nl(self) -> int:
        """Links so far (attach adds links as the model steps)."""
        return self._L.vxo_num_links(self._h)

    @staticmethod
    def _tokens(expr):
        """A formula as a nested tuple, e.g.
        ("mtMUL", ("mtCONST", 0.01), ("mtSIN", ("mtVAR", "t"))), linearised in
        the reference's token order: breadth-first over the formula's subtree,
        reversed, END last (VX3_SimulationManager::ParseMathTree)."""
        bfs = [expr]
        for node in bfs:
            if node[0] not in ("mtVAR", "mtCONST"):
                bfs.extend(node[1:])
        ops, vals = [], []
        for node in reversed(bfs):
            ops.append(MT_CODES[node[0]])
            if node[0] == "mtVAR":
                vals.append(float(MT_VARS.index(node[1])))
            elif node[0] == "mtCONST":
                vals.append(float(node[1]))
            else:
                vals.append(0.0)
        ops.append(MT_CODES["mtEND"])
        vals.append(0.0)
        return np.asarray(ops, dtype=np.int32), np.asarray(vals, dtype=np.float64)

    def set_force_field(self, axis: int, expr) -> None:
        """Force-field formula of one axis (a nested tuple, see _tokens)."""
        o, v = self._tokens(expr)
        self._L.vxo_set_force_field(self._h, int(axis), len(o), _ptr(o, C.c_int), _ptr(v, C.c_double))

    def set_attach_condition(self, k: int, expr) -> None:
        """AttachCondition Condition_k (a nested tuple, see _tokens)."""
        o, v = self._tokens(expr)
        self._L.vxo_set_attach_condition(self._h, int(k), len(o), _ptr(o, C.c_int), _ptr(v, C.c_double))

    def samples(self) -> dict:
        """The last CycleStep samples: recentAngle, targetCloseness,
        numClosePairs, angleSampleTimes and the sampled centre of mass."""
        o = np.zeros(7)
        self._L.vxo_get_samples(self._h, _ptr(o, C.c_double))
        return dict(angle=o[0], closeness=o[1], pairs=int(o[2]), count=int(o[3]), com=o[4:7].copy())

    def attach_state(self):
        """Per voxel the enableAttach flag, per link the updates left whose
        forces are scaled by 0.01."""
        vf = np.zeros(self.nv, dtype=np.int32); ln = np.zeros(max(1, self.nl), dtype=np.int32)
        self._L.vxo_get_attach(self._h, _ptr(vf, C.c_int), _ptr(ln, C.c_int))
        return vf, ln[:self.nl]

    def link_curve(self, l: int) -> np.ndarray:
        st, ss = np.zeros(72, dtype=np.float32), np.zeros(72, dtype=np.float32)
        n = self._L.vxo_get_link_curve(self._h, int(l), _ptr(st, C.c_float), _ptr(ss, C.c_float))
        return np.stack([st[:n], ss[:n]], axis=1)

    def set_external(self, voxel: int, dof: int = 0, translation=(0, 0, 0), rotation=(0, 0, 0),
                     force=(0, 0, 0), moment=(0, 0, 0)) -> None:
        """External of a voxel: fixed degrees of freedom (bits 1,2,4 = x,y,z
        translation, 8,16,32 = x,y,z rotation), prescribed translation / rotation
        vector, force and moment."""
        t = np.ascontiguousarray(translation, dtype=np.float64)
        r = np.ascontiguousarray(rotation, dtype=np.float64)
        f = np.ascontiguousarray(force, dtype=np.float32)
        m = np.ascontiguousarray(moment, dtype=np.float32)
        self._L.vxo_set_external(self._h, int(voxel), int(dof), _ptr(t, C.c_double), _ptr(r, C.c_double),
                                 _ptr(f, C.c_float), _ptr(m, C.c_float))

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.vxo_free(self._h)
            self._h = None

    @property
    def time(self) -> float:
        return self._L.vxo_time(self._h)

    def recommended_dt(self) -> float:
        return self._L.vxo_recommended_dt(self._h)

    def step_dt(self) -> float:
        return self._L.vxo_step_dt(self._h)

    def step(self, nsteps: int = 1, dt: float = -1.0, flags: int = 0) -> int:
        return self._L.vxo_step(self._h, dt, nsteps, flags)

    def update_links(self) -> np.ndarray:
        out = np.zeros((self.nl, 12), dtype=np.float64)
        self._L.vxo_update_links(self._h, _ptr(out, C.c_double))
        return out

    def state(self):
        pos = np.zeros((self.nv, 3)); q = np.zeros((self.nv, 4))
        lm = np.zeros((self.nv, 3)); am = np.zeros((self.nv, 3))
        self._L.vxo_get_voxels(self._h, _ptr(pos, C.c_double), _ptr(q, C.c_double),
                               _ptr(lm, C.c_double), _ptr(am, C.c_double))
        return dict(pos=pos, quat=q, linmom=lm, angmom=am)

    def set_state(self, pos=None, quat=None, linmom=None, angmom=None):
        c = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (pos, quat, linmom, angmom)]
        self._L.vxo_set_voxels(self._h, *[_ptr(a, C.c_double) for a in c])

    def flags(self):
        vf = np.zeros(self.nv, dtype=np.int32); lf = np.zeros(self.nl, dtype=np.int32)
        self._L.vxo_get_flags(self._h, _ptr(vf, C.c_int), _ptr(lf, C.c_int))
        return vf, lf

    def removed(self):
        """Per voxel and per link: the step (1-based) at whose end it was
        removed, 0 while present."""
        vr = np.zeros(self.nv, dtype=np.int32); lr = np.zeros(max(1, self.nl), dtype=np.int32)
        self._L.vxo_get_removed(self._h, _ptr(vr, C.c_int), _ptr(lr, C.c_int))
        return vr, lr[:self.nl]

    def collision_count(self) -> int:
        return int(self._L.vxo_collision_count(self._h))

    def signals(self) -> np.ndarray:
        """Each voxel's localSignal."""
        o = np.zeros(self.nv)
        self._L.vxo_get_signals(self._h, _ptr(o, C.c_double))
        return o

    def contact(self) -> np.ndarray:
        """Each voxel's contact force of the last step."""
        o = np.zeros((self.nv, 3))
        self._L.vxo_get_contact(self._h, _ptr(o, C.c_double))
        return o

    def contact_pairs(self) -> np.ndarray:
        """(voxel1, voxel2) of every pair in contact in the last step, in the
        order the oracle applied them (ascending voxel1 surface index, then
        ascending voxel2)."""
        n = self._L.vxo_contact_pairs(self._h, None, 0)
        out = np.zeros((n, 2), dtype=np.int32)
        if n:
            self._L.vxo_contact_pairs(self._h, _ptr(out, C.c_int), n)
        return out

    def link_state(self):
        s = np.zeros((self.nl, 9)); st = np.zeros(self.nl, dtype=np.float32)
        self._L.vxo_get_link_state(self._h, _ptr(s, C.c_double), _ptr(st, C.c_float))
        return s, st

    def temperature(self):
        t = np.zeros(self.nv, dtype=np.float32); r = np.zeros(self.nl)
        self._L.vxo_get_temperature(self._h, _ptr(t, C.c_float), _ptr(r, C.c_double))
        return t, r

    def set_temperature(self, t):
        t = np.ascontiguousarray(t, dtype=np.float32)
        self._L.vxo_set_temperature(self._h, _ptr(t, C.c_float))

    def topology(self):
        lat = np.zeros((self.nv, 3), dtype=np.int32)
        ln = np.zeros(self.nl, dtype=np.int32); lp = np.zeros(self.nl, dtype=np.int32)
        self.ns = self._L.vxo_num_surface(self._h)
        la = np.zeros(self.nl, dtype=np.int32); sf = np.zeros(self.ns, dtype=np.int32)
        self._L.vxo_get_topology(self._h, *[_ptr(a, C.c_int) for a in (lat, ln, lp, la, sf)])
        return dict(lattice=lat, link_neg=ln, link_pos=lp, link_axis=la, surface=sf)

    def com(self) -> np.ndarray:
        c = np.zeros(3)
        self._L.vxo_center_of_mass(self._h, _ptr(c, C.c_double))
        return c

    def voxel_material(self, v: int) -> np.ndarray:
        o = np.zeros(12, dtype=np.float32)
        self._L.vxo_get_voxel_material(self._h, v, _ptr(o, C.c_float))
        return o

    def link_material(self, l: int) -> np.ndarray:
        o = np.zeros(11, dtype=np.float32)
        self._L.vxo_get_link_material(self._h, l, _ptr(o, C.c_float))
        return o


class OracleDiff:
    """FP64 oracle of the differentiable mode (softplus floor, per-voxel E and phase)."""

    def __init__(self, problem: dict, phase, E, dt: float, floor_k: float = 1e5, floor_w: float = 1e-4,
                 fric_v: float = 1e-3):
        p = dict(SPEC_DEFAULTS)
        p.update(problem)
        self.model = OracleModel(problem)
        self._phase = np.ascontiguousarray(phase, dtype=np.float64)
        self._E = np.ascontiguousarray(E, dtype=np.float64)
        vary = int(p["vary_temp"]) and p["temp_period"] > 0
        o = _DiffOpts(floor_k, floor_w, fric_v, int(p["floor_enabled"]), int(vary), int(p["enable_expansion"]),
                      p["temp_amplitude"], p["temp_period"], dt)
        self._L = lib()
        self._h = self._L.vxo_diff_new(self.model._h, _ptr(self._phase, C.c_double), _ptr(self._E, C.c_double),
                                       C.byref(o))
        self.nv = self.model.nv

    def __del__(self):
        if getattr(self, "_h", None):
            self._L.vxo_diff_free(self._h)
            self._h = None

    def step(self, n: int = 1):
        self._L.vxo_diff_step(self._h, int(n))

    def state(self):
        pos = np.zeros((self.nv, 3)); q = np.zeros((self.nv, 4))
        lm = np.zeros((self.nv, 3)); am = np.zeros((self.nv, 3))
        self._L.vxo_diff_get(self._h, *[_ptr(a, C.c_double) for a in (pos, q, lm, am)])
        return dict(pos=pos, quat=q, linmom=lm, angmom=am)

    def set_state(self, pos=None, quat=None, linmom=None, angmom=None):
        c = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (pos, quat, linmom, angmom)]
        self._L.vxo_diff_set(self._h, *[_ptr(a, C.c_double) for a in c])

    def masses(self) -> np.ndarray:
        return np.array([self._L.vxo_diff_mass(self._h, v) for v in range(self.nv)])
