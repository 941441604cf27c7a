"""Python binding of the MI355X-native VX3 timestep (libvx3mi.so, include/vx3mi.h).

The binding is a thin ctypes layer over the C ABI.  It never falls back to a
CPU implementation: if the HIP library is missing, ``lib()`` raises.

A *problem* is a plain dict describing one robot the way a merged VXA file
does (lattice dims, palette, per-voxel phase, simulator and environment
settings); see ``PROBLEM_DEFAULTS`` and ``MATERIAL_DEFAULTS``.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Iterable, Sequence

import numpy as np

__all__ = ["lib", "Model", "Batch", "VX3MIError", "FP32", "FP64", "MODE_AUTO", "MODE_PLAIN", "MODE_FLAT",
           "MODE_PERSISTENT", "MODE_TILED", "Diff", "run", "FORMULA_STOP", "FORMULA_FITNESS", "PROBLEM_DEFAULTS", "MATERIAL_DEFAULTS", "library_path"]

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# the in-tree build; VX3MI_LIB names another build of the same library
_LIB_PATH = os.environ.get("VX3MI_LIB") or os.path.join(_PKG, "lib", "libvx3mi.so")

FP32, FP64 = 32, 64
MODE_AUTO, MODE_PLAIN, MODE_FLAT, MODE_PERSISTENT, MODE_TILED = 0, 1, 2, 3, 4
# fixed degrees of freedom of a voxel external (CVX_External dofComponent bits)
DOF_X, DOF_Y, DOF_Z, DOF_TX, DOF_TY, DOF_TZ, DOF_ALL = 1, 2, 4, 8, 16, 32, 63
OK, DIVERGED = 0, 1

PROBLEM_DEFAULTS = dict(lattice_dim=0.01, dt_frac=0.9, bond_damp_z=0.1, col_damp_z=1.0, slow_damp_z=1.0,
                        volume_effects=0, enable_collision=1, enable_expansion=0, grav_acc=-9.81,
                        floor_enabled=1, temp_enabled=1, vary_temp=0, temp_base=25.0,
                        temp_amplitude=0.0, temp_period=0.1, phase=None, voxel_E=None,
                        enable_cilia=0, secondary_experiment=0, cilia_base=None, cilia_shift=None, enable_signals=0,
                        enable_detach=0, max_dist_pairs=0.0, enable_attach=0, safety_guard=500)
MATERIAL_DEFAULTS = dict(E=1e6, rho=1000.0, nu=0.0, cte=0.0, mu_s=0.0, mu_k=0.0, fail_stress=-1.0,
                         plastic_mod=0.0, yield_stress=0.0,
                         fixed=0, is_measured=1, is_target=0, thermal_on_after=0.0,
                         cilia=0.0, cilia_on_after=0.0, remove_after=0.0, ss_data=None,
                         is_pacemaker=0, pacemaker_period=0.0, signal_decay=0.9, signal_delay=0.03,
                         inactive_period=0.03, rgba=(0.5, 0.5, 0.5, 1.0), sticky=0)


class VX3MIError(RuntimeError):
    pass


def material_model(m: dict) -> int:
    """MatModel of a material dict: 3 (stress-strain data) with "ss_data"
    (strains, stresses), 2 (bilinear) with a yield stress, 1 (linear with
    failure) with a positive failure stress, else 0 (linear)."""
    if m.get("ss_data") is not None:
        return 3
    if float(m.get("yield_stress", 0.0)) > 0:
        return 2
    return 1 if float(m.get("fail_stress", -1.0)) > 0 else 0


class _Material(C.Structure):
    _fields_ = [("elastic_mod", C.c_float), ("density", C.c_float), ("poissons_ratio", C.c_float),
                ("cte", C.c_float), ("u_static", C.c_float), ("u_dynamic", C.c_float),
                ("fail_stress", C.c_float), ("plastic_mod", C.c_float), ("yield_stress", C.c_float),
                ("mat_model", C.c_int), ("fixed", C.c_int),
                ("is_measured", C.c_int), ("is_target", C.c_int), ("thermal_on_after", C.c_double),
                ("red", C.c_float), ("green", C.c_float), ("blue", C.c_float), ("alpha", C.c_float),
                ("cilia", C.c_double), ("cilia_on_after", C.c_double), ("remove_after", C.c_double),
                ("ss_points", C.c_int), ("ss_strain", C.c_float * 32), ("ss_stress", C.c_float * 32),
                ("is_pacemaker", C.c_int), ("pacemaker_period", C.c_double), ("signal_value_decay", C.c_double),
                ("signal_time_delay", C.c_double), ("inactive_period", C.c_double), ("sticky", C.c_int)]


class _Settings(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("lattice_dim", C.c_double),
                ("dt_frac", C.c_double), ("floor_enabled", C.c_int), ("enable_collision", C.c_int),
                ("enable_expansion", C.c_int), ("vary_temp", C.c_int), ("temp_base", C.c_double),
                ("temp_amplitude", C.c_double), ("temp_period", C.c_double), ("initial_temp", C.c_float),
                ("record_step_size", C.c_int), ("record_voxel", C.c_int), ("record_link", C.c_int),
                ("has_stop_formula", C.c_int), ("has_fitness_formula", C.c_int),
                ("enable_cilia", C.c_int), ("secondary_experiment", C.c_int), ("enable_signals", C.c_int),
                ("enable_detach", C.c_int), ("max_dist_pairs", C.c_double), ("enable_attach", C.c_int),
                ("safety_guard", C.c_int)]


FORMULA_STOP, FORMULA_FITNESS = 0, 1
FORMULA_VARS = ("x", "y", "z", "hit", "t", "angle", "targetCloseness", "numClosePairs", "num_voxel")


class _DiffOpts(C.Structure):
    _fields_ = [("floor_stiffness", C.c_double), ("floor_width", C.c_double), ("friction_speed", C.c_double),
                ("dt", C.c_double), ("e_max", C.c_double), ("checkpoint_every", C.c_int),
                ("hbm_budget", C.c_double), ("reverse_threads", C.c_int)]


class _Result(C.Structure):
    _fields_ = [("current_time", C.c_double), ("fitness_score", C.c_double), ("initial_com", C.c_double * 3),
                ("current_com", C.c_double * 3), ("total_distance_of_all_voxels", C.c_double),
                ("vox_size", C.c_double), ("recommended_dt", C.c_double), ("num_voxel", C.c_int),
                ("num_measured_voxel", C.c_int), ("num_close_pairs", C.c_int), ("diverged", C.c_int),
                ("angle_sample_times", C.c_int), ("real_stepsize", C.c_int), ("steps", C.c_longlong)]


_TEXT_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_char), C.c_size_t)


class _Lattice(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int), ("cells", C.POINTER(C.c_int8)),
                ("phase", C.POINTER(C.c_double)), ("voxel_E", C.POINTER(C.c_float)), ("nmat", C.c_int),
                ("mats", C.POINTER(_Material)), ("lattice_dim", C.c_double), ("dt_frac", C.c_double),
                ("bond_damp_z", C.c_float), ("col_damp_z", C.c_float), ("slow_damp_z", C.c_float),
                ("volume_effects", C.c_int), ("enable_collision", C.c_int), ("enable_expansion", C.c_int),
                ("grav_acc", C.c_float), ("floor_enabled", C.c_int), ("temp_enabled", C.c_int),
                ("vary_temp", C.c_int), ("temp_base", C.c_double), ("temp_amplitude", C.c_double),
                ("temp_period", C.c_double), ("enable_cilia", C.c_int), ("secondary_experiment", C.c_int),
                ("cilia_base", C.POINTER(C.c_double)), ("cilia_shift", C.POINTER(C.c_double)),
                ("enable_signals", C.c_int), ("enable_detach", C.c_int), ("max_dist_pairs", C.c_double),
                ("enable_attach", C.c_int), ("safety_guard", C.c_int)]


def library_path() -> str:
    return _LIB_PATH


_lib = None


def lib():
    """Load libvx3mi.so (built in-tree by ``make -C voxcraft-sim_amd``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise VX3MIError(f"HIP library not built: {_LIB_PATH} (run make -C voxcraft-sim_amd)")
    L = C.CDLL(_LIB_PATH)
    vp, ip, dp, fp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_float)
    pvp = C.POINTER(C.c_void_p)
    sigs = {
        "vx3mi_last_error": (C.c_char_p, []),
        "vx3mi_version": (C.c_char_p, []),
        "vx3mi_model_from_vxa": (C.c_int, [C.c_char_p, C.c_char_p, pvp]),
        "vx3mi_model_from_xml": (C.c_int, [C.c_char_p, C.c_char_p, pvp]),
        "vx3mi_model_from_lattice": (C.c_int, [C.POINTER(_Lattice), pvp]),
        "vx3mi_model_free": (None, [vp]),
        "vx3mi_model_counts": (C.c_int, [vp, ip, ip, ip]),
        "vx3mi_model_settings": (C.c_int, [vp, C.POINTER(_Settings)]),
        "vx3mi_model_phase": (C.c_int, [vp, dp]),
        "vx3mi_model_eval_formula": (C.c_int, [vp, C.c_int, dp, dp]),
        "vx3mi_model_recommended_dt": (C.c_int, [vp, dp]),
        "vx3mi_model_topology": (C.c_int, [vp, ip, ip, ip, ip]),
        "vx3mi_model_voxel_constants": (C.c_int, [vp, C.c_int, fp]),
        "vx3mi_model_link_constants": (C.c_int, [vp, C.c_int, fp]),
        "vx3mi_model_link_curve": (C.c_int, [vp, C.c_int, fp, fp]),
        "vx3mi_model_set_force_field": (C.c_int, [vp, C.c_int, C.c_char_p]),
        "vx3mi_model_set_attach_condition": (C.c_int, [vp, C.c_int, C.c_char_p]),
        "vx3mi_batch_attach_state": (C.c_int, [vp, C.c_int, ip, ip, ip, ip, ip, C.c_int, ip]),
        "vx3mi_model_set_external": (C.c_int, [vp, C.c_int, C.c_int, dp, dp, dp, dp]),
        "vx3mi_model_get_external": (C.c_int, [vp, C.c_int, ip, dp, dp, dp, dp]),
        "vx3mi_batch_create": (C.c_int, [pvp, C.c_int, C.c_int, pvp]),
        "vx3mi_batch_destroy": (None, [vp]),
        "vx3mi_batch_counts": (C.c_int, [vp, ip, ip, ip]),
        "vx3mi_batch_recommended_dt": (C.c_int, [vp, dp]),
        "vx3mi_step": (C.c_int, [vp, C.c_int, C.c_float, C.c_int]),
        "vx3mi_batch_sync": (C.c_int, [vp]),
        "vx3mi_get_state": (C.c_int, [vp, C.c_int, dp, dp, dp, dp]),
        "vx3mi_set_state": (C.c_int, [vp, C.c_int, dp, dp, dp, dp]),
        "vx3mi_get_voxel_flags": (C.c_int, [vp, C.c_int, ip]),
        "vx3mi_get_link_flags": (C.c_int, [vp, C.c_int, ip, fp]),
        "vx3mi_batch_removed": (C.c_int, [vp, C.c_int, ip, ip]),
        "vx3mi_batch_signals": (C.c_int, [vp, C.c_int, dp]),
        "vx3mi_batch_hits": (C.c_int, [vp, ip]),
        "vx3mi_model_cilia": (C.c_int, [vp, dp, dp]),
        "vx3mi_model_material": (C.c_int, [vp, C.c_int, C.POINTER(_Material)]),
        "vx3mi_batch_record_contacts": (C.c_int, [vp, C.c_int]),
        "vx3mi_batch_record_contacts_static": (C.c_int, [vp, C.c_int]),
        "vx3mi_batch_listings": (C.c_int, [vp, ip]),
        "vx3mi_batch_contact_pairs": (C.c_int, [vp, C.c_int, ip, C.c_int]),
        "vx3mi_batch_time": (C.c_int, [vp, dp]),
        "vx3mi_batch_diverged": (C.c_int, [vp, ip]),
        "vx3mi_center_of_mass": (C.c_int, [vp, dp]),
        "vx3mi_eval_link_forces": (C.c_int, [vp, dp]),
        "vx3mi_last_step_ms": (C.c_int, [vp, fp]),
        "vx3mi_last_step_launches": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
        "vx3mi_batch_autotune": (C.c_int, [vp, C.c_int, dp, ip]),
        "vx3mi_batch_configure": (C.c_int, [vp, C.c_int, C.c_int, C.c_int]),
        "vx3mi_run": (C.c_int, [vp, pvp, C.c_int, C.c_longlong, C.c_int, _TEXT_FN, vp, C.POINTER(_Result),
                                C.POINTER(dp)]),
        "vx3mi_model_raw_print": (C.c_char_p, [vp]),
        "vx3mi_model_palette_index": (C.c_int, [vp, ip]),
        "vx3mi_model_save_positions": (C.c_int, [vp]),
        "vx3mi_diff_default_opts": (None, [C.POINTER(_DiffOpts)]),
        "vx3mi_diff_create": (C.c_int, [pvp, C.c_int, C.c_int, C.POINTER(_DiffOpts), pvp]),
        "vx3mi_diff_destroy": (None, [vp]),
        "vx3mi_diff_counts": (C.c_int, [vp, ip, ip, ip]),
        "vx3mi_diff_set_params": (C.c_int, [vp, dp, dp]),
        "vx3mi_diff_forward": (C.c_int, [vp, C.c_int]),
        "vx3mi_diff_get_state": (C.c_int, [vp, C.c_int, dp, dp, dp, dp]),
        "vx3mi_diff_backward": (C.c_int, [vp, dp, dp, dp, dp, dp, dp]),
        "vx3mi_diff_info": (C.c_int, [vp, dp]),
    }
    for name, (res, args) in sigs.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def _check(rc: int, allow=(OK,)) -> int:
    if rc not in allow:
        msg = lib().vx3mi_last_error().decode(errors="replace")
        raise VX3MIError(f"vx3mi error {rc}: {msg}")
    return rc


def _ptr(a, ct):
    return None if a is None else a.ctypes.data_as(C.POINTER(ct))


class Model:
    """Host model of one robot (topology, materials, initial state)."""

    def __init__(self, handle: int, keep=()):
        self._h = C.c_void_p(handle)
        self._keep = keep
        nv, nl, ns = C.c_int(), C.c_int(), C.c_int()
        _check(lib().vx3mi_model_counts(self._h, C.byref(nv), C.byref(nl), C.byref(ns)))
        self.nv, self.nl, self.ns = nv.value, nl.value, ns.value

    @classmethod
    def from_problem(cls, problem: dict) -> "Model":
        p = dict(PROBLEM_DEFAULTS)
        p.update(problem)
        nx, ny, nz = (int(d) for d in p["dims"])
        cells = np.ascontiguousarray(p["cells"], dtype=np.int8).reshape(-1)
        if cells.size != nx * ny * nz:
            raise VX3MIError("cells size does not match dims")
        mats = [dict(MATERIAL_DEFAULTS, **m) for m in p["materials"]]
        marr = (_Material * max(1, len(mats)))()
        for i, m in enumerate(mats):
            r, g, b, a = m["rgba"]
            ss = m["ss_data"] if m["ss_data"] is not None else ((), ())
            if len(ss[0]) > 32 or len(ss[0]) != len(ss[1]):
                raise VX3MIError("ss_data: equal numbers of strains and stresses, at most 32")
            marr[i] = _Material(m["E"], m["rho"], m["nu"], m["cte"], m["mu_s"], m["mu_k"], float(m["fail_stress"]),
                                float(m["plastic_mod"]), float(m["yield_stress"]), material_model(m),
                                int(m["fixed"]), int(m["is_measured"]), int(m["is_target"]), m["thermal_on_after"],
                                r, g, b, a, float(m["cilia"]), float(m["cilia_on_after"]), float(m["remove_after"]),
                                len(ss[0]), (C.c_float * 32)(*ss[0]), (C.c_float * 32)(*ss[1]), int(m["is_pacemaker"]),
                                float(m["pacemaker_period"]), float(m["signal_decay"]), float(m["signal_delay"]),
                                float(m["inactive_period"]), int(m["sticky"]))
        phase = None if p["phase"] is None else np.ascontiguousarray(p["phase"], dtype=np.float64)
        E = None if p["voxel_E"] is None else np.ascontiguousarray(p["voxel_E"], dtype=np.float32)
        cb, cs = [None if p[k] is None else np.ascontiguousarray(p[k], dtype=np.float64).reshape(-1)
                  for k in ("cilia_base", "cilia_shift")]
        spec = _Lattice(nx, ny, nz, _ptr(cells, C.c_int8), _ptr(phase, C.c_double), _ptr(E, C.c_float),
                        len(mats), marr, p["lattice_dim"], p["dt_frac"], p["bond_damp_z"], p["col_damp_z"],
                        p["slow_damp_z"], int(p["volume_effects"]), int(p["enable_collision"]),
                        int(p["enable_expansion"]), p["grav_acc"], int(p["floor_enabled"]),
                        int(p["temp_enabled"]), int(p["vary_temp"]), p["temp_base"], p["temp_amplitude"],
                        p["temp_period"], int(p["enable_cilia"]), int(p["secondary_experiment"]),
                        _ptr(cb, C.c_double), _ptr(cs, C.c_double), int(p["enable_signals"]),
                        int(p["enable_detach"]), float(p["max_dist_pairs"]), int(p["enable_attach"]),
                        int(p["safety_guard"]))
        h = C.c_void_p()
        _check(lib().vx3mi_model_from_lattice(C.byref(spec), C.byref(h)))
        m = cls(h.value)
        for e in p.get("externals") or []:
            m.set_external(**e)
        for axis, expr in (p.get("force_field") or {}).items():
            m.set_force_field(axis, expr)
        for k, expr in enumerate(p.get("attach_conditions") or []):
            if expr is not None:
                m.set_attach_condition(k, expr)
        return m

    def set_attach_condition(self, k: int, expr) -> None:
        """AttachCondition Condition_k (0..4) as a nested tuple (the format of
        vx3mi.vxa.formula_xml) or MathTree XML text; None removes it."""
        from .vxa import formula_xml
        text = "" if expr is None else (expr if isinstance(expr, str) else formula_xml(expr))
        _check(lib().vx3mi_model_set_attach_condition(self._h, int(k), text.encode()))

    def set_force_field(self, axis, expr) -> None:
        """Force-field formula of axis ("x", "y", "z" or 0..2) as a nested tuple
        (the format of vx3mi.vxa.formula_xml) or MathTree XML text; None removes it."""
        from .vxa import formula_xml
        a = "xyz".index(axis) if isinstance(axis, str) else int(axis)
        text = "" if expr is None else (expr if isinstance(expr, str) else formula_xml(expr))
        _check(lib().vx3mi_model_set_force_field(self._h, a, text.encode()))

    @classmethod
    def from_vxa(cls, base_vxa: str, vxd: str | None = None) -> "Model":
        h = C.c_void_p()
        _check(lib().vx3mi_model_from_vxa(base_vxa.encode(), None if vxd is None else vxd.encode(), C.byref(h)))
        return cls(h.value)

    @classmethod
    def from_xml(cls, base_text: str, vxd_text: str | None = None) -> "Model":
        h = C.c_void_p()
        _check(lib().vx3mi_model_from_xml(base_text.encode(), None if vxd_text is None else vxd_text.encode(),
                                          C.byref(h)))
        return cls(h.value)

    def recommended_dt(self) -> float:
        d = C.c_double()
        _check(lib().vx3mi_model_recommended_dt(self._h, C.byref(d)))
        return d.value

    def settings(self) -> dict:
        st = _Settings()
        _check(lib().vx3mi_model_settings(self._h, C.byref(st)))
        return {name: getattr(st, name) for name, _ in _Settings._fields_}

    def cilia(self):
        """(BaseCiliaForce, ShiftCiliaForce), each voxels x 3."""
        b, sh = np.zeros((self.nv, 3)), np.zeros((self.nv, 3))
        _check(lib().vx3mi_model_cilia(self._h, _ptr(b, C.c_double), _ptr(sh, C.c_double)))
        return b, sh

    def material(self, index: int) -> dict:
        """Palette entry `index` (1-based) as the model holds it."""
        o = _Material()
        _check(lib().vx3mi_model_material(self._h, int(index), C.byref(o)))
        return {name: getattr(o, name) for name, _ in _Material._fields_}

    def phase(self) -> np.ndarray:
        o = np.zeros(self.nv)
        _check(lib().vx3mi_model_phase(self._h, _ptr(o, C.c_double)))
        return o

    def eval_formula(self, which: int, **vars_) -> float:
        """Evaluate the stop (FORMULA_STOP) or fitness (FORMULA_FITNESS) formula;
        unnamed variables are 0."""
        v = np.array([float(vars_.get(n, 0.0)) for n in FORMULA_VARS])
        out = C.c_double()
        _check(lib().vx3mi_model_eval_formula(self._h, which, _ptr(v, C.c_double), C.byref(out)))
        return out.value

    def topology(self) -> dict:
        lat = np.zeros((self.nv, 3), dtype=np.int32)
        ln, lp, la = (np.zeros(self.nl, dtype=np.int32) for _ in range(3))
        _check(lib().vx3mi_model_topology(self._h, _ptr(lat, C.c_int), _ptr(ln, C.c_int), _ptr(lp, C.c_int),
                                          _ptr(la, C.c_int)))
        return dict(lattice=lat, link_neg=ln, link_pos=lp, link_axis=la)

    def voxel_constants(self, v: int) -> np.ndarray:
        o = np.zeros(12, dtype=np.float32)
        _check(lib().vx3mi_model_voxel_constants(self._h, int(v), _ptr(o, C.c_float)))
        return o

    def link_constants(self, l: int) -> np.ndarray:
        o = np.zeros(11, dtype=np.float32)
        _check(lib().vx3mi_model_link_constants(self._h, int(l), _ptr(o, C.c_float)))
        return o

    def link_curve(self, l: int) -> np.ndarray:
        """(n, 2) strain / stress points of a non-linear link material (n = 0: linear)."""
        st, ss = np.zeros(72, dtype=np.float32), np.zeros(72, dtype=np.float32)
        n = lib().vx3mi_model_link_curve(self._h, int(l), _ptr(st, C.c_float), _ptr(ss, C.c_float))
        _check(min(n, 0))
        return np.stack([st[:n], ss[:n]], axis=1)

    def set_external(self, voxel: int, dof: int = 0, translation=(0, 0, 0), rotation=(0, 0, 0),
                     force=(0, 0, 0), moment=(0, 0, 0)) -> None:
        """External of a voxel (CVX_Voxel::external()): dof bits 1,2,4 fix x,y,z
        translation and 8,16,32 fix x,y,z rotation (DOF_* constants); translation
        (m) / rotation vector (rad) where the fixed ones are held; force (N) and
        moment (N m).  dof < 0 removes the external."""
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (translation, rotation, force, moment)]
        _check(lib().vx3mi_model_set_external(self._h, int(voxel), int(dof), *[_ptr(x, C.c_double) for x in a]))

    def external(self, voxel: int) -> dict | None:
        """The voxel's external as a dict (None when it has none)."""
        dof = C.c_int()
        a = [np.zeros(3) for _ in range(4)]
        _check(lib().vx3mi_model_get_external(self._h, int(voxel), C.byref(dof),
                                              *[_ptr(x, C.c_double) for x in a]))
        if dof.value < 0:
            return None
        return dict(dof=dof.value, translation=a[0], rotation=a[1], force=a[2], moment=a[3])

    def __del__(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().vx3mi_model_free(self._h)
            self._h = C.c_void_p()


class Batch:
    """A device-resident batch of robots (CSR-packed)."""

    def __init__(self, models: Sequence[Model] | Model, precision: int = FP32):
        if isinstance(models, Model):
            models = [models]
        self.models = list(models)
        arr = (C.c_void_p * len(self.models))(*[m._h.value for m in self.models])
        h = C.c_void_p()
        _check(lib().vx3mi_batch_create(arr, len(self.models), int(precision), C.byref(h)))
        self._h = h
        self.precision = precision
        r, v, l = C.c_int(), C.c_int(), C.c_int()
        _check(lib().vx3mi_batch_counts(self._h, C.byref(r), C.byref(v), C.byref(l)))
        self.nrobot, self.nv, self.nl = r.value, v.value, l.value
        self.voff = np.concatenate([[0], np.cumsum([m.nv for m in self.models])]).astype(np.int64)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().vx3mi_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def step(self, nsteps: int = 1, dt: float = -1.0, mode: int = MODE_AUTO) -> int:
        return _check(lib().vx3mi_step(self._h, int(nsteps), float(dt), int(mode)), allow=(OK, DIVERGED))

    def sync(self):
        _check(lib().vx3mi_batch_sync(self._h))

    def last_step_ms(self) -> float:
        ms = C.c_float()
        _check(lib().vx3mi_last_step_ms(self._h, C.byref(ms)))
        return ms.value

    def last_step_launches(self) -> int:
        n = C.c_longlong()
        _check(lib().vx3mi_last_step_launches(self._h, C.byref(n)))
        return n.value

    def state(self, robot: int = -1) -> dict:
        n = self.nv if robot < 0 else self.models[robot].nv
        pos, lm, am = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 3))
        q = np.zeros((n, 4))
        _check(lib().vx3mi_get_state(self._h, int(robot), _ptr(pos, C.c_double), _ptr(q, C.c_double),
                                     _ptr(lm, C.c_double), _ptr(am, C.c_double)))
        return dict(pos=pos, quat=q, linmom=lm, angmom=am)

    def set_state(self, robot: int = -1, pos=None, quat=None, linmom=None, angmom=None):
        c = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (pos, quat, linmom, angmom)]
        _check(lib().vx3mi_set_state(self._h, int(robot), *[_ptr(a, C.c_double) for a in c]))

    def voxel_flags(self, robot: int = -1) -> np.ndarray:
        n = self.nv if robot < 0 else self.models[robot].nv
        f = np.zeros(n, dtype=np.int32)
        _check(lib().vx3mi_get_voxel_flags(self._h, int(robot), _ptr(f, C.c_int)))
        return f

    def record_contacts(self, cap: int = 1 << 20):
        """Record the pairs in contact during each later step call (cap pairs per
        call; 0 stops recording)."""
        _check(lib().vx3mi_batch_record_contacts(self._h, int(cap)))

    def record_contacts_static(self, cap: int = 1 << 20):
        """record_contacts into the library's static per-GPU queue (allocates nothing)."""
        _check(lib().vx3mi_batch_record_contacts_static(self._h, int(cap)))

    def listings(self) -> np.ndarray:
        """Per robot: the steps whose collision phase rebuilt the candidate lists."""
        out = np.zeros(self.nrobot, dtype=np.int32)
        _check(lib().vx3mi_batch_listings(self._h, _ptr(out, C.c_int)))
        return out

    def contact_pairs(self, robot: int = -1) -> np.ndarray:
        """(n, 2) voxel index pairs in contact in the last step call, smaller
        index first, sorted (robot-local indices, batch indices for robot -1)."""
        n = lib().vx3mi_batch_contact_pairs(self._h, int(robot), None, 0)
        if n < 0:
            _check(n)
        out = np.zeros((max(n, 1), 2), dtype=np.int32)
        if n:
            _check(min(0, lib().vx3mi_batch_contact_pairs(self._h, int(robot), _ptr(out, C.c_int), n)))
        return out[:n]

    def link_flags(self, robot: int = -1):
        """(flags, axial strain) per link: bit 0 local velocity valid, bit 1 small angle."""
        n = self.nl if robot < 0 else self.models[robot].nl
        f, st = np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.float32)
        _check(lib().vx3mi_get_link_flags(self._h, int(robot), _ptr(f, C.c_int), _ptr(st, C.c_float)))
        return f, st

    def removed(self, robot: int = -1):
        """(voxel, link) arrays: the step count at whose end each was removed
        (SecondaryExperiment material removal), 0 while present."""
        nv = self.nv if robot < 0 else self.models[robot].nv
        nl = self.nl if robot < 0 else self.models[robot].nl
        vr, lr = np.zeros(nv, dtype=np.int32), np.zeros(max(1, nl), dtype=np.int32)
        _check(lib().vx3mi_batch_removed(self._h, int(robot), _ptr(vr, C.c_int), _ptr(lr, C.c_int)))
        return vr, lr[:nl]

    def hits(self) -> np.ndarray:
        """collisionCount per robot (target / non-target contacts over the steps)."""
        out = np.zeros(self.nrobot, dtype=np.int32)
        _check(lib().vx3mi_batch_hits(self._h, _ptr(out, C.c_int)))
        return out

    def signals(self, robot: int = -1) -> np.ndarray:
        """Each voxel's localSignal (EnableSignals)."""
        o = np.zeros(self.nv if robot < 0 else self.models[robot].nv)
        _check(lib().vx3mi_batch_signals(self._h, int(robot), _ptr(o, C.c_double)))
        return o

    def time(self) -> np.ndarray:
        t = np.zeros(self.nrobot)
        _check(lib().vx3mi_batch_time(self._h, _ptr(t, C.c_double)))
        return t

    def diverged(self) -> np.ndarray:
        d = np.zeros(self.nrobot, dtype=np.int32)
        _check(lib().vx3mi_batch_diverged(self._h, _ptr(d, C.c_int)), allow=(OK, DIVERGED))
        return d

    def com(self) -> np.ndarray:
        c = np.zeros((self.nrobot, 3))
        _check(lib().vx3mi_center_of_mass(self._h, _ptr(c, C.c_double)))
        return c

    def recommended_dt(self) -> np.ndarray:
        d = np.zeros(self.nrobot)
        _check(lib().vx3mi_batch_recommended_dt(self._h, _ptr(d, C.c_double)))
        return d

    CANDIDATES = ("tiled-768", "tiled-384x2", "persistent-128", "persistent-256", "persistent-512", "flat-graph")

    def autotune(self, trial_steps: int = 100) -> dict:
        """Times every launch configuration on trial steps (state restored after)
        and keeps the fastest for MODE_AUTO."""
        ms = np.zeros(6)
        ch = np.zeros(3, dtype=np.int32)
        _check(lib().vx3mi_batch_autotune(self._h, int(trial_steps), _ptr(ms, C.c_double), _ptr(ch, C.c_int)))
        return dict(ms={k: (None if v < 0 else float(v)) for k, v in zip(self.CANDIDATES, ms)},
                    mode=int(ch[0]), tile=int(ch[1]), threads=int(ch[2]))

    def configure(self, mode: int, tile: int = 0, threads: int = 0):
        _check(lib().vx3mi_batch_configure(self._h, int(mode), int(tile), int(threads)))

    def attach_state(self, robot: int = 0) -> dict:
        """Per voxel the enableAttach flag; the links attach made (robot voxel
        indices: negative end, positive end, axis, updates left at 0.01 force);
        overflow when an attach candidate was lost."""
        nv = self.models[robot].nv
        vf = np.zeros(nv, dtype=np.int32)
        ov = C.c_int()
        n = _check(lib().vx3mi_batch_attach_state(self._h, int(robot), _ptr(vf, C.c_int), None, None, None, None, 0,
                                                  C.byref(ov)), allow=range(0, 1 << 30))
        links = np.zeros((n, 4), dtype=np.int32)
        cols = [np.zeros(max(n, 1), dtype=np.int32) for _ in range(4)]
        if n:
            _check(lib().vx3mi_batch_attach_state(self._h, int(robot), None, *[_ptr(c, C.c_int) for c in cols], n,
                                                  C.byref(ov)), allow=range(0, 1 << 30))
            links = np.stack([c[:n] for c in cols], axis=1)
        return dict(flags=vf, links=links, overflow=int(ov.value))

    def eval_link_forces(self) -> np.ndarray:
        out = np.zeros((self.nl, 12))
        _check(lib().vx3mi_eval_link_forces(self._h, _ptr(out, C.c_double)))
        return out


class Diff:
    """Differentiable batch (softplus floor, per-voxel phase and modulus, adjoint).

    Options (keyword): floor_stiffness, floor_width, friction_speed, dt, e_max,
    checkpoint_every, hbm_budget, reverse_threads (see vx3mi_diff_opts in
    include/vx3mi.h).
    """

    def __init__(self, models: Sequence[Model] | Model, precision: int = FP32, **opts):
        if isinstance(models, Model):
            models = [models]
        self._models = list(models)
        o = _DiffOpts()
        lib().vx3mi_diff_default_opts(C.byref(o))
        for k, v in opts.items():
            if not hasattr(o, k):
                raise TypeError(f"unknown option {k}")
            setattr(o, k, v)
        arr = (C.c_void_p * len(self._models))(*[m._h.value for m in self._models])
        h = C.c_void_p()
        _check(lib().vx3mi_diff_create(arr, len(self._models), precision, C.byref(o), C.byref(h)))
        self._h = h
        nr, nv, nl = C.c_int(), C.c_int(), C.c_int()
        _check(lib().vx3mi_diff_counts(self._h, C.byref(nr), C.byref(nv), C.byref(nl)))
        self.nr, self.nv, self.nl = nr.value, nv.value, nl.value
        self.precision = precision
        self.voxel_offsets = np.cumsum([0] + [m.nv for m in self._models])

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().vx3mi_diff_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        self.close()

    def set_params(self, phase, E):
        ph = np.ascontiguousarray(phase, dtype=np.float64).reshape(-1)
        e = np.ascontiguousarray(E, dtype=np.float64).reshape(-1)
        if ph.size != self.nv or e.size != self.nv:
            raise VX3MIError("phase and E need one value per voxel of the batch")
        _check(lib().vx3mi_diff_set_params(self._h, _ptr(ph, C.c_double), _ptr(e, C.c_double)))

    def forward(self, nsteps: int):
        _check(lib().vx3mi_diff_forward(self._h, int(nsteps)))

    def state(self, robot: int = -1) -> dict:
        n = self.nv if robot < 0 else self._models[robot].nv
        pos, q, lm, am = np.zeros((n, 3)), np.zeros((n, 4)), np.zeros((n, 3)), np.zeros((n, 3))
        _check(lib().vx3mi_diff_get_state(self._h, int(robot), *[_ptr(a, C.c_double) for a in (pos, q, lm, am)]))
        return dict(pos=pos, quat=q, linmom=lm, angmom=am)

    def backward(self, grad_pos=None, grad_quat=None, grad_linmom=None, grad_angmom=None):
        """dL/dphase (per cycle) and dL/dE per voxel from dL/d(final state)."""
        g = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
             for a in (grad_pos, grad_quat, grad_linmom, grad_angmom)]
        gp, ge = np.zeros(self.nv), np.zeros(self.nv)
        _check(lib().vx3mi_diff_backward(self._h, *[_ptr(a, C.c_double) for a in g], _ptr(gp, C.c_double),
                                         _ptr(ge, C.c_double)))
        return gp, ge

    def info(self) -> dict:
        o = np.zeros(10)
        _check(lib().vx3mi_diff_info(self._h, _ptr(o, C.c_double)))
        return dict(checkpoint_every=int(o[0]), checkpoint_bytes=o[1], tape_bytes=o[2], dt=o[3], forward_ms=o[4],
                    backward_ms=o[5], steps=int(o[6]), precision=int(o[7]), recompute_ms=o[8], reverse_ms=o[9])


def run(batch: "Batch", models: Sequence[Model], max_steps: int = 0, mode: int = MODE_AUTO, history: bool = True):
    """The simulation driver (CUDA_Simulation) over a batch: returns per robot a
    result dict, its .history text (header and frames) and its final positions."""
    models = list(models)
    n = len(models)
    texts = [[] for _ in range(n)]

    def on_text(user, robot, text, length):
        texts[robot].append(C.string_at(text, length).decode())

    cb = _TEXT_FN(on_text) if history else _TEXT_FN()
    res = (_Result * n)()
    finals = [np.zeros((m.nv, 3)) for m in models]
    fp = (C.POINTER(C.c_double) * n)(*[_ptr(f, C.c_double) for f in finals])
    arr = (C.c_void_p * n)(*[m._h.value for m in models])
    _check(lib().vx3mi_run(batch._h, arr, n, int(max_steps), int(mode), cb, None, res, fp))
    out = []
    for r in range(n):
        d = {name: getattr(res[r], name) for name, _ in _Result._fields_}
        d["initial_com"] = list(res[r].initial_com)
        d["current_com"] = list(res[r].current_com)
        out.append(d)
    return out, ["".join(t) for t in texts], finals
