"""Host-side checks (no GPU): the product's model builder against the oracle's
independent builder, the C ABI surface, and error handling of the host calls."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
import vx3mi
from vx3mi import configs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PROBLEMS = {
    "config0": configs.config0,
    "config1_block": lambda: configs.config1(seed=1, dims=(12, 10, 6)),
    "config2_robot": lambda: configs.config2(seed=3, n=1)[0],
    "config3": configs.config3,
    # MatModel 3 palette (stress-strain data), the data-model golden robot
    "data_model": lambda: dict(configs.config0(), materials=[
        dict(rho=1000.0, cte=0.001, ss_data=([0.0, 0.002, 0.006, 0.015, 0.05], [0.0, 2000.0, 4400.0, 7000.0, 9000.0]))]),
}


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_topology_matches_oracle(name):
    prob = PROBLEMS[name]()
    m, o = vx3mi.Model.from_problem(prob), O.OracleModel(prob)
    assert (m.nv, m.nl, m.ns) == (o.nv, o.nl, o.ns)
    t, u = m.topology(), o.topology()
    for k in ("lattice", "link_neg", "link_pos", "link_axis"):
        assert np.array_equal(t[k], u[k]), k


@pytest.mark.parametrize("name", list(PROBLEMS))
def test_material_constants_bit_equal(name):
    prob = PROBLEMS[name]()
    m, o = vx3mi.Model.from_problem(prob), O.OracleModel(prob)
    f = np.float32
    zI, zG, zC = f(prob["bond_damp_z"]), f(prob["slow_damp_z"]), f(prob["col_damp_z"])
    gm = f(-f(prob["grav_acc"]) / 9.80665)
    for v in range(0, m.nv, max(1, m.nv // 37)):
        E, nu, rho, cte, muS, muK, mass, minv, smass, iinv, c2m, c2i = o.voxel_material(v)
        expect = np.array([mass, minv, iinv, f(f(2) * smass) * zI, zG * c2m, zG * c2i, zC * c2m,
                           f(f(-mass) * f(9.80665)) * gm, f(float(f(2) * E) * prob["lattice_dim"]), muS, muK, cte],
                          dtype=np.float32)
        assert np.array_equal(m.voxel_constants(v), expect), v
    for l in range(0, m.nl, max(1, m.nl // 41)):
        assert np.array_equal(m.link_constants(l), o.link_material(l)), l
    assert m.recommended_dt() == o.recommended_dt()


def test_config1_counts():
    """The 64 x 64 x 16 large-body workload (seed 0): 45,800 voxels, 93,183 links,
    41,479 surface voxels."""
    m = vx3mi.Model.from_problem(configs.config1())
    assert (m.nv, m.nl, m.ns) == (45800, 93183, 41479)


def test_abi_exports_every_header_symbol():
    hdr = open(os.path.join(ROOT, "include", "vx3mi.h")).read()
    names = set(re.findall(r"\b(vx3mi_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 25
    L = vx3mi.lib()
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_host_error_paths():
    L = vx3mi.lib()
    h = C.c_void_p()
    assert L.vx3mi_model_from_lattice(None, C.byref(h)) == -1
    assert b"invalid" in L.vx3mi_last_error()
    big = dict(configs.config0(), dims=(1024, 1, 1), cells=np.ones(1024, dtype=np.int8), phase=None)
    with pytest.raises(vx3mi.VX3MIError, match="-5"):
        vx3mi.Model.from_problem(big)
    # an empty robot is rejected before any device is touched (CUDA_Simulation aborts it)
    empty = dict(configs.config0(), cells=np.zeros(64, dtype=np.int8), phase=None)
    em = vx3mi.Model.from_problem(empty)
    assert em.nv == 0
    with pytest.raises(vx3mi.VX3MIError, match="no voxels"):
        vx3mi.Batch([em])
    with pytest.raises(vx3mi.VX3MIError):
        vx3mi.Batch([vx3mi.Model.from_problem(configs.config0())], precision=16)


def test_palette_index_is_signed_char():
    """Palette values above 127 read back negative in the reference
    (src/VXA/VX_Object.h:160) and are treated as empty cells."""
    p = configs.config0()
    cells = np.ones(64, dtype=np.int16)
    cells[:8] = 200
    p["cells"] = cells.astype(np.uint8).view(np.int8)
    p["phase"] = None
    assert vx3mi.Model.from_problem(p).nv == 56 == O.OracleModel(p).nv
