"""Golden fixtures (tests/golden/, written by make_golden.py from the CPU
restatement): the oracle must keep reproducing them (CPU), and the HIP
library must reproduce them bit for bit in its direct and sorted forms
(GPU)."""
import os

import numpy as np
import pytest

import mlsmpm
from oracle import oracle

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    return np.load(os.path.join(HERE, name + ".npz"))


def test_oracle_reproduces_tiny8():
    d = fixture("tiny8")
    o = oracle.OracleScene(grid=tuple(d["grid"]))
    o.reset(d["pos3"], d["vel3"])
    for k in range(3):
        o.step(1)
        np.testing.assert_array_equal(o.cell_mass, d[f"cell_mass_{k}"])
        np.testing.assert_array_equal(o.cell_mv, d[f"cell_mv_{k}"])
        np.testing.assert_array_equal(o.cell_v, d[f"cell_v_{k}"])
        np.testing.assert_array_equal(o.pos4, d[f"pos4_{k}"])
        np.testing.assert_array_equal(o.affine, d[f"C9_{k}"])


def test_oracle_reproduces_dam16_with_diffuse():
    d = fixture("dam16_d")
    D = int(d["diffuse_max"][0])
    o = oracle.OracleScene(grid=tuple(d["grid"]), diffuse_max=D, seed=1)
    o.reset(d["pos3"])
    for k in range(4):
        o.step(1)
        np.testing.assert_array_equal(o.pos4, d[f"pos4_{k}"])
        np.testing.assert_array_equal(o.vel4, d[f"vel4_{k}"])
    np.testing.assert_array_equal(o.affine, d["C9_3"])
    assert int(o.top[0]) == int(d["top"][0])
    np.testing.assert_array_equal(o.free_list, d["free_list"])
    for f in ("pos", "vel", "ptype", "alive", "lifetime"):
        np.testing.assert_array_equal(o.pool[f], d[f"pool_{f}"])


FORMS = {
    "direct": dict(flags=mlsmpm.MPM_FLAG_KEEP_GRID | mlsmpm.MPM_FLAG_NO_LDS, sort_period=0),
    "sorted1": dict(flags=mlsmpm.MPM_FLAG_KEEP_GRID, sort_period=1),
}


@pytest.mark.gpu
@pytest.mark.parametrize("form", list(FORMS))
def test_gpu_reproduces_tiny8(form):
    d = fixture("tiny8")
    pos = d["pos3"]
    with mlsmpm.Scene(mlsmpm.default_config(*d["grid"], len(pos), **FORMS[form])) as s:
        s.reset(pos, d["vel3"])
        for k in range(3):
            s.step(1)
            g = s.read_grid()
            p = s.read_particles()
            np.testing.assert_array_equal(g["cell_mass"], d[f"cell_mass_{k}"])
            np.testing.assert_array_equal(g["cell_mv"], d[f"cell_mv_{k}"])
            np.testing.assert_array_equal(g["cell_mass_f"], d[f"cell_mass_f_{k}"])
            np.testing.assert_array_equal(g["cell_v"], d[f"cell_v_{k}"])
            np.testing.assert_array_equal(p["pos4"], d[f"pos4_{k}"])
            np.testing.assert_array_equal(p["vel4"], d[f"vel4_{k}"])
            np.testing.assert_array_equal(p["C9"], d[f"C9_{k}"])
            np.testing.assert_array_equal(p["density"], d[f"density_{k}"])


@pytest.mark.gpu
@pytest.mark.parametrize("form", list(FORMS))
def test_gpu_reproduces_dam16_with_diffuse(form):
    d = fixture("dam16_d")
    pos = d["pos3"]
    D = int(d["diffuse_max"][0])
    cfg = mlsmpm.default_config(*d["grid"], len(pos), diffuse_capacity=D, seed=1, **FORMS[form])
    with mlsmpm.Scene(cfg) as s:
        s.reset(pos)
        for k in range(4):
            s.step(1)
            p = s.read_particles()
            np.testing.assert_array_equal(p["pos4"], d[f"pos4_{k}"])
            np.testing.assert_array_equal(p["vel4"], d[f"vel4_{k}"])
        np.testing.assert_array_equal(s.read_particles()["C9"], d["C9_3"])
        rec, fl, top = s.read_diffuse()
        assert top == int(d["top"][0])
        np.testing.assert_array_equal(fl, d["free_list"])
        for f in ("pos", "vel", "ptype", "alive", "lifetime"):
            np.testing.assert_array_equal(rec[f], d[f"pool_{f}"])
