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


def _interact16():
    import sys
    sys.path.insert(0, HERE)
    import make_golden
    return make_golden


def _check_interact16(d, scene, mg):
    def on_step(k):
        if k in mg.KEPT_STEPS:
            if isinstance(scene, oracle.OracleScene):
                p4, v4 = scene.pos4, scene.vel4
            else:
                p = scene.read_particles(C9=False, density=False)
                p4, v4 = p["pos4"], p["vel4"]
            np.testing.assert_array_equal(p4, d[f"pos4_{k}"], err_msg=f"pos step {k}")
            np.testing.assert_array_equal(v4, d[f"vel4_{k}"], err_msg=f"vel step {k}")

    hit, t = mg.run_interact16(scene, d["pos3"], d["add3"], d["addv3"], on_step)
    assert hit == bool(d["ray"][0]) and np.float32(t) == d["ray"][1]


def test_oracle_reproduces_interact16():
    d, mg = fixture("interact16"), _interact16()
    o = oracle.OracleScene(grid=tuple(d["grid"]))
    _check_interact16(d, o, mg)
    np.testing.assert_array_equal(o.cell_mass, d["cell_mass"])
    np.testing.assert_array_equal(o.cell_v, d["cell_v"])
    np.testing.assert_array_equal(o.affine, d["C9"])


@pytest.mark.gpu
@pytest.mark.parametrize("form", list(FORMS))
def test_gpu_reproduces_interact16(form):
    d, mg = fixture("interact16"), _interact16()
    cap = len(d["pos3"]) + len(d["add3"])
    with mlsmpm.Scene(mlsmpm.default_config(*d["grid"], cap, **FORMS[form])) as s:
        _check_interact16(d, s, mg)
        g, p = s.read_grid(), s.read_particles()
        np.testing.assert_array_equal(g["cell_mass"], d["cell_mass"])
        np.testing.assert_array_equal(g["cell_mv"], d["cell_mv"])
        np.testing.assert_array_equal(g["cell_v"], d["cell_v"])
        np.testing.assert_array_equal(p["C9"], d["C9"])
        np.testing.assert_array_equal(p["density"], d["density"])
