"""CPU checks against the committed golden fixtures (tests/golden/, written by
tests/golden/make_fixtures.py from the oracle): the oracle still reproduces its fluxes, metric terms,
right-hand sides, time steps and RK4 steps on the stored inputs, and the product's host layer
reproduces the stored mesh, metric terms and initial conditions. Any edit that changes a result by
more than rounding fails here."""
import ctypes as C
import json
import os

import numpy as np
import pytest

from oracle import pyoracle as PO
from tests.conftest import rel_l2
from tests.golden import make_fixtures as MF

GOLD = os.path.dirname(MF.__file__)
_dp = C.POINTER(C.c_double)


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def test_fluxes_match_fixture():
    d = load("fluxes.npz")
    lib = PO.load()
    ul, ur, nrm, blend, av = d["ul"], d["ur"], d["normal"], d["blend"], d["avars"]
    f = np.zeros(5)
    for s in range(10):
        ph = MF._phys(scheme=s)
        for i in range(len(ul)):
            lib.oracle_inv_surf_flux(C.byref(ph), s, blend[i], PO.p(ul[i]), PO.p(ur[i]), PO.p(nrm[i]),
                                     f.ctypes.data_as(_dp))
            g = d["surf"][s, i]
            assert np.abs(f - g).max() <= 1e-14 * np.abs(g).max(), (s, i)
    for v in range(3):
        ph = MF._phys(vol=v)
        for i in range(len(ul)):
            lib.oracle_vol_flux(C.byref(ph), PO.p(ul[i]), PO.p(ur[i]), PO.p(nrm[i]), f.ctypes.data_as(_dp))
            g = d["vol"][v, i]
            assert np.abs(f - g).max() <= 1e-14 * np.abs(g).max(), (v, i)
    for i in range(len(ul)):
        lib.oracle_dif_flux(PO.p(ul[i]), PO.p(av[i]), PO.p(nrm[i]), f.ctypes.data_as(_dp))
        g = d["dif"][i]
        assert np.abs(f - g).max() <= 1e-14 * np.abs(g).max(), i


def test_metrics_match_fixture():
    d = load("metrics_perturbed_2x2x2_N3.npz")
    assert json.loads(str(d["mesh"])) == json.loads(json.dumps(MF.METRIC_MESH))
    m = MF.metric_mesh()
    assert np.array_equal(m.vertices, d["vertices"])  # the host mesh generator
    for arr, key in ((m.Ja, "Ja"), (m.detJ, "detJ"), (m.face_normals, "face_normals")):
        assert np.abs(arr - d[key]).max() <= 1e-15 * np.abs(d[key]).max(), key  # the host metric terms
    got = MF.metrics()  # the oracle's metric terms
    for key in ("Ja", "detJ", "face_normals"):
        assert np.abs(got[key] - d[key]).max() <= 1e-15 * np.abs(d[key]).max(), key


@pytest.mark.parametrize("stem,builder,kwargs,warm,noise", MF.RHS_CASES, ids=[c[0] for c in MF.RHS_CASES])
def test_rhs_match_fixture(stem, builder, kwargs, warm, noise):
    d = load(f"rhs_{stem}.npz")
    assert str(d["builder"]) == builder and json.loads(str(d["kwargs"])) == kwargs
    case = MF.build_case(builder, kwargs)
    if warm == 0:
        assert np.array_equal(case.Q0, d["Q"])  # the host initial condition
    o = PO.OraclePlens(case)
    o.set_state(d["Q"])
    rhs, a, ad = o.calc_rhs()
    assert rel_l2(rhs, d["rhs"]) <= 1e-14
    assert np.allclose(a, d["alpha"], rtol=1e-13, atol=1e-15)
    assert np.allclose(ad, d["alpha_d"], rtol=1e-13, atol=1e-15)
    # the stored state is well conditioned: rounding-level input changes move the RHS far less than
    # the 1e-12 bar the GPU is held to
    cond = MF.rhs_conditioning(o, d["Q"], d["rhs"])
    assert cond <= 5e-13 and abs(cond - float(d["conditioning"])) <= 0.1 * cond
    dt = o.calc_time_step(float(d["courant"]))
    assert abs(dt - float(d["dt"])) <= 1e-14 * dt
    o.update(float(d["courant"]))
    assert rel_l2(o.get_state(), d["Q1"]) <= 1e-14
    o.close()


def test_fixture_cases_exercise_the_blender_and_walls():
    # the shock case has blended cells with alpha strictly between the bounds, the wall flow a
    # non-trivial viscous right-hand side
    s = load("rhs_shock_tube_6x2x2_N2.npz")
    assert ((s["alpha"] > 0) & (s["alpha"] < 0.5)).any()
    w = load("rhs_wall_flow_4x2x2_N4.npz")
    assert np.abs(w["rhs"]).max() > 1e3
    assert not np.array_equal(w["Q"], w["Q1"])
