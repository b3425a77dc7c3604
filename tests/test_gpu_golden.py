"""The GPU path through the C ABI against the committed golden fixtures (tests/golden/, written by the
oracle): one right-hand side with blender values, the time step and one RK4 step on each stored
state. Tolerances: relative L2 <= 1e-12 for the right-hand side and the step (the plain-FP64 bar),
alpha within rtol 1e-10, dt within 1e-12."""
import json
import os

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
STEMS = ["periodic_cube_2x2x2_N3", "shock_tube_6x2x2_N2", "wall_flow_4x2x2_N4"]


@pytest.mark.parametrize("stem", STEMS)
def test_gpu_matches_golden_rhs(stem):
    from plens_amd import GpuPlens, cases
    d = np.load(os.path.join(GOLD, f"rhs_{stem}.npz"), allow_pickle=False)
    print(f"[golden {stem}]", flush=True)
    case = getattr(cases, str(d["builder"]))(**json.loads(str(d["kwargs"])))
    g = GpuPlens(case)
    g.set_state(d["Q"])
    rhs, a, ad = g.calc_rhs()
    err = rel_l2(rhs, d["rhs"])
    print(f"[golden {stem}] rhs rel L2 {err:.3e}, blended cells {(a > 0).sum()}", flush=True)
    assert np.array_equal(a > 0, d["alpha"] > 0)
    assert np.allclose(a, d["alpha"], rtol=1e-10, atol=1e-13)
    assert np.allclose(ad, d["alpha_d"], rtol=1e-10, atol=1e-13)
    assert err <= 1e-12
    courant = float(d["courant"])
    dt = g.calc_time_step(courant)
    assert abs(dt - float(d["dt"])) <= 1e-12 * dt
    g.set_state(d["Q"])
    g.update(courant)
    err1 = rel_l2(g.get_state(), d["Q1"])
    print(f"[golden {stem}] one step rel L2 {err1:.3e}", flush=True)
    assert err1 <= 1e-12
