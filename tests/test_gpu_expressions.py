"""Expression-valued Courant number, wall blender limit and viscous blending region on the GPU
(plens_gpu_set_expressions, evaluated on the device once per step at the step's start time) against
the CPU oracle, which evaluates the same strings with its own evaluator."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

EXPR = dict(courant="0.3 + 20*t", wall_blender_limit="0.05 + 100*t", viscous_blending_region="y > 0.05")


def test_expressions_parity_wall_flow():
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow(nx=6, ny=4, nz=2)
    g, o, ref = GpuPlens(case), OraclePlens(case), GpuPlens(case)
    g.set_expressions(**EXPR)
    o.set_expressions(**EXPR)
    tg = g.run(10, 0.77, 0.0)  # the numeric Courant number is ignored
    to = o.run(10, 0.77, 0.0)
    ref.run(10, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    moved = rel_l2(g.get_state(), ref.get_state())
    _, ag, adg = g.calc_rhs()
    o.set_time(tg)
    _, ao, ado = o.calc_rhs()
    print(f"expressions, 10 steps: rel L2 vs oracle {err:.3e}, change from constant parameters {moved:.3e}, "
          f"t gpu {tg:.15g} cpu {to:.15g}, cells with alpha_d = 0: {int((adg == 0).sum())}", flush=True)
    assert abs(tg - to) <= 1e-12 * to
    assert err <= 1e-10
    assert moved > 1e-6
    assert np.allclose(ag, ao, rtol=1e-10, atol=1e-13)
    assert np.array_equal(adg == 0, ado == 0)


def test_expression_errors_rejected():
    from plens_amd import GpuPlens, wall_flow
    g = GpuPlens(wall_flow(nx=4, ny=2, nz=2))
    for bad in ("0.3 +", "sin(", "foo(t)"):
        with pytest.raises(RuntimeError):
            g.set_expressions(courant=bad)
    g.set_expressions()  # clears
