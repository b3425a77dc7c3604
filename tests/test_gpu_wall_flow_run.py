"""The measured wall flow (128x64x2 cells, N=4, 2,048,000 DoF) for its full 500 RK4 steps on the GPU,
plain FP64 and FP32-bulk: the run completes without a positivity failure, every state is finite with
positive density and pressure, the step count and the solution time agree between the two, and the
blender is active. The oracle is not run for 500 full-size steps (about ten minutes of host time);
parity against it at this size is covered for one RHS, 20 steps and one step by test_gpu_fullsize,
and the FP32-bulk error where rounding is not amplified by bench.py. After 500 steps the blender's
switching has amplified rounding (the FP64 path's own spread under a 2^-24 input perturbation is
7.5e-5 after 55 steps), so the FP32-bulk difference is reported, not held to 1e-6."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

GAMMA = 1.4


def _check_state(Q):
    assert np.all(np.isfinite(Q))
    rho = Q[0]
    p = (GAMMA - 1) * (Q[4] - 0.5 * (Q[1] ** 2 + Q[2] ** 2 + Q[3] ** 2) / rho)
    assert rho.min() > 0 and p.min() > 0


def test_wall_flow_500_steps_fp64_and_fp32_bulk():
    from plens_amd import GpuPlens, wall_flow, with_precision
    case = wall_flow()
    assert case.mesh.n_dof == 2_048_000
    g = GpuPlens(case)
    f = GpuPlens(with_precision(case, "fp32_bulk"))
    tg = g.run(500, case.courant)
    tf = f.run(500, case.courant)
    Qg, Qf = g.get_state(), f.get_state()
    _check_state(Qg)
    _check_state(Qf)
    blended = g.blended_cell_stages()
    d = rel_l2(Qf, Qg)
    print(f"wall flow 500 steps: t fp64 {tg:.6e}, fp32_bulk {tf:.6e}; blended cell-stages {blended}; "
          f"FP32-bulk vs FP64 rel L2 {d:.2e}", flush=True)
    assert blended > 0
    assert abs(tf - tg) <= 1e-3 * tg  # the same dt within the FP32 error of the state it comes from
    assert d < 1e-2
