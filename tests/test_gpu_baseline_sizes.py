"""Parity at the sizes and step counts of the measured configurations, through the C ABI against the
CPU oracle:
  - the periodic cube at full size (8^3 cells, N=3, 32,768 DoF) for 100 RK4 steps, with alpha = 0 in
    every cell at every stage on both sides (the pure DG path);
  - the shock tube at full size (128x8x8 cells, N=3, 524,288 DoF): one RHS at the initial state and
    one after 20 steps with the blender active, and one RK step from there;
  - the full shock tube for its 200 steps, and 100 steps of the reduced shock tube (16x2x2) and
    wall flow (6x4x2), with the blender switching.

Bars: relative L2 <= 1e-12 per RHS evaluation and <= 1e-10 after 100 steps. Where the blender is
active, the RHS difference is split as in test_gpu_fullsize: the operator with the oracle given the
GPU's alpha (<= 1e-12) and the oracle's own response to the alpha difference."""
import os

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
RUN_TOL = 1e-10


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _pair(case):
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens
    return GpuPlens(case), OraclePlens(case, threads=_threads())


def test_periodic_cube_full_size_100_steps_pure_dg():
    from plens_amd import periodic_cube
    case = periodic_cube()  # 8^3 cells, N=3
    assert case.mesh.n_dof == 32_768
    g, o = _pair(case)
    tg = g.run(100, case.courant)
    to = o.run(100, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    bg, bo = g.blended_cell_stages(), o.blended_cell_stages()
    print(f"cube 8^3 N=3, 100 steps: rel L2 {err:.3e}, t gpu {tg:.15g} cpu {to:.15g}, "
          f"blended cell-stages gpu {bg} cpu {bo}", flush=True)
    assert bg == 0 and bo == 0  # alpha = 0 in every cell at all 500 stages
    assert abs(tg - to) <= 1e-12 * to
    assert err <= RUN_TOL


def _rhs_split(g, o, label):
    rg, ag, adg = g.calc_rhs()
    ro, ao, ado = o.calc_rhs()
    assert np.array_equal(ag > 0, ao > 0)
    assert np.allclose(ag, ao, rtol=1e-10, atol=1e-13)
    err = rel_l2(rg, ro)
    o.set_alpha(ag, adg)
    ro_ga, _, _ = o.calc_rhs()
    o.set_alpha()
    err_op = rel_l2(rg, ro_ga)
    err_alpha = rel_l2(ro_ga, ro)
    nb = int((ag > 0).sum())
    print(f"[{label}] rhs rel L2 {err:.3e} (operator {err_op:.3e}, alpha difference {err_alpha:.3e}), "
          f"blended cells {nb}", flush=True)
    assert err_op <= RHS_TOL
    assert err <= RHS_TOL + err_alpha * (1 + 1e-6)
    return nb


def test_shock_tube_full_size_rhs_and_step():
    from plens_amd import shock_tube
    case = shock_tube()  # 128x8x8 cells, N=3
    assert case.mesh.n_dof == 524_288
    g, o = _pair(case)
    assert _rhs_split(g, o, "shock tube 128x8x8, initial state") > 0
    g.run(20, case.courant)
    o.set_state(g.get_state())
    assert _rhs_split(g, o, "shock tube 128x8x8, after 20 steps") > 0
    dtg, dto = g.calc_time_step(case.courant), o.calc_time_step(case.courant)
    assert abs(dtg - dto) <= 1e-12 * dto
    g.update(case.courant)
    o.update(case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"shock tube 128x8x8: one step from the 20-step state: rel L2 {err:.3e}", flush=True)
    assert err <= RHS_TOL


@pytest.mark.parametrize("which", ["shock_tube_16x2x2", "wall_flow_6x4x2"])
def test_100_steps_with_blending(which):
    from plens_amd import shock_tube, wall_flow
    case = shock_tube(nx=16, ny=2, nz=2) if which.startswith("shock") else wall_flow(nx=6, ny=4, nz=2)
    g, o = _pair(case)
    tg = g.run(100, case.courant)
    to = o.run(100, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    bg, bo = g.blended_cell_stages(), o.blended_cell_stages()
    print(f"[{which}] 100 steps: rel L2 {err:.3e}, t rel diff {abs(tg - to) / to:.2e}, "
          f"blended cell-stages gpu {bg} cpu {bo}", flush=True)
    assert bg > 0 and bg == bo
    assert err <= RUN_TOL


def test_shock_tube_full_size_200_steps():
    """the shock tube at full size for its 200 steps (about a minute of oracle time on 16 threads)"""
    from plens_amd import shock_tube
    case = shock_tube()
    g, o = _pair(case)
    tg = g.run(200, case.courant)
    to = o.run(200, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    bg, bo = g.blended_cell_stages(), o.blended_cell_stages()
    print(f"shock tube 128x8x8, 200 steps: rel L2 {err:.3e}, t rel diff {abs(tg - to) / to:.2e}, "
          f"blended cell-stages gpu {bg} cpu {bo}", flush=True)
    assert bg > 0 and bg == bo
    assert err <= RUN_TOL
