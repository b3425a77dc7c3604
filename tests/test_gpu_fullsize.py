"""The measured workload at its full size (wall flow 128x64x2 hexahedra, N=4, 2,048,000 DoF) through the
C ABI against the CPU oracle: one RHS at the initial state, one RHS and one RK step from a state in
which the blender is active (20 steps in), and the bitwise repeatability of a graph-replayed run.

Tolerances: relative L2 <= 1e-12 for the RHS (the plain-FP64 bar) and alpha within rtol 1e-10. Where
the blender is active, the Hennemann sigmoid multiplies an error in the troubled-cell value by
alpha(1-alpha) s / threshold (about 2000 at N=4). The GPU sums the modal energies of the
sum-factorised transform and the oracle those of the dense change-of-basis product, so the two alphas
differ in the last few digits (about 1e-12 here). The operator is therefore checked at 1e-12 with the
oracle given the GPU's alpha, and the full difference is bounded by that plus the oracle's own
response to the alpha difference (triangle inequality)."""
import os

import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12


def _threads():
    return max(1, min(16, os.cpu_count() or 1))


def _check_rhs(g, o, label, want_blended):
    rg, ag, adg = g.calc_rhs()
    ro, ao, ado = o.calc_rhs()
    nb = int((ag > 0).sum())
    assert np.array_equal(ag > 0, ao > 0)
    assert np.allclose(ag, ao, rtol=1e-10, atol=1e-13)
    assert np.allclose(adg, ado, rtol=1e-10, atol=1e-13)
    if want_blended:
        assert nb > 0
    err = rel_l2(rg, ro)
    o.set_alpha(ag, adg)  # the oracle's operator with the GPU's blender values
    ro_ga, _, _ = o.calc_rhs()
    o.set_alpha()
    err_op = rel_l2(rg, ro_ga)
    err_alpha = rel_l2(ro_ga, ro)  # the oracle's own response to the alpha difference
    print(f"[{label}] rhs rel L2 {err:.3e} (operator {err_op:.3e}, alpha difference {err_alpha:.3e}), "
          f"max |d alpha| {np.abs(ag - ao).max():.3e}, blended cells {nb}", flush=True)
    assert err_op <= RHS_TOL
    assert err <= RHS_TOL + err_alpha * (1 + 1e-6)


def test_full_size_wall_flow():
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow()
    assert case.mesh.n_dof == 2_048_000
    g = GpuPlens(case)
    o = OraclePlens(case, threads=_threads())
    _check_rhs(g, o, "full size, initial state", want_blended=False)

    g.run(20, case.courant)
    q20 = g.get_state()
    o.set_state(q20)
    _check_rhs(g, o, "full size, after 20 steps", want_blended=True)

    dtg, dto = g.calc_time_step(case.courant), o.calc_time_step(case.courant)
    print(f"dt gpu {dtg:.17e} cpu {dto:.17e}", flush=True)
    assert abs(dtg - dto) <= 1e-12 * dto
    g.update(case.courant)
    o.update(case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"one step from the 20-step state: rel L2 {err:.3e}", flush=True)
    assert err <= RHS_TOL

    g2 = GpuPlens(case)  # same run again: the graph replay is deterministic
    g2.run(20, case.courant)
    assert np.array_equal(g2.get_state(), q20)
