"""The time step on the GPU: the caller-supplied step (plens_gpu_rk4_step_dt, the path of a caller that
reduces the step over partitions as dgsem/plens.cc:3872-3875 does), the per-cell steps of
plens_gpu_calc_dt against the oracle's, and the closed form of calc_time_step (dgsem/plens.cc:3818-3862)
on a sheared periodic lattice."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def test_given_step_equals_computed_step_bitwise():
    """calc_dt then rk4_step_dt(dt) gives the state of rk4_step(courant), bit for bit, over 3 steps of
    a blended viscous shock tube"""
    from plens_amd import GpuPlens, shock_tube
    case = shock_tube(nx=16, ny=2, nz=2)
    a, b = GpuPlens(case), GpuPlens(case)
    t = 0.0
    for _ in range(3):
        dt = a.calc_time_step(case.courant)
        a.update_with_step(dt, t)
        tb = b.update(case.courant, t)  # (t + dt) - t: dt up to the rounding of the sum
        assert abs(tb - dt) <= 4e-16 * (t + dt)
        assert np.array_equal(a.get_state(), b.get_state())
        t += dt


def test_given_step_scales_the_update_and_rejects_bad_steps():
    """half the step gives a different state than the full step; zero, negative and non-finite steps
    and local time stepping are rejected"""
    from plens_amd import GpuPlens, periodic_cube
    case = periodic_cube(n=3, degree=3)
    a, b = GpuPlens(case), GpuPlens(case)
    dt = a.calc_time_step(case.courant)
    a.update_with_step(dt)
    b.update_with_step(0.5 * dt)
    assert rel_l2(a.get_state(), b.get_state()) > 1e-6
    for bad in (0.0, -dt, float("inf"), float("nan")):
        with pytest.raises(RuntimeError):
            a.update_with_step(bad)
    a.set_local_stepping(True, 0.0)
    with pytest.raises(RuntimeError):
        a.update_with_step(dt)


def test_cell_steps_match_oracle():
    """dt_cell = each cell's own stable step: the oracle's local steps with the limiter off, and their
    minimum is dt_global"""
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow(nx=6, ny=4, nz=2)
    g, o = GpuPlens(case), OraclePlens(case)
    dg, cells = g.calc_time_step(case.courant, cell_steps=True)
    o.set_local_stepping(True, -1.0, limit=False)
    o.update(case.courant)
    oc = o.loc_dt()
    print(f"cell steps: max rel diff {np.abs(cells / oc - 1).max():.3e}", flush=True)
    assert np.allclose(cells, oc, rtol=1e-13, atol=0)
    assert dg == cells.min()


def test_time_step_closed_form_on_sheared_lattice():
    """the GPU step on a periodic lattice of sheared parallelepipeds (trilinear map path) against the
    closed form (numpy inverse of the cell matrix, brute-force minimum vertex distance), with a
    Courant number above 1 (capped to 0.9)"""
    from plens_amd import _lib as L
    from plens_amd import GpuPlens, cases
    from tests.test_oracle_step_bc_blender_pins import DT_CASES, closed_form_dt, lattice_mesh
    for name in ("sheared", "courant_cap", "negative_max"):
        A, N, rho, vel, p, co, mu0 = DT_CASES[name]
        lm = lattice_mesh(A, N)
        m = cases.Mesh.from_cells(N, lm.vertices, lm.face_nbr, lm.face_bid)
        E = p / 0.4 + 0.5 * rho * float(np.dot(vel, vel))
        Q = np.tile(np.array([rho, rho * vel[0], rho * vel[1], rho * vel[2], E])[:, None], (1, m.n_dof))
        g = GpuPlens(mesh=m, physics=cases.make_physics(mu0=mu0), bcs=[cases.make_bc(L.BC_PERIODIC)])
        g.set_state(Q)
        got, want = g.calc_time_step(co), closed_form_dt(A, N, rho, vel, p, co, mu0)
        print(f"[{name}] dt gpu {got:.17g} closed form {want:.17g}", flush=True)
        assert abs(got - want) <= 1e-13 * want
