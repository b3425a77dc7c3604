"""FP32-bulk variant (fluxes, viscous auxiliary variables and finished face terms in FP32; states, RK
registers, residual accumulation, metrics, blender, time step and the near-equal log-mean branch in
FP64) against the FP64 oracle on identical meshes and states.

Bars: the blender values and the time step do not depend on the flux arithmetic and must match the
oracle to FP64 round-off. The RHS error is bounded relative to the RHS's own sensitivity to FP32
round-off: the oracle's RHS change when the state is perturbed by the FP32 unit round-off 2^-24
(relative, uniform) measures how much the cancellation of the DG sums amplifies FP32-level errors of
the fluxes; the variant must stay within 10 times that and within relative L2 1e-4 absolute. The
solution after the listed steps must stay within relative L2 1e-6."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-4
SENS_FACTOR = 10.0
RUN_TOL = 1e-6


def _pair(case):
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, with_precision
    return GpuPlens(with_precision(case, "fp32_bulk")), OraclePlens(case)


def _fp32_sensitivity(case, ro):
    """relative L2 change of the oracle RHS under a 2^-24 relative perturbation of the state"""
    from oracle.pyoracle import OraclePlens
    rng = np.random.default_rng(5)
    o = OraclePlens(case)
    o.set_state(case.Q0 * (1.0 + 2.0 ** -24 * rng.uniform(-1.0, 1.0, case.Q0.shape)))
    r, _, _ = o.calc_rhs()
    return rel_l2(r, ro)


def _rhs(case, label):
    g, o = _pair(case)
    rg, ag, adg = g.calc_rhs()
    ro, ao, ado = o.calc_rhs()
    err = rel_l2(rg, ro)
    sens = _fp32_sensitivity(case, ro)
    print(f"[{label}] fp32-bulk rhs rel L2 = {err:.3e}  fp32 input sensitivity {sens:.3e}  "
          f"ratio {err / sens:.1f}", flush=True)
    assert np.allclose(ag, ao, rtol=1e-10, atol=1e-13)
    assert np.allclose(adg, ado, rtol=1e-10, atol=1e-13)
    assert err <= RHS_TOL, f"{label}: rhs rel L2 {err:.3e}"
    assert err <= SENS_FACTOR * sens, f"{label}: rhs rel L2 {err:.3e} vs sensitivity {sens:.3e}"
    return err


@pytest.mark.parametrize("flux", ["HLLC", "AUSM+-up", "Chandrashekhar", "Rusanov"])
def test_fp32_rhs_periodic_cube(flux):
    from plens_amd import periodic_cube
    _rhs(periodic_cube(n=4, degree=3, flux=flux), f"cube-{flux}")


@pytest.mark.parametrize("degree", [1, 2, 4, 5])
def test_fp32_rhs_degrees(degree):
    from plens_amd import periodic_cube
    _rhs(periodic_cube(n=3, degree=degree, flux="Chandrashekhar"), f"cube-N{degree}")


@pytest.mark.parametrize("flux", ["HLLC", "Chandrashekhar"])
def test_fp32_rhs_shock_tube_blended(flux):
    from plens_amd import shock_tube
    _rhs(shock_tube(nx=16, ny=2, nz=2, flux=flux), f"shock-{flux}")


@pytest.mark.parametrize("surface,volume", [("Rusanov-HLLC", "Chandrashekhar"),
                                            ("Rusanov-AUSM+-up", "Chandrashekhar"), ("Modified SW", "KG"),
                                            ("KG", "KG"), ("Rusanov-KG", "KG"), ("IR", "IR")])
def test_fp32_rhs_remaining_fluxes(surface, volume):
    from plens_amd import periodic_cube, shock_tube, with_fluxes
    _rhs(with_fluxes(periodic_cube(n=4, degree=3), surface=surface, volume=volume), f"cube-{surface}-{volume}")
    _rhs(with_fluxes(shock_tube(nx=16, ny=2, nz=2), surface=surface, volume=volume), f"shock-{surface}-{volume}")


def test_fp32_dt_matches_fp64():
    from plens_amd import wall_flow
    case = wall_flow(nx=6, ny=4, nz=2)
    g, o = _pair(case)
    dg, do = g.calc_time_step(case.courant), o.calc_time_step(case.courant)
    assert abs(dg - do) <= 1e-13 * do


def test_fp32_steps_shock_tube():
    from plens_amd import shock_tube
    case = shock_tube(nx=16, ny=2, nz=2)
    g, o = _pair(case)
    tg = g.run(20, case.courant)
    to = o.run(20, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"fp32-bulk shock 20 steps rel L2 {err:.3e}  t {tg:.10g} / {to:.10g}", flush=True)
    assert err <= RUN_TOL


def test_fp32_steps_wall_flow():
    from plens_amd import wall_flow
    case = wall_flow(nx=6, ny=4, nz=2)
    g, o = _pair(case)
    g.run(10, case.courant)
    o.run(10, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"fp32-bulk wall 10 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


def test_fp32_run_100_steps_periodic_cube():
    from plens_amd import periodic_cube
    case = periodic_cube(n=4, degree=3)
    g, o = _pair(case)
    g.run(100, case.courant)
    o.run(100, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"fp32-bulk cube 100 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


def test_fp32_full_size_wall_flow():
    """the measured wall flow at full size (128x64x2 cells, N=4, 2,048,000 DoF), whose handle runs the
    streaming face and volume kernels: RHS against the oracle, and 5 steps against the FP64 path
    (before the blender starts switching cells, so rounding is not amplified yet)"""
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow()
    assert case.mesh.n_dof == 2_048_000
    _rhs(case, "wall flow full size")
    g32, _ = _pair(case)
    g64 = GpuPlens(case)
    g32.run(5, case.courant)
    g64.run(5, case.courant)
    err = rel_l2(g32.get_state(), g64.get_state())
    print(f"wall flow full size, 5 steps: fp32-bulk vs fp64 rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL
