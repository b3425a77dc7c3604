"""GPU parity of the HIP path (through the C ABI) against the CPU oracle on identical meshes and
states. Tolerances from the plain-FP64 bar: relative L2 <= 1e-12 per RHS evaluation and <= 1e-10
after 100 RK4 steps."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

RHS_TOL = 1e-12
RUN_TOL = 1e-10


def _pair(case):
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens
    return GpuPlens(case), OraclePlens(case)


def _rhs_parity(case, label):
    g, o = _pair(case)
    rg, ag, adg = g.calc_rhs()
    ro, ao, ado = o.calc_rhs()
    err = rel_l2(rg, ro)
    print(f"[{label}] rhs rel L2 = {err:.3e}  alpha max gpu/cpu = {ag.max():.4f}/{ao.max():.4f}", flush=True)
    assert np.array_equal(ag > 0, ao > 0)
    assert np.allclose(ag, ao, rtol=1e-10, atol=1e-13)
    assert np.allclose(adg, ado, rtol=1e-10, atol=1e-13)
    assert err <= RHS_TOL, f"{label}: rhs rel L2 {err:.3e}"
    return g, o


@pytest.mark.parametrize("flux", ["HLLC", "AUSM+-up", "Chandrashekhar", "Rusanov"])
def test_rhs_periodic_cube(flux):
    from plens_amd import periodic_cube
    _rhs_parity(periodic_cube(n=4, degree=3, flux=flux), f"cube-{flux}")


@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5])
def test_rhs_periodic_cube_degrees(degree):
    from plens_amd import periodic_cube
    _rhs_parity(periodic_cube(n=3, degree=degree, flux="Chandrashekhar"), f"cube-N{degree}")


@pytest.mark.parametrize("flux", ["HLLC", "AUSM+-up", "Chandrashekhar"])
def test_rhs_shock_tube_blended(flux):
    from plens_amd import shock_tube
    case = shock_tube(nx=16, ny=2, nz=2, flux=flux)
    g, o = _rhs_parity(case, f"shock-{flux}")
    _, a, _ = o.calc_rhs()
    assert (a > 0).any(), "the blender must be active near the shock"


@pytest.mark.parametrize("flux", ["HLLC", "Chandrashekhar"])
def test_rhs_wall_flow_curvilinear(flux):
    from plens_amd import wall_flow
    _rhs_parity(wall_flow(nx=6, ny=4, nz=2, flux=flux), f"wall-{flux}")


REMAINING_SURFACE = ["Rusanov-HLLC", "Rusanov-AUSM+-up", "Modified SW", "KG", "Rusanov-KG", "IR"]


@pytest.mark.parametrize("flux", REMAINING_SURFACE)
def test_rhs_remaining_surface_fluxes(flux):
    """Surface schemes 4-9 in the face kernel and the subcell FV of blended cells (shock tube), and
    through the curvilinear face rotations (wall flow)."""
    from plens_amd import shock_tube, wall_flow
    _, o = _rhs_parity(shock_tube(nx=16, ny=2, nz=2, flux=flux), f"shock-{flux}")
    assert (o.calc_rhs()[1] > 0).any()
    _rhs_parity(wall_flow(nx=6, ny=4, nz=2, flux=flux), f"wall-{flux}")


@pytest.mark.parametrize("volume", ["KG", "IR"])
def test_rhs_volume_fluxes(volume):
    """Kennedy-Gruber and Ismail-Roe two-point fluxes of the split-form volume term: curvilinear
    metrics, blended cells, every degree's pair tables."""
    from plens_amd import periodic_cube, shock_tube, wall_flow, with_fluxes
    _rhs_parity(with_fluxes(wall_flow(nx=6, ny=4, nz=2), volume=volume), f"wall-vol-{volume}")
    _rhs_parity(with_fluxes(shock_tube(nx=16, ny=2, nz=2), surface=volume, volume=volume),
                f"shock-{volume}-{volume}")
    for degree in (1, 2, 3, 5):
        _rhs_parity(with_fluxes(periodic_cube(n=3, degree=degree), volume=volume), f"cube-N{degree}-vol-{volume}")


# (IR surface fluxes lose positivity on the Mach-6 isothermal wall within 10 steps in the oracle as
# well: that pairing is stepped on the shock tube only)
@pytest.mark.parametrize("name,surface,volume", [("shock", "IR", "IR"), ("wall", "Rusanov-HLLC", "IR"),
                                                 ("shock", "Modified SW", "KG"), ("wall", "Modified SW", "KG"),
                                                 ("wall", "Rusanov-KG", "KG"), ("shock", "Rusanov-AUSM+-up", "KG")])
def test_steps_remaining_fluxes(name, surface, volume):
    from plens_amd import shock_tube, wall_flow, with_fluxes
    base = shock_tube(nx=16, ny=2, nz=2) if name == "shock" else wall_flow(nx=6, ny=4, nz=2)
    case = with_fluxes(base, surface=surface, volume=volume)
    g, o = _pair(case)
    g.run(10, case.courant)
    o.run(10, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"{name} {surface}/{volume} 10 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


def _with_jump(case, x0):
    """density doubled beyond x = x0 at the same velocity and pressure: blended cells at the jump"""
    Q = case.Q0.copy()
    x = case.mesh.node_coords[:, 0, :].reshape(-1)
    f = np.where(x > x0, 2.0, 1.0)
    v = Q[1:4] / Q[0]
    p = 0.4 * (Q[4] - 0.5 * Q[0] * (v * v).sum(0))
    Q[0] *= f
    Q[1:4] = Q[0] * v
    Q[4] = p / 0.4 + 0.5 * Q[0] * (v * v).sum(0)
    case.Q0 = Q
    return case


@pytest.mark.parametrize("geometry,N,q", [("flare", 3, 2), ("flare", 4, 3), ("flare-perturbed", 4, 2),
                                          ("nose", 3, 2), ("nose", 4, 3)])
def test_curved_meshes(geometry, N, q):
    """MappingQ(q) meshes on the cylinder and nose-cylinder manifold descriptions, metric terms
    handed over stored: RHS (smooth, and with a density jump that blends cells), 10 steps."""
    from plens_amd import cylinder_flare, nose_cone
    make = (lambda: nose_cone(degree=N, mapping_degree=q)) if geometry == "nose" else \
        (lambda: cylinder_flare(degree=N, mapping_degree=q, perturb=0.2 if "perturbed" in geometry else 0.0))
    _rhs_parity(make(), f"{geometry}-N{N}-q{q}")
    x0 = -0.2 if geometry == "nose" else 1.1
    _, o = _rhs_parity(_with_jump(make(), x0), f"{geometry}-N{N}-q{q}-jump")
    assert (o.calc_rhs()[1] > 0).any()
    case = make()
    g, o = _pair(case)
    g.run(10, case.courant)
    o.run(10, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"{geometry} N{N} q{q} 10 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


@pytest.mark.parametrize("exprs", ["default", "front"])
def test_varying_inflow(exprs):
    """Varying inflow: prescribed p, T, velocity from expressions of the face DoF location and the
    step's start time; RHS at two BC times and 10 steps from t = 0.3. "front": a moving
    discontinuity written as in the double-Mach-reflection input (ternary in y and t)."""
    from plens_amd import varying_inflow_tube
    kw = {} if exprs == "default" else dict(
        p_expr="( y - 0.05 - 0.5*t > 0 ? 1 : 1.3 )", T_expr="( y - 0.05 - 0.5*t > 0 ? 1 : 1.1 )",
        vel_expr="( y - 0.05 - 0.5*t > 0 ? 2 : 2.2 ); ( y - 0.05 - 0.5*t > 0 ? 0 : -0.1 ); 0")
    case = varying_inflow_tube(**kw)
    g, o = _rhs_parity(case, f"vinflow-{exprs}-t0")
    for h in (g, o):
        h.set_time(0.2)
    rg, _, _ = g.calc_rhs()
    ro, _, _ = o.calc_rhs()
    err = rel_l2(rg, ro)
    print(f"vinflow-{exprs} t=0.2 rhs rel L2 {err:.3e}", flush=True)
    assert err <= RHS_TOL
    g, o = _pair(case)
    g.run(10, case.courant, t0=0.3)
    o.run(10, case.courant, t0=0.3)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"vinflow-{exprs} 10 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


def test_constant_varying_inflow_is_bitwise_the_uniform_inflow():
    from plens_amd import GpuPlens, varying_inflow_tube
    from plens_amd import _lib as L
    from plens_amd.cases import make_bc
    case = varying_inflow_tube()
    uni = make_bc(L.BC_UNIFORM_INFLOW, p=1.2, T=0.9, vel=(2.0, 0.1, -0.05))
    var = make_bc(L.BC_VARYING_INFLOW, p_expr="1.2", T_expr="0.9", vel_expr="2; 0.1; -0.05")
    a = GpuPlens(mesh=case.mesh, physics=case.physics, bcs=[uni] + case.bcs[1:])
    b = GpuPlens(mesh=case.mesh, physics=case.physics, bcs=[var] + case.bcs[1:])
    for h in (a, b):
        h.set_state(case.Q0)
        h.run(5, case.courant)
    assert np.array_equal(a.get_state(), b.get_state())


def test_varying_inflow_rejections():
    from plens_amd import GpuBatch, GpuPlens, varying_inflow_tube
    from plens_amd import _lib as L
    from plens_amd.cases import make_bc
    case = varying_inflow_tube()
    with pytest.raises(L.PlensError):
        GpuPlens(mesh=case.mesh, physics=case.physics,
                 bcs=[make_bc(L.BC_VARYING_INFLOW, p_expr="1 +", T_expr="1", vel_expr="1; 0; 0")] + case.bcs[1:])
    with pytest.raises(L.PlensError):
        GpuPlens(mesh=case.mesh, physics=case.physics,
                 bcs=[make_bc(L.BC_VARYING_INFLOW, p_expr="1", T_expr="1", vel_expr="1; 0")] + case.bcs[1:])
    with pytest.raises(L.PlensError):
        GpuBatch([GpuPlens(case)])


def test_dt_parity():
    from plens_amd import wall_flow
    case = wall_flow(nx=6, ny=4, nz=2)
    g, o = _pair(case)
    dg, do = g.calc_time_step(case.courant), o.calc_time_step(case.courant)
    print(f"dt gpu {dg:.17g} cpu {do:.17g}", flush=True)
    assert abs(dg - do) <= 1e-13 * do


def test_run_100_steps_periodic_cube():
    from plens_amd import periodic_cube
    case = periodic_cube(n=4, degree=3)
    g, o = _pair(case)
    tg = g.run(100, case.courant)
    to = o.run(100, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"100 steps: rel L2 {err:.3e}  t gpu {tg:.12g} cpu {to:.12g}", flush=True)
    assert abs(tg - to) <= 1e-12 * to
    assert err <= RUN_TOL


def test_graph_run_matches_stepwise():
    from plens_amd import GpuPlens, shock_tube
    case = shock_tube(nx=16, ny=2, nz=2)
    a = GpuPlens(case)
    b = GpuPlens(case)
    ta = a.run(5, case.courant)
    tb = sum(b.update(case.courant) for _ in range(5))
    assert ta == tb
    assert np.array_equal(a.get_state(), b.get_state())


def test_steps_shock_tube_blended():
    from plens_amd import shock_tube
    case = shock_tube(nx=16, ny=2, nz=2)
    g, o = _pair(case)
    g.run(20, case.courant)
    o.run(20, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"shock 20 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


def test_steps_wall_flow():
    from plens_amd import wall_flow
    case = wall_flow(nx=6, ny=4, nz=2)
    g, o = _pair(case)
    g.run(10, case.courant)
    o.run(10, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"wall 10 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


def test_positivity_failure_reported():
    from plens_amd import GpuPlens, PositivityError, periodic_cube
    case = periodic_cube(n=2, degree=2)
    Q = case.Q0.copy()
    bad = 37
    Q[0, bad] = -1.0  # negative density planted on the host; kernels only flag it
    g = GpuPlens(case)
    g.set_state(Q)
    with pytest.raises(PositivityError) as ei:
        g.calc_rhs()
    assert ei.value.dof == bad and ei.value.kind == "density"


def test_rhs_many_blended_cells():
    """noisy state on a 16^3 periodic cube: thousands of blended cells, more than one pass of the
    subcell kernel's grid over the blended-cell list"""
    from plens_amd import periodic_cube
    case = periodic_cube(n=16, degree=3, flux="HLLC")
    rng = np.random.default_rng(11)
    Q = case.Q0.copy()
    Q[0] *= 1 + 0.2 * rng.uniform(-1, 1, Q.shape[1])
    case.Q0 = Q
    g, o = _rhs_parity(case, "many-blended")
    _, a, _ = o.calc_rhs()
    assert (a > 0).sum() > 2048, (a > 0).sum()


@pytest.mark.parametrize("flux", ["HLLC", "Chandrashekhar"])
def test_rhs_wall_flow_stored_metrics(flux):
    """the stored-metric path (Ja and detJ read from the host arrays) against the oracle"""
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow(nx=6, ny=4, nz=2, flux=flux)
    g, o = GpuPlens(case, metrics="stored"), OraclePlens(case)
    rg, _, _ = g.calc_rhs()
    ro, _, _ = o.calc_rhs()
    err = rel_l2(rg, ro)
    print(f"[wall-{flux} stored metrics] rhs rel L2 = {err:.3e}", flush=True)
    assert err <= RHS_TOL
    g.run(10, case.courant)
    o.run(10, case.courant)
    assert rel_l2(g.get_state(), o.get_state()) <= RUN_TOL


def test_map_must_reproduce_given_metrics():
    """vertices handed over together with metric terms that no trilinear map of them produces"""
    from plens_amd import GpuPlens, wall_flow
    from plens_amd._lib import PLENS_E_INVALID, PlensError
    case = wall_flow(nx=4, ny=2, nz=2)
    case.mesh.Ja[1, 4, 7] *= 1.001
    with pytest.raises(PlensError) as ei:
        GpuPlens(case)
    assert ei.value.code == PLENS_E_INVALID
    GpuPlens(case, metrics="stored")  # the arrays alone are accepted as given


# ---- entropy-variable lifting (the reference's default build, calc_cell_evar_grad) ------------
def _entropy(case):
    from plens_amd import with_gradient_variables
    return with_gradient_variables(case, "entropy")


@pytest.mark.parametrize("flux", ["HLLC", "Chandrashekhar"])
def test_rhs_entropy_lifting_cube(flux):
    from plens_amd import periodic_cube
    _rhs_parity(_entropy(periodic_cube(n=4, degree=3, flux=flux)), f"evar-cube-{flux}")


def test_rhs_entropy_lifting_shock_tube():
    from plens_amd import shock_tube
    _rhs_parity(_entropy(shock_tube(nx=16, ny=2, nz=2)), "evar-shock")


@pytest.mark.parametrize("degree", [2, 4])
def test_rhs_entropy_lifting_wall_flow(degree):
    from plens_amd import wall_flow
    _rhs_parity(_entropy(wall_flow(degree=degree, nx=6, ny=4, nz=2)), f"evar-wall-N{degree}")


def test_run_entropy_lifting():
    from plens_amd import periodic_cube, wall_flow
    for case, steps in ((_entropy(periodic_cube(n=4, degree=3)), 100), (_entropy(wall_flow(nx=6, ny=4, nz=2)), 10)):
        g, o = _pair(case)
        g.run(steps, case.courant)
        o.run(steps, case.courant)
        err = rel_l2(g.get_state(), o.get_state())
        print(f"evar {case.name} {steps} steps rel L2 {err:.3e}", flush=True)
        assert err <= RUN_TOL


# ---- the remaining boundary conditions on the y faces, with y velocity and y variation ----------
@pytest.mark.parametrize("bc_type", ["free", "symmetry", "empty", "insulated_wall", "zpg_inflow"])
@pytest.mark.parametrize("wall_limit", [0.0, 0.5])
def test_rhs_and_steps_remaining_bcs(bc_type, wall_limit):
    from plens_amd import _lib as L
    from plens_amd.cases import make_bc
    from tests.test_host_and_assembly import y_faces_case
    bc = {"free": make_bc(L.BC_FREE), "symmetry": make_bc(L.BC_SYMMETRY), "empty": make_bc(L.BC_EMPTY),
          "insulated_wall": make_bc(L.BC_INSULATED_WALL, vel=(0.3, 0.0, 0.1)),
          "zpg_inflow": make_bc(L.BC_ZPG_INFLOW, vel=(0.5, 0.05, 0.0))}[bc_type]
    case = y_faces_case(bc, n=3, N=3, mu0=2e-3)
    case.physics.wall_blender_limit = wall_limit
    rng = np.random.default_rng(7)
    X, Y = (case.mesh.node_coords[:, d, :].reshape(-1) for d in (0, 1))
    Q = case.Q0.copy()
    Q[0] *= 1 + 0.05 * np.cos(np.pi * Y) + 0.01 * rng.uniform(-1, 1, X.size)
    Q[2] = Q[0] * (0.1 * np.sin(2 * np.pi * X) + 0.05)
    Q[4] = 1.0 / 0.4 + 0.5 * (Q[1] ** 2 + Q[2] ** 2) / Q[0]
    case.Q0 = Q
    g, o = _rhs_parity(case, f"bc-{bc_type}-limit{wall_limit}")
    g.run(10, case.courant)
    o.run(10, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"bc-{bc_type} 10 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


# ---- local time stepping and the steady-state error -------------------------------------------
@pytest.mark.parametrize("limit,start", [(False, -1.0), (True, -1.0), (False, "mid")])
def test_local_stepping_parity(limit, start):
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens
    from tests.test_host_and_assembly import _stretched_tube
    case = _stretched_tube()
    if start == "mid":  # switch on after three global steps, inside one graph-replayed run
        probe = OraclePlens(case)
        start = probe.run(3, case.courant) * 0.999
    g, o = GpuPlens(case), OraclePlens(case)
    kw = dict(start_time=start, limit=limit, neighbor_factor=1.1, global_factor=3.0)
    g.set_local_stepping(True, **kw)
    o.set_local_stepping(True, **kw)
    tg = g.run(6, case.courant)
    to = o.run(6, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    eg, eo = g.ss_error(), o.ss_error()
    print(f"LTS limit={limit} start={start}: 6 steps rel L2 {err:.3e}, ss error gpu {eg:.10e} cpu {eo:.10e}",
          flush=True)
    assert abs(tg - to) <= 1e-12 * to
    assert err <= RUN_TOL
    assert abs(eg - eo) <= 1e-10 * eo


def test_ss_error_global_stepping_and_measurement_calls():
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, shock_tube
    from plens_amd._lib import PlensError
    case = shock_tube(nx=16, ny=2, nz=2)
    g, o = GpuPlens(case), OraclePlens(case)
    with pytest.raises(PlensError):
        g.ss_error()  # no step yet
    g.run(3, case.courant)
    o.run(3, case.courant)
    e0 = g.ss_error()
    g.calc_rhs()
    g.time_kernel("grad", reps=2)
    g.time_kernel("lo", reps=2)
    g.time_kernel("vol", reps=2)
    e1 = g.ss_error()
    print(f"ss error gpu {e0:.10e} cpu {o.ss_error():.10e}", flush=True)
    assert e0 == e1  # RHS evaluations and kernel timing leave the last step's record alone
    assert abs(e0 - o.ss_error()) <= 1e-10 * o.ss_error()
    # calc_time_step replaces the steps the error divides by, as it replaces loc_time_steps
    g.calc_time_step(case.courant)
    o.calc_time_step(case.courant)
    assert abs(g.ss_error() - o.ss_error()) <= 1e-10 * o.ss_error()


def test_batch_rejects_local_stepping():
    from plens_amd import GpuBatch, GpuPlens, periodic_cube
    from plens_amd._lib import PLENS_E_INVALID, PlensError
    case = periodic_cube(n=2, degree=2)
    a, b = GpuPlens(case), GpuPlens(case)
    b.set_local_stepping(True)
    with pytest.raises(PlensError) as ei:
        GpuBatch([a, b])
    assert ei.value.code == PLENS_E_INVALID


@pytest.mark.parametrize("label", ["cell-N3", "cell-N5", "shock-2cells"])
def test_smallest_meshes(label):
    """one fully periodic cell (every face its own neighbour) and a two-cell blended shock tube:
    the smallest grids of every kernel (one face-task group, one k_alpha block, one blended cell)"""
    from plens_amd import periodic_cube, shock_tube
    case = {"cell-N3": lambda: periodic_cube(n=1, degree=3), "cell-N5": lambda: periodic_cube(n=1, degree=5),
            "shock-2cells": lambda: shock_tube(nx=2, ny=1, nz=1)}[label]()
    g, o = _rhs_parity(case, label)
    g.run(10, case.courant)
    o.run(10, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"[{label}] 10 steps rel L2 {err:.3e}", flush=True)
    assert err <= RUN_TOL


def test_hllc_normals_near_the_x_axis():
    """HLLC where face and subcell normals lie within 1e-3 of the x axis (a wall flow perturbed by
    2e-4 of the spacing): the reference's rotation takes such a normal as (sign n_x, 0, 0), and the
    face and subcell kernels must do the same; the operator is checked with the GPU's blender values
    (their own tolerance: rtol 1e-10)"""
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow(nx=12, ny=6, nz=2, perturb=2e-4)
    nr = case.mesh.face_normals.reshape(case.mesh.n_cells, 6, 3, -1)[:, 0:2]
    tilt = np.sqrt(nr[:, :, 1] ** 2 + nr[:, :, 2] ** 2)
    assert ((tilt > 0) & (tilt <= 1e-3)).sum() > 1000
    _rhs_parity(case, "near-x normals, initial")
    g, o = GpuPlens(case), OraclePlens(case)
    g.run(15, case.courant)
    o.set_state(g.get_state())
    rg, ag, adg = g.calc_rhs()
    _, ao, ado = o.calc_rhs()
    assert (ag > 0).sum() > 0 and np.array_equal(ag > 0, ao > 0)
    assert np.allclose(ag, ao, rtol=1e-10, atol=1e-13) and np.allclose(adg, ado, rtol=1e-10, atol=1e-13)
    o.set_alpha(ag, adg)
    ro, _, _ = o.calc_rhs()
    err = rel_l2(rg, ro)
    print(f"near-x normals, 15 steps in ({int((ag > 0).sum())} blended cells): operator rel L2 {err:.3e}", flush=True)
    assert err <= RHS_TOL


# every cells-per-workgroup count of at most four wavefronts (the autotuner picks among the largest
# count for each wavefront count; the others check that partial cells in a wavefront are handled)
_SHAPES = [(2, c) for c in range(1, 10)] + [(3, c) for c in range(1, 5)] + [(4, 1), (4, 2), (5, 1)]


@pytest.mark.parametrize("degree,cpb", _SHAPES)
def test_cells_per_workgroup_shapes(degree, cpb):
    """RHS and 5 steps of the blended shock tube with cpb cells per workgroup against the oracle"""
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, shock_tube
    case = shock_tube(degree=degree, nx=16, ny=2, nz=2)
    nn = (degree + 1) ** 3
    assert ((cpb * nn + 63) // 64) * 64 <= 256
    g, o = GpuPlens(case, cells_per_block=cpb), OraclePlens(case)
    rg, ag, _ = g.calc_rhs()
    ro, ao, _ = o.calc_rhs()
    err_rhs = rel_l2(rg, ro)
    assert (ao > 0).sum() > 0 and np.allclose(ag, ao, rtol=1e-10, atol=1e-13)
    g.run(5, case.courant)
    o.run(5, case.courant)
    err = rel_l2(g.get_state(), o.get_state())
    print(f"N={degree} cpb={cpb}: rhs rel L2 {err_rhs:.3e}, 5 steps {err:.3e}", flush=True)
    assert err_rhs <= RHS_TOL
    assert err <= RUN_TOL
