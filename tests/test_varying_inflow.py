"""Varying inflow BC (IBCs/varying_inflow.cc): the expression compiler of the library and the
oracle's evaluator against closed forms of muParser/deal.II FunctionParser semantics and against
each other, and the oracle's BC against the uniform inflow it generalises. CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import pyoracle as PO
from plens_amd import _lib as L
from plens_amd.cases import make_bc, varying_inflow_tube

dp = C.POINTER(C.c_double)

# (expression, (x, y, z, t), value): precedence, associativity and every function
CASES = [
    ("-2^2", (0, 0, 0, 0), -4.0), ("2^3^2", (0, 0, 0, 0), 512.0), ("2^-1", (0, 0, 0, 0), 0.5),
    ("-x^2", (3, 0, 0, 0), -9.0), ("1+2*3", (0, 0, 0, 0), 7.0), ("(1+2)*3", (0, 0, 0, 0), 9.0),
    ("8/2/2", (0, 0, 0, 0), 2.0), ("8-2-2", (0, 0, 0, 0), 4.0), ("2*-3", (0, 0, 0, 0), -6.0), ("+x", (5, 0, 0, 0), 5.0),
    ("1 < 2 ? 3 : 4", (0, 0, 0, 0), 3.0), ("0 ? 1 : 0 ? 2 : 3", (0, 0, 0, 0), 3.0),
    ("x > 1 ? y : z", (2, 5, 7, 0), 5.0), ("2 > 1 && 0 || 1", (0, 0, 0, 0), 1.0), ("1 && 0", (0, 0, 0, 0), 0.0),
    ("1 <= 1", (0, 0, 0, 0), 1.0), ("1 >= 2", (0, 0, 0, 0), 0.0), ("3 == 3", (0, 0, 0, 0), 1.0),
    ("3 != 3", (0, 0, 0, 0), 0.0), ("1 + 2 < 4", (0, 0, 0, 0), 1.0),
    ("min(3, 1, 2)", (0, 0, 0, 0), 1.0), ("max(3, 1, 2)", (0, 0, 0, 0), 3.0), ("sum(1, 2, 3)", (0, 0, 0, 0), 6.0),
    ("avg(1, 2, 6)", (0, 0, 0, 0), 3.0), ("if(0, 1, 2)", (0, 0, 0, 0), 2.0), ("pow(2, 10)", (0, 0, 0, 0), 1024.0),
    ("atan2(1, 1)", (0, 0, 0, 0), math.pi / 4), ("log(_e)", (0, 0, 0, 0), 1.0), ("ln(_e^2)", (0, 0, 0, 0), 2.0),
    ("log10(1000)", (0, 0, 0, 0), 3.0), ("log2(8)", (0, 0, 0, 0), 3.0), ("sign(-2)", (0, 0, 0, 0), -1.0),
    ("sign(0)", (0, 0, 0, 0), 0.0), ("rint(2.5)", (0, 0, 0, 0), 2.0), ("abs(-1.5)", (0, 0, 0, 0), 1.5),
    ("floor(-1.5)", (0, 0, 0, 0), -2.0), ("ceil(-1.5)", (0, 0, 0, 0), -1.0), ("int(-2.7)", (0, 0, 0, 0), -2.0),
    ("cot(_pi/4)", (0, 0, 0, 0), 1.0), ("csc(_pi/2)", (0, 0, 0, 0), 1.0), ("sec(0)", (0, 0, 0, 0), 1.0),
    ("erfc(0)", (0, 0, 0, 0), 1.0), ("sqrt(16) + exp(0)", (0, 0, 0, 0), 5.0),
    ("sin(_pi/6) + cos(0) + tan(_pi/4)", (0, 0, 0, 0), 2.5), ("asin(1) + acos(1) + atan(1)", (0, 0, 0, 0), 3 * math.pi / 4),
    ("sinh(0) + cosh(0) + tanh(0)", (0, 0, 0, 0), 1.0), ("asinh(0) + acosh(1) + atanh(0)", (0, 0, 0, 0), 0.0),
    ("x + 2*y - z/t", (1, 2, 3, 4), 4.25), ("1.5e-3*1e3 + .5", (0, 0, 0, 0), 2.0),
    # the double-Mach-reflection inflow of study1/dmr/scripts/base_input.prm
    ("( x - 1.3213672 - 11.547005*t > 0 ? 1 : 116.5 )", (2.0, 0, 0, 0.01), 1.0),
    ("( x - 1.3213672 - 11.547005*t > 0 ? 1 : 116.5 )", (1.0, 0, 0, 0.01), 116.5),
]


def lib_eval(expr, xyzt, n=1):
    out = np.zeros(n)
    rc = L.load().plens_expr_eval(expr.encode(), n, *map(float, xyzt), out.ctypes.data_as(dp))
    return rc, out


def oracle_eval(expr, xyzt):
    return PO.load().oracle_eval_expression(expr.encode(), *map(float, xyzt))


@pytest.mark.parametrize("expr,xyzt,value", CASES)
def test_expression_semantics(expr, xyzt, value):
    rc, v = lib_eval(expr, xyzt)
    assert rc == 0 and abs(v[0] - value) <= 1e-14 * max(1.0, abs(value)), (expr, v[0])
    assert abs(oracle_eval(expr, xyzt) - value) <= 1e-14 * max(1.0, abs(value)), expr


def test_components_and_errors():
    rc, v = lib_eval("1; x + 1; t*2", (1, 0, 0, 3), 3)
    assert rc == 0 and list(v) == [1.0, 2.0, 6.0]
    assert lib_eval("1; 2", (0, 0, 0, 0), 3)[0] == -1  # wrong number of components
    for bad in ("1 +", "(1", "foo(1)", "q", "1 ? 2", "sin(1, 2)", "2 3", ""):
        assert lib_eval(bad, (0, 0, 0, 0))[0] == -1, bad
        assert math.isnan(oracle_eval(bad, (0, 0, 0, 0))), bad


def test_compiler_agrees_with_the_oracle_on_random_expressions():
    rng = np.random.default_rng(11)
    ops = ["+", "-", "*", "/", "^", "<", ">", "&&", "||"]
    fns = ["sin", "cos", "exp", "abs", "sqrt", "tanh", "atan"]
    leaves = ["x", "y", "z", "t", "1.5", "2", "0.25", "_pi"]

    def gen(d):
        if d == 0 or rng.random() < 0.3:
            return str(rng.choice(leaves))
        r = rng.random()
        if r < 0.5:
            return f"({gen(d - 1)} {rng.choice(ops)} {gen(d - 1)})"
        if r < 0.7:
            return f"-{gen(d - 1)}"
        if r < 0.85:
            return f"{rng.choice(fns)}({gen(d - 1)})"
        return f"({gen(d - 1)} > {gen(d - 1)} ? {gen(d - 1)} : {gen(d - 1)})"

    n = 0
    for _ in range(300):
        e = gen(4)
        xyzt = rng.uniform(0.1, 1.5, 4)
        rc, v = lib_eval(e, xyzt)
        o = oracle_eval(e, xyzt)
        assert rc == 0, e
        if math.isfinite(o):
            assert abs(v[0] - o) <= 1e-13 * max(1.0, abs(o)), (e, v[0], o)
            n += 1
        else:
            assert not math.isfinite(v[0]) or math.isnan(o), e
    assert n > 200


def test_oracle_node_locations_are_the_hosts():
    from oracle.pyoracle import OraclePlens
    case = varying_inflow_tube(degree=3, nx=2, ny=2, nz=1)
    o = OraclePlens(case)
    m = case.mesh
    X = np.zeros(3)
    for c in range(m.n_cells):
        for node in range(m.n):
            o.lib.oracle_node_location(o._h, c, node, X.ctypes.data_as(dp))
            assert np.abs(X - m.node_coords[c, :, node]).max() < 1e-15


def test_constant_varying_inflow_is_the_uniform_inflow():
    """p, T and velocity given as constant expressions reproduce the uniform inflow (same
    prim_to_cons with rho = p / (R T)); a time-dependent expression equals it at the matching time
    only, with the BC time set by set_time (calc_rhs) and by the step's start time (steps)."""
    from oracle.pyoracle import OraclePlens
    base = varying_inflow_tube(degree=3, nx=4, ny=2, nz=2)
    uni = make_bc(L.BC_UNIFORM_INFLOW, p=1.2, T=0.9, vel=(2.0, 0.1, -0.05))
    r_uni = OraclePlens(mesh=base.mesh, physics=base.physics, bcs=[uni] + base.bcs[1:])
    r_uni.set_state(base.Q0)
    ru = r_uni.calc_rhs()[0]
    const = make_bc(L.BC_VARYING_INFLOW, p_expr="1.2", T_expr="0.9", vel_expr="2; 0.1; -0.05")
    o = OraclePlens(mesh=base.mesh, physics=base.physics, bcs=[const] + base.bcs[1:])
    o.set_state(base.Q0)
    assert np.array_equal(o.calc_rhs()[0], ru)
    timed = make_bc(L.BC_VARYING_INFLOW, p_expr="1.2 + (t - 0.5)", T_expr="0.9", vel_expr="2; 0.1; -0.05")
    o = OraclePlens(mesh=base.mesh, physics=base.physics, bcs=[timed] + base.bcs[1:])
    o.set_state(base.Q0)
    o.set_time(0.5)
    assert np.array_equal(o.calc_rhs()[0], ru)
    o.set_time(0.7)
    assert not np.array_equal(o.calc_rhs()[0], ru)
