"""The flop-counting build of the oracle (oracle/liboracle_flops.so, Real = the counting double of
oracle/flopcount.h) computes what the oracle proper computes, bit for bit, and its counters follow
the reference's loop structure; the algorithmic flop figure of the bench is derived from them."""
import numpy as np

from oracle.pyoracle import FLOPS_LIB_PATH, OraclePlens
from plens_amd import periodic_cube


def test_counting_build_matches_the_oracle_bitwise():
    case = periodic_cube(n=2, degree=3)
    a = OraclePlens(case)
    b = OraclePlens(case, lib_path=FLOPS_LIB_PATH)
    b.flop_counts(reset=True)
    a.update(case.courant)
    b.update(case.courant)
    assert np.array_equal(a.get_state(), b.get_state())
    f, k = b.flop_counts(reset=True)
    N, n_dof, n_cells = case.mesh.degree, case.mesh.n_dof, case.mesh.n_cells
    # calls of the nested buckets per RK step (5 RHS evaluations): one two-point volume flux for
    # every node, line partner (itself included) and direction; the same count of viscous fluxes in
    # the ho viscous volume term; one dense change of basis per cell
    assert k[10] == 5 * n_dof * 3 * (N + 1)
    assert k[13] == 5 * n_dof * 3 * (N + 1)
    assert k[11] == 5 * n_cells
    # the dense change of basis is n^2 multiply-adds per cell (2 flops each)
    n = (N + 1) ** 3
    assert f[11] == 5 * n_cells * 2 * n * n
    # the RK update per value (rk4_step): the time-step product at every stage, then 2 operations at
    # stage 1 (mul, add), 5 at stage 2 (coefficient difference, 2 mul, 2 add), 8 at stages 3-5
    # (2 coefficient differences, 3 mul, 3 add)
    assert f[9] == 5 * n_dof * (5 * 1 + 2 + 5 + 3 * 8)
    a.close()
    b.close()


def test_counter_counts_a_closed_form():
    # the viscous flux along a direction (navier_stokes.cc:1644-1652) has a fixed operation count:
    # the counting build attributes exactly that many flops to every call
    case = periodic_cube(n=2, degree=2)
    b = OraclePlens(case, lib_path=FLOPS_LIB_PATH)
    b.flop_counts(reset=True)
    b.calc_rhs()
    f, k = b.flop_counts(reset=True)
    assert k[13] > 0 and f[13] % k[13] == 0  # every call the same number of operations
    per_call = f[13] // k[13]
    # stress times direction (9 mul, 9 add into the zeroed vector), 1/rho (1 div), per direction
    # the velocity (1 mul) and the energy term -q_d dir_d + v_d (tau dir)_d added in (2 mul, 2 add):
    # 18 + 1 + 3 x 5 = 34; negation is free
    assert per_call == 34, per_call
    b.close()
