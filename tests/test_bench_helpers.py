"""CPU checks of the measurement helpers of bench.py: the algorithmic byte counts per DoF, and the
parsers that read the committed rocprofv3 summaries (calibrated traffic, VALU flops, VALU busy)."""
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_bytes_per_dof_at_degree_4():
    # FP64, viscous: k_grad Q 5 + map 24/125 doubles + avars 9; k_vol reads Q 5, R' 4, R'' 3 (stage
    # averages), avars 9, face terms 5 x 6 x 25 / 125 and the map, writes Q 5 and R 5; k_face per
    # face node (0.6 per DoF)
    # 13 doubles of states and normal, the two cells' maps (48 doubles per 25-node face), 18 avars, 10
    # written terms
    b = bench.kernel_bytes_per_dof(4)
    m = 24 * 8 / 125
    assert b["grad"] == pytest.approx(8 * 14 + m)
    assert b["vol"] == pytest.approx(8 * (5 + 4 + 3 + 9 + 6 + 5 + 5) + m)
    assert b["face"] == pytest.approx(0.6 * (8 * 13 + 2 * 24 * 8 / 25 + 8 * 28))
    assert b["face"] < bench.kernel_bytes_per_dof(4, face_maps=False)["face"]
    # inviscid: no avars anywhere
    bi = bench.kernel_bytes_per_dof(4, viscous=False)
    assert b["grad"] - bi["grad"] == pytest.approx(72)
    # the two-pass stage figure: Q in and out, R out, R' and R'' in, avars out and in, the maps
    assert bench.stage_bytes_per_dof(4) == pytest.approx(8 * 22 + 8 * 18 + m)
    # FP32-bulk: avars in FP32, the RK registers FP64
    assert bench.stage_bytes_per_dof(4, fb=4) == pytest.approx(8 * 22 + 4 * 18 + m)
    assert bench.kernel_bytes_per_dof(4, fb=4)["vol"] == pytest.approx(8 * 22 + 4 * 15 + m)


# the measured wall flow (2,048,000 DoF) runs the streaming volume kernel in both precisions: its
# name carries the streaming template argument
@pytest.mark.parametrize("precision,kernel", [("fp64", "k_vol<4,double,false,true,true>"),
                                              ("fp32_bulk", "k_vol<4,float,false,true,true>")])
def test_profile_parsers_find_the_dominant_kernel(precision, kernel):
    assert bench.streaming_handle(2_048_000, 8 if precision == "fp64" else 4)
    traffic, src = bench.measured_traffic(precision, kernel)
    assert src is not None and traffic > 1e8
    flops, fsrc, busy, _ = bench.measured_compute(precision, kernel)
    assert fsrc is not None and flops["F64"] > 1e9
    if precision == "fp32_bulk":
        assert flops["F32"] > 1e8
    else:
        assert 0.0 < busy <= 1.0


def test_profile_parsers_return_none_for_unknown_kernels():
    assert bench.measured_traffic("fp64", "k_nothing<4>") == (None, None)
    flops, src, busy, bsrc = bench.measured_compute("fp64", "k_nothing<4>")
    assert flops == {} and src is None and busy is None and bsrc is None


def test_algorithmic_flops_of_a_small_case():
    # the reduction from the reference-as-written count to the minimal algorithm on a viscous
    # periodic cube: 3 (N+1) ordered pair fluxes per node against 3 (N/2 + 2/(N+1)); each saving is
    # non-negative, and the subcell residual is kept only for the blended fraction of the cells
    from plens_amd import periodic_cube
    case = periodic_cube(n=4, degree=3)
    af = bench.algorithmic_flops(case, case.Q0)
    N = 3
    assert af["pair_fluxes_per_node_ref"] == pytest.approx(3 * (N + 1))
    assert af["pair_fluxes_per_node_min"] == pytest.approx(3 * (N / 2 + 2 / (N + 1)))
    sv = af["minimal_savings"]
    assert all(v >= 0 for v in sv.values())
    fb = af["blended_cell_fraction"]
    assert 0.0 <= fb <= 1.0
    assert sv["subcell_residual_blended_only"] == pytest.approx(af["sections_ref"]["lo inv residual"] * (1 - fb))
    assert af["min_flop_per_dof_stage"] == pytest.approx(af["ref_flop_per_dof_stage"] - sum(sv.values()))
    assert 0 < af["min_flop_per_dof_stage"] < af["ref_flop_per_dof_stage"]
    # the dense change of basis is 2 n flops per DoF and RHS; its sum-factorised form 3 (N+1) x 2
    assert af["sections_ref"]["dense change of basis"] == pytest.approx(2 * (N + 1) ** 3)
    assert sv["sum_factorised_change_of_basis"] == pytest.approx(2 * (N + 1) ** 3 - 6 * (N + 1))


def test_streaming_rule_of_the_kernel_names():
    # the measured wall flow (2,048,000 DoF) runs the streaming face/volume kernels in both
    # precisions; the shock tube (524,288 DoF) fits in the Infinity Cache with its RK registers and a
    # 34^3 N=4 cube (4,913,000 DoF) cannot keep its gathered data there either way
    assert bench.streaming_handle(2_048_000, 8) and bench.streaming_handle(2_048_000, 4)
    assert not bench.streaming_handle(524_288, 8)
    assert not bench.streaming_handle(4_913_000, 8)
    # at the measured crossovers: N=4 cubes of 22^3 cells (1,331,000 DoF, FP64) stream, 21^3
    # (1,157,625 DoF, FP32-bulk) do not
    assert bench.streaming_handle(1_331_000, 8) and not bench.streaming_handle(1_157_625, 4)
