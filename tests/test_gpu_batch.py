"""Batched uneven solves (plens_gpu_batch_*): instances of different degrees, sizes, boundary
conditions and viscosities advanced by one graph launch per step must reproduce the same instances
advanced one by one bit for bit (same kernels, per-instance time step and viscosity), and the
oracle; the autotuner's tuning file is written once and reused."""
import numpy as np
import pytest

from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu


def _small_instances(k=10, max_n=8):
    from plens_amd import batch_instance
    return [batch_instance(i, max_n=max_n) for i in range(k)]


def test_batch_matches_sequential_bitwise(tmp_path):
    from plens_amd import GpuBatch, GpuPlens
    cases = _small_instances()
    degrees = {c.meta["degree"] for c in cases}
    assert len(degrees) >= 3, degrees
    seq = [GpuPlens(c) for c in cases]
    batch = GpuBatch([GpuPlens(c) for c in cases], tuning_db=str(tmp_path / "tune.txt"))
    tb = batch.run(5, [c.courant for c in cases])
    for i, (c, g) in enumerate(zip(cases, seq)):
        ts = g.run(5, c.courant)
        print(f"instance {i} {c.meta['kind']} N={c.meta['degree']} cells={c.meta['cells']} "
              f"t batch {tb[i]:.12g} seq {ts:.12g}", flush=True)
        assert tb[i] == ts
        assert np.array_equal(batch.get_state(i), g.get_state())
    print(batch.groups(), flush=True)


def test_batch_matches_oracle():
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuBatch, GpuPlens
    cases = _small_instances(k=4, max_n=4)
    batch = GpuBatch([GpuPlens(c) for c in cases])
    tb = batch.run(10, [c.courant for c in cases])
    for i, c in enumerate(cases):
        o = OraclePlens(c)
        to = o.run(10, c.courant)
        err = rel_l2(batch.get_state(i), o.get_state())
        print(f"instance {i}: rel L2 vs oracle {err:.3e}", flush=True)
        assert abs(tb[i] - to) <= 1e-12 * to
        assert err <= 1e-10


def test_autotuner_file_reused(tmp_path):
    from plens_amd import GpuBatch, GpuPlens
    cases = _small_instances(k=6)
    db = tmp_path / "tune.txt"
    b1 = GpuBatch([GpuPlens(c) for c in cases], tuning_db=str(db))
    shapes1 = b1.groups()
    text = db.read_text()
    lines = [ln for ln in text.splitlines() if ln and not ln.startswith("#")]
    assert len(lines) == len(shapes1)
    b2 = GpuBatch([GpuPlens(c) for c in cases], tuning_db=str(db))
    assert b2.groups() == shapes1
    assert db.read_text() == text


def test_batch_fp32_bulk_error_per_instance():
    from plens_amd import GpuBatch, GpuPlens, with_precision
    cases = _small_instances(k=8)
    ref = GpuBatch([GpuPlens(c) for c in cases])
    mix = GpuBatch([GpuPlens(with_precision(c, "fp32_bulk")) for c in cases])
    ref.run(10, 0.4)
    mix.run(10, 0.4)
    for i in range(len(cases)):
        err = rel_l2(mix.get_state(i), ref.get_state(i))
        print(f"instance {i}: fp32-bulk vs fp64 rel L2 {err:.3e}", flush=True)
        assert err <= 1e-6


def test_batch_positivity_names_instance():
    from plens_amd import GpuBatch, GpuPlens, PositivityError, periodic_cube
    cases = _small_instances(k=3)
    bad = periodic_cube(n=2, degree=2, flux="HLLC")  # same surface flux as the other instances
    Q = bad.Q0.copy()
    Q[0, 5] = -1.0  # planted on the host; the kernels only flag it
    bad.Q0 = Q
    batch = GpuBatch([GpuPlens(c) for c in cases + [bad]])
    with pytest.raises(PositivityError) as ei:
        batch.run(1, 0.4)
    assert ei.value.instance == 3 and ei.value.dof == 5 and ei.value.kind == "density"


def test_streaming_kernels_match_the_cached_ones_bitwise():
    """The measured wall flow alone (2,048,000 DoF) runs the streaming face and volume kernels (its
    states and auxiliary variables, 112 B per DoF, fit the 256 MB Infinity Cache, with the RK
    registers they exceed the 176 MiB of it that holds them in practice); merged in a batch with a
    second wall flow (524,288 DoF) its group no longer fits and the cached kernels run. The cache
    hints change no arithmetic: the same state bit for bit."""
    from plens_amd import GpuBatch, GpuPlens, wall_flow
    cache, stream_from = 256 * 2 ** 20, 176 * 2 ** 20

    def streams(n):
        return 112 * n <= cache and 192 * n > stream_from

    wf, wf2 = wall_flow(), wall_flow(nx=64, ny=32, nz=2)
    assert streams(wf.mesh.n_dof) and not streams(wf.mesh.n_dof + wf2.mesh.n_dof)
    solo = GpuPlens(wf)
    batch = GpuBatch([GpuPlens(wf), GpuPlens(wf2)])
    ts = solo.run(5, wf.courant)
    tb = batch.run(5, [wf.courant, wf2.courant])
    assert tb[0] == ts
    assert np.array_equal(solo.get_state(), batch.get_state(0))


@pytest.mark.parametrize("degree,precision,n_main,n_extra", [(2, "fp64", 40, 30), (3, "fp64", 32, 20),
                                                            (5, "fp32_bulk", 22, 20)])
def test_streaming_kernels_match_the_cached_ones_per_degree(degree, precision, n_main, n_extra):
    """As above for periodic cubes at other degrees and in FP32-bulk (auxiliary variables of
    9 sizeof(TF) bytes): the cube alone streams, merged with a second cube its group is cached."""
    from plens_amd import GpuBatch, GpuPlens, periodic_cube, with_precision
    cache, stream_from, fb = 256 * 2 ** 20, 176 * 2 ** 20, 8 if precision == "fp64" else 4

    def streams(n):
        return (40 + 9 * fb) * n <= cache and (120 + 9 * fb) * n > stream_from

    main = with_precision(periodic_cube(degree=degree, n=n_main, flux="HLLC"), precision)
    extra = with_precision(periodic_cube(degree=degree, n=n_extra, flux="HLLC", seed=3), precision)
    assert streams(main.mesh.n_dof) and not streams(main.mesh.n_dof + extra.mesh.n_dof)
    solo = GpuPlens(main)
    batch = GpuBatch([GpuPlens(main), GpuPlens(extra)])
    ts = solo.run(5, main.courant)
    tb = batch.run(5, [main.courant, extra.courant])
    assert tb[0] == ts
    assert np.array_equal(solo.get_state(), batch.get_state(0))
