"""Data output and checkpoints on the GPU (through the C ABI): the fields of PLENS::write() against
the oracle, no effect of an output on the run, the written VTU, and checkpoint/restart continuity
and rejection of foreign or damaged files."""
import numpy as np
import pytest

from tests.conftest import rel_l2
from tests.vtu_read import read_vtu

pytestmark = pytest.mark.gpu


# relative L2 by block: stress tensor, heat flux, rho, p, T, velocity (a block near zero, such as
# the stresses of a uniform state, is measured against the block's own size, not per component)
BLOCKS = {"tau": slice(0, 6), "q": slice(6, 9), "rho": slice(9, 10), "p": slice(10, 11), "T": slice(11, 12),
          "vel": slice(12, 15)}


def _compare(g, o, courant, t, label, tol=1e-12, tol_av=1e-12):
    ng, cg = g.output_fields(courant, t)
    no, co = o.output_fields(courant, t)
    errs = {k: rel_l2(ng[b], no[b]) for k, b in BLOCKS.items()}
    print(f"[{label}] output fields rel L2 " + " ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert max(errs[k] for k in ("rho", "p", "T", "vel")) <= tol, errs
    assert max(errs["tau"], errs["q"]) <= tol_av, errs
    assert np.allclose(cg[0], co[0], rtol=1e-10, atol=1e-13) and np.allclose(cg[1], co[1], rtol=1e-10, atol=1e-13)
    assert np.abs(cg[2] - co[2]).max() <= 1e-12 * np.abs(co[2]).max()
    return ng, cg


@pytest.mark.parametrize("which", ["wall", "shock", "curved"])
def test_output_fields_parity(which):
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, cylinder_flare, shock_tube, wall_flow
    case = {"wall": lambda: wall_flow(nx=6, ny=4, nz=2), "shock": lambda: shock_tube(nx=16, ny=2, nz=2),
            "curved": lambda: cylinder_flare(degree=3, mapping_degree=2, n_axial=6, n_normal=2, n_azim=2)}[which]()
    g, o = GpuPlens(case), OraclePlens(case)
    _compare(g, o, case.courant, 0.0, which)
    g.run(3, case.courant)
    o.run(3, case.courant)
    _compare(g, o, case.courant, 0.1, which + "-3steps")


def test_output_fields_with_local_stepping():
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow(nx=6, ny=4, nz=2)
    g, o = GpuPlens(case), OraclePlens(case)
    for h in (g, o):
        h.set_local_stepping(True, start_time=-1.0, limit=True, neighbor_factor=1.05, global_factor=3.0)
    _, cg = _compare(g, o, case.courant, 0.0, "lts")
    assert cg[2].max() > cg[2].min()


def test_output_fields_fp32_bulk():
    from oracle.pyoracle import OraclePlens
    from plens_amd import GpuPlens, wall_flow, with_precision
    case = wall_flow(nx=6, ny=4, nz=2)
    _compare(GpuPlens(with_precision(case, "fp32_bulk")), OraclePlens(case), case.courant, 0.0, "fp32", tol=1e-12,
             tol_av=1e-4)


def test_output_does_not_disturb_the_run():
    from plens_amd import GpuPlens, shock_tube
    case = shock_tube(nx=16, ny=2, nz=2)
    a, b = GpuPlens(case), GpuPlens(case)
    a.run(4, case.courant)
    b.run(4, case.courant)
    q = b.get_state()
    b.output_fields(case.courant, 0.0)
    assert np.array_equal(b.get_state(), q)
    a.run(4, case.courant)
    b.run(4, case.courant)
    assert np.array_equal(a.get_state(), b.get_state())


def test_write_output_files(tmp_path):
    from plens_amd import GpuPlens, wall_flow
    case = wall_flow(nx=4, ny=2, nz=2)
    g = GpuPlens(case)
    nodal, cell = g.write_output(tmp_path, "wall", 0, case.courant, 0.0, [0.0])
    v = read_vtu(tmp_path / "wall_000000.0.vtu")
    for f, nm in enumerate(["txx", "txy", "txz", "tyy", "tyz", "tzz", "qx", "qy", "qz", "rho", "p", "T", "u", "v", "w"]):
        assert np.array_equal(v["point_data"][nm], nodal[f])
    assert np.array_equal(v["point_data"]["alpha"], np.repeat(cell[0], case.mesh.n))


def test_checkpoint_restart_is_bitwise(tmp_path):
    from plens_amd import GpuPlens, shock_tube
    case = shock_tube(nx=16, ny=2, nz=2)
    a = GpuPlens(case)
    t = a.run(5, case.courant)
    a.write_checkpoint(tmp_path / "s.rst", t, step=5, output_counter=2)
    a.run(5, case.courant)
    b = GpuPlens(case)
    tb, step, oc = b.read_checkpoint(tmp_path / "s.rst")
    assert (tb, step, oc) == (t, 5, 2)
    b.run(5, case.courant)
    assert np.array_equal(a.get_state(), b.get_state())


def test_checkpoint_rejects_foreign_or_damaged_files(tmp_path):
    from plens_amd import GpuPlens, wall_flow
    from plens_amd._lib import PlensError
    case = wall_flow(nx=6, ny=4, nz=2)
    g = GpuPlens(case)
    g.write_checkpoint(tmp_path / "c.rst", 1.0)
    raw = (tmp_path / "c.rst").read_bytes()
    other = GpuPlens(wall_flow(nx=6, ny=4, nz=2, seed=9))  # same sizes, other mesh
    with pytest.raises(PlensError):
        other.read_checkpoint(tmp_path / "c.rst")
    q = g.get_state()
    bad = bytearray(raw)
    bad[200] ^= 1
    (tmp_path / "bad.rst").write_bytes(bytes(bad))
    (tmp_path / "short.rst").write_bytes(raw[:-8])
    (tmp_path / "long.rst").write_bytes(raw + b"\\0")
    for name in ("bad.rst", "short.rst", "long.rst", "missing.rst"):
        with pytest.raises(PlensError):
            g.read_checkpoint(tmp_path / name)
    assert np.array_equal(g.get_state(), q)
    assert g.read_checkpoint(tmp_path / "c.rst")[0] == 1.0
