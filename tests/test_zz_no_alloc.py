"""The allocation-free steady-state error (plens_gpu_ss_error_noalloc, called through the C ABI)
against the allocating entry point plens_gpu_ss_error: equal bit for bit, on two calls in a row
whose second has fewer cells than the first (the device scratch keeps the first call's contents)."""
import ctypes as C

import pytest

pytestmark = pytest.mark.gpu


def _ss_both(g):
    a, b = C.c_double(), C.c_double()
    assert g.lib.plens_gpu_ss_error(g._h, C.byref(a)) == 0
    assert g.lib.plens_gpu_ss_error_noalloc(g._h, C.byref(b)) == 0
    return a.value, b.value


def test_noalloc_ss_error_equals_allocating_entry_point():
    from plens_amd import GpuPlens, shock_tube, wall_flow
    big = wall_flow(nx=12, ny=8, nz=2)      # 192 cells, N=4
    small = shock_tube(nx=16, ny=2, nz=2)   # 64 cells, N=3
    results = []
    for case in (big, small):
        g = GpuPlens(case)
        e = C.c_double()
        # before any step there is no steady-state error to report
        assert g.lib.plens_gpu_ss_error_noalloc(g._h, C.byref(e)) == -1
        g.run(2, case.courant)
        a, b = _ss_both(g)
        print(f"[{case.name}] cells {case.mesh.n_cells}: ss error {a:.17g} / noalloc {b:.17g}", flush=True)
        assert a > 0 and a == b
        results.append(b)
        g.close()
    assert results[0] != results[1]
