"""plens_gpu_create rejects connectivity the kernels cannot take (checked before any device call, so
these run on the CPU): non-reciprocal neighbours, an interior face whose neighbour is the cell
itself, a periodic face whose partner face is not periodic, and neighbour faces whose nodes do not
coincide in the standard orientation (the face-node maps of both sides assume it; the reference
matches any orientation by location, dgsem/plens.cc:1574-1627). plens_mesh_build_curved rejects
shared faces whose vertex ids differ in the standard order."""
import ctypes as C

import numpy as np
import pytest

from plens_amd import _lib as L
from plens_amd import cases


# vertex a + 2b + 4c of a cell turned by 90 degrees about its x axis is vertex a + 2(1-c) + 4b of the
# unturned cell
ROT_VERTICES = [a + 2 * (1 - c) + 4 * b for c in (0, 1) for b in (0, 1) for a in (0, 1)]


def create_rc(mesh, bcs, with_vertices=True, mutate=None):
    """return code of plens_gpu_create for the mesh, after mutate(arrays) edits copies of its tables"""
    lib = L.load()
    nbr, bid, V = mesh.face_nbr.copy(), mesh.face_bid.copy(), mesh.vertices.copy()
    X = mesh.node_coords.copy()
    if mutate:
        mutate(nbr, bid, V, X)
    d = mesh.desc(with_vertices=with_vertices)
    ip, dp = C.POINTER(C.c_int32), C.POINTER(C.c_double)
    d.face_nbr, d.face_bid = nbr.ctypes.data_as(ip), bid.ctypes.data_as(ip)
    if with_vertices:
        d.vertices = V.ctypes.data_as(dp)
    d.node_coords = X.ctypes.data_as(dp)
    arr = (L.BC * len(bcs))(*bcs)
    h = C.c_void_p()
    rc = lib.plens_gpu_create(C.byref(d), C.byref(cases.make_physics()), arr, len(bcs), C.byref(h))
    if rc == 0:
        lib.plens_gpu_destroy(h)
    return rc


def box():
    m = cases.Mesh(2, 3, 2, 2, (0, 0, 0), (1, 1, 1), (1, 0, 1), [-1, -1, 1, 1, -1, -1], (0, -1, 0),
                   perturb_frac=0.05, seed=3)
    return m, [cases.make_bc(L.BC_PERIODIC), cases.make_bc(L.BC_FREE)]


def test_valid_box_passes_validation():
    m, bcs = box()
    for wv in (True, False):  # vertices, or DoF locations only
        assert create_rc(m, bcs, with_vertices=wv) in (L.PLENS_OK, L.PLENS_E_NODEVICE)


def _non_reciprocal(nbr, bid, V, X):
    c = 1  # interior x+ neighbour of cell 1 is cell 2; point it at cell 0's row instead
    assert bid[c, 1] < 0
    nbr[c, 1] = 4


def _self_neighbour(nbr, bid, V, X):
    nbr[1, 1] = 1


def _periodic_to_interior(nbr, bid, V, X):
    # the x- face of cell 0 is periodic with cell 2's x+ face; make that face interior
    assert bid[0, 0] >= 0 and nbr[0, 0] == 2
    bid[2, 1] = -1


def _rotated_neighbour(nbr, bid, V, X):
    # cell 2 with its local frame turned by 90 degrees about x (new y = old z, new z = 1 - old y):
    # the same right-handed cell, but its x- face vertices come in another order than cell 1's x+
    # face vertices
    V[2] = V[2][ROT_VERTICES]
    n = X.shape[2]
    NP = round(n ** (1 / 3))
    idx = [i + NP * ((NP - 1 - k) + NP * j) for k in range(NP) for j in range(NP) for i in range(NP)]
    X[2] = X[2][:, idx]


@pytest.mark.parametrize("mutate", [_non_reciprocal, _self_neighbour, _periodic_to_interior, _rotated_neighbour],
                         ids=["non_reciprocal", "self_neighbour", "periodic_partner_interior", "rotated_neighbour"])
@pytest.mark.parametrize("with_vertices", [True, False], ids=["vertices", "node_coords"])
def test_bad_connectivity_rejected(mutate, with_vertices):
    m, bcs = box()
    assert create_rc(m, bcs, with_vertices=with_vertices, mutate=mutate) == L.PLENS_E_INVALID


def test_curved_mesh_rejects_misoriented_shared_face():
    """two cells along x sharing the face of vertex ids (1,3,5,7) / (0,2,4,6): accepted in the
    standard order, rejected when the second cell lists its vertices turned by 90 degrees about x"""
    X = np.array([[x, y, z] for z in (0, 1) for y in (0, 1) for x in (1.0, 2.0, 3.0)])
    vid = lambda i, j, k: i + 3 * (j + 2 * k)
    cells = np.array([[vid(i + a, b, c) for c in (0, 1) for b in (0, 1) for a in (0, 1)] for i in (0, 1)], dtype=np.int32)
    nbr = np.full((2, 6), -1, dtype=np.int32)
    bid = np.zeros((2, 6), dtype=np.int32)
    nbr[0, 1], nbr[1, 0] = 1, 0
    bid[0, 1] = bid[1, 0] = -1
    md = L.ManifoldDesc()
    md.type = L.MANIFOLD_CYLINDER
    md.axis_dir[0] = 1.0
    md.axis_point[1], md.axis_point[2] = -5.0, -5.0
    m = cases.Mesh.curved(2, 2, X, cells, nbr, bid, md)
    assert m.n_cells == 2
    bad = cells.copy()
    bad[1] = bad[1][ROT_VERTICES]
    with pytest.raises(RuntimeError):
        cases.Mesh.curved(2, 2, X, bad, nbr, bid, md)
