"""Workload definitions: the structured meshes, gas models, boundary conditions and seeded initial
states of the measured configurations, built through the host C ABI (include/plens_host.h).

The gas is the reference's 'custom' gas ("Navier-Stokes" subsection, dgsem/plens.cc:185-223) with
molecular weight R0 = 8.314462618, i.e. a nondimensional gas constant R = 1. Blender parameters are
the reference defaults (dgsem/plens.cc:420-515) unless a workload says otherwise.
"""
import ctypes as C
import math
import dataclasses
from dataclasses import dataclass, field

import numpy as np

from . import _lib as L

R0 = 8.314462618
SUTH_RATIO = 110.4 / 273.15  # Sutherland S/T0 of air, used for all nondimensional workloads


class Mesh:
    """Owning wrapper of a host plens_mesh: a structured box (constructor) or a curved mesh of given
    hexahedra on a manifold description (Mesh.curved)."""

    def __init__(self, degree, nx, ny, nz, lo, hi, periodic, side_bid, periodic_bid,
                 y_stretch_ratio=1.0, perturb_frac=0.0, seed=0, transpose_y_face_normals=True, _curved=None):
        lib = L.load()
        if _curved is not None:
            self._h = _curved
            self.degree = degree
            self.n_cells = nx
            self.shape = (nx, 1, 1)
            self.transpose_y_face_normals = bool(transpose_y_face_normals)
            self._load(lib)
            return
        p = L.BoxParams()
        p.degree, p.nx, p.ny, p.nz = degree, nx, ny, nz
        for d in range(3):
            p.lo[d], p.hi[d] = lo[d], hi[d]
            p.periodic[d] = int(periodic[d])
            p.periodic_bid[d] = periodic_bid[d]
        for f in range(6):
            p.side_bid[f] = side_bid[f]
        p.y_stretch_ratio = y_stretch_ratio
        p.perturb_frac = perturb_frac
        p.seed = seed
        p.transpose_y_face_normals = int(transpose_y_face_normals)
        self.params = p
        self._h = C.c_void_p()
        L.check(lib.plens_mesh_build_box(C.byref(p), C.byref(self._h)), "plens_mesh_build_box")
        self.degree = degree
        self.n_cells = nx * ny * nz
        self.shape = (nx, ny, nz)
        self.transpose_y_face_normals = bool(transpose_y_face_normals)
        self._load(lib)

    @classmethod
    def curved(cls, degree, mapping_degree, vertices, cells, face_nbr, face_bid, manifold,
               transpose_y_face_normals=True):
        """Curved mesh (plens_mesh_build_curved): vertices [n_v][3], cells [n_c][8] global vertex
        ids in deal.II order, face_nbr / face_bid [n_c][6], manifold an L.ManifoldDesc."""
        lib = L.load()
        X = np.ascontiguousarray(vertices, dtype=np.float64)
        cv = np.ascontiguousarray(cells, dtype=np.int32)
        nb = np.ascontiguousarray(face_nbr, dtype=np.int32)
        bd = np.ascontiguousarray(face_bid, dtype=np.int32)
        h = C.c_void_p()
        ip = C.POINTER(C.c_int32)
        L.check(lib.plens_mesh_build_curved(degree, mapping_degree, cv.shape[0], X.shape[0],
                                            X.ctypes.data_as(C.POINTER(C.c_double)), cv.ctypes.data_as(ip),
                                            nb.ctypes.data_as(ip), bd.ctypes.data_as(ip), C.byref(manifold),
                                            int(transpose_y_face_normals), C.byref(h)), "plens_mesh_build_curved")
        return cls(degree, cv.shape[0], 1, 1, None, None, None, None, None,
                   transpose_y_face_normals=transpose_y_face_normals, _curved=h)

    @classmethod
    def from_cells(cls, degree, cell_vertices, face_nbr, face_bid, transpose_y_face_normals=True):
        """Trilinear (MappingQ1) mesh of the given cells [n_c][8][3] (deal.II vertex order, each cell
        its own vertices) and connectivity: plens_mesh_build_curved at mapping degree 1, whose
        support points are the vertices, so the manifold plays no part."""
        V = np.ascontiguousarray(cell_vertices, dtype=np.float64).reshape(-1, 8, 3)
        md = L.ManifoldDesc()
        md.type = L.MANIFOLD_CYLINDER
        md.axis_dir[0] = 1.0
        cells = np.arange(V.shape[0] * 8, dtype=np.int32).reshape(-1, 8)
        return cls.curved(degree, 1, V.reshape(-1, 3), cells, face_nbr, face_bid, md,
                          transpose_y_face_normals=transpose_y_face_normals)

    def _load(self, lib):
        degree = self.degree
        self.n = (degree + 1) ** 3
        self.nf = (degree + 1) ** 2
        self.n_dof = self.n_cells * self.n
        self.mapping_degree = int(lib.plens_mesh_mapping_degree(self._h))

        def arr(fn, shape, dtype=np.float64):
            ptr = fn(self._h)
            return np.ctypeslib.as_array(ptr, shape=(int(np.prod(shape)),)).reshape(shape).copy()

        nc, n, nf = self.n_cells, self.n, self.nf
        self.vertices = arr(lib.plens_mesh_vertices, (nc, 8, 3))
        self.node_coords = arr(lib.plens_mesh_node_coords, (nc, 3, n))
        self.Ja = arr(lib.plens_mesh_Ja, (nc, 9, n))
        self.detJ = arr(lib.plens_mesh_detJ, (nc, n))
        self.face_normals = arr(lib.plens_mesh_face_normals, (nc, 6, 3, nf))
        self.face_nbr = arr(lib.plens_mesh_face_nbr, (nc, 6), np.int32)
        self.face_bid = arr(lib.plens_mesh_face_bid, (nc, 6), np.int32)
        self.min_vertex_dist = arr(lib.plens_mesh_min_vertex_dist, (nc,))
        qp = self.mapping_degree + 1
        self.support_points = arr(lib.plens_mesh_support_points, (nc, qp ** 3, 3))

    @property
    def handle(self):
        return self._h

    def desc(self, with_vertices=True):
        """plens_mesh_desc pointing into this object's (contiguous) numpy arrays; with_vertices
        hands over the cell vertices too, so the kernels derive the metric terms from the trilinear
        maps instead of reading the stored Ja and detJ."""
        d = L.MeshDesc()
        d.n_cells, d.degree = self.n_cells, self.degree
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int32)
        d.Ja = self.Ja.ctypes.data_as(dp)
        d.detJ = self.detJ.ctypes.data_as(dp)
        d.face_normals = self.face_normals.ctypes.data_as(dp)
        d.min_vertex_dist = self.min_vertex_dist.ctypes.data_as(dp)
        d.face_nbr = self.face_nbr.ctypes.data_as(ip)
        d.face_bid = self.face_bid.ctypes.data_as(ip)
        # a curved (q > 1) map is not trilinear: its metric terms are always handed over stored
        d.vertices = self.vertices.ctypes.data_as(dp) if (with_vertices and self.mapping_degree == 1) else None
        d.node_coords = self.node_coords.ctypes.data_as(dp)
        return d

    def close(self):
        if self._h:
            L.load().plens_mesh_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_physics(gamma=1.4, prandtl=0.72, mu0=1e-3, T0=1.0, S=SUTH_RATIO, flux="HLLC",
                 precision=L.PREC_FP64, wall_blender_limit=0.5, alpha_max=0.5, alpha_min=1e-3,
                 threshold_factor=0.5, threshold_exp_factor=1.8, sharpness_factor=9.21024,
                 gradient_variables=L.GRAD_CONSERVATIVE, vol_flux="Chandrashekhar"):
    ph = L.Physics()
    ph.gamma, ph.molecular_weight, ph.prandtl = gamma, R0, prandtl
    ph.mu0, ph.T0, ph.S = mu0, T0, S
    ph.inv_surf_flux = L.FLUX[flux] if isinstance(flux, str) else int(flux)
    ph.precision = precision
    ph.threshold_factor, ph.threshold_exp_factor = threshold_factor, threshold_exp_factor
    ph.sharpness_factor, ph.alpha_min, ph.alpha_max = sharpness_factor, alpha_min, alpha_max
    ph.wall_blender_limit = wall_blender_limit
    ph.gradient_variables = gradient_variables
    ph.inv_vol_flux = L.VOL_FLUX[vol_flux] if isinstance(vol_flux, str) else int(vol_flux)
    return ph


def make_bc(type_, p=0.0, T=0.0, vel=(0.0, 0.0, 0.0), p_expr=None, T_expr=None, vel_expr=None):
    """plens_bc; a varying inflow takes the prm strings "prescribed p", "prescribed T" and
    "prescribed velocity" (components separated by ';')"""
    b = L.BC()
    b.type, b.p, b.T = type_, p, T
    for d in range(3):
        b.vel[d] = vel[d]
    for name, v in (("p_expr", p_expr), ("T_expr", T_expr), ("vel_expr", vel_expr)):
        setattr(b, name, v.encode() if v is not None else None)
    return b


@dataclass
class Case:
    name: str
    mesh: Mesh
    physics: L.Physics
    bcs: list
    Q0: np.ndarray
    courant: float
    n_steps: int
    meta: dict = field(default_factory=dict)

    def bc_array(self):
        arr = (L.BC * max(len(self.bcs), 1))()
        for i, b in enumerate(self.bcs):
            arr[i] = b
        return arr


def _ic(fn, mesh, *args):
    Q = np.zeros((5, mesh.n_dof))
    L.check(fn(mesh.handle, *args, Q.ctypes.data_as(C.POINTER(C.c_double))), fn.__name__)
    return Q


def periodic_cube(degree=3, n=8, seed=0, flux="Chandrashekhar", mu0=1e-3, steps=100, velocity=(0.5, 0.3, 0.2)):
    """Periodic unit cube, density wave rho = 1 + 0.1 sin(2 pi k.x), p = 1, u = (0.5, 0.3, 0.2),
    seeded integer k with 0 < |k| <= 2; CFL 0.4. The blender stays at its floor."""
    lib = L.load()
    mesh = Mesh(degree, n, n, n, (0, 0, 0), (1, 1, 1), (1, 1, 1), [-1] * 6, (0, 0, 0))
    ph = make_physics(mu0=mu0, flux=flux)
    vel = (C.c_double * 3)(*velocity)
    kk = (C.c_int32 * 3)()
    Q = np.zeros((5, mesh.n_dof))
    L.check(lib.plens_ic_density_wave(mesh.handle, 1.4, 1.0, vel, 1.0, 0.1, 2.0, seed, kk,
                                      Q.ctypes.data_as(C.POINTER(C.c_double))), "ic")
    return Case("periodic_cube", mesh, ph, [make_bc(L.BC_PERIODIC)], Q, 0.4, steps,
                {"k": list(kk), "seed": seed, "cells": (n, n, n), "degree": degree})


def shock_tube(degree=3, nx=128, ny=8, nz=8, seed=1, flux="HLLC", mach=5.0, reynolds=1e4, steps=200):
    """Stationary Mach-5 normal shock in [0,4]x[0,1/4]^2 (Rankine-Hugoniot, pre-shock rho=p=1),
    seeded 1% nodal density noise downstream; uniform inflow at x=0, outflow at x=4 with the
    post-shock pressure, periodic in y and z, Re=1e4 per unit length."""
    lib = L.load()
    g = 1.4
    a1 = math.sqrt(g)
    u1 = mach * a1
    p2 = 1.0 + 2.0 * g / (g + 1.0) * (mach * mach - 1.0)
    mesh = Mesh(degree, nx, ny, nz, (0, 0, 0), (4.0, 0.25, 0.25), (0, 1, 1),
                [0, 1, -1, -1, -1, -1], (-1, 2, 2))
    ph = make_physics(mu0=u1 / reynolds, T0=1.0, flux=flux)
    bcs = [make_bc(L.BC_UNIFORM_INFLOW, p=1.0, T=1.0, vel=(u1, 0, 0)), make_bc(L.BC_OUTFLOW, p=p2),
           make_bc(L.BC_PERIODIC)]
    dx = 4.0 / nx
    x_shock = 1.0 + 0.5 * dx
    Q = np.zeros((5, mesh.n_dof))
    L.check(lib.plens_ic_normal_shock(mesh.handle, g, 1.0, 1.0, mach, x_shock, 0.01, seed,
                                      Q.ctypes.data_as(C.POINTER(C.c_double))), "ic")
    return Case("shock_tube", mesh, ph, bcs, Q, 0.4, steps,
                {"seed": seed, "cells": (nx, ny, nz), "degree": degree, "x_shock": x_shock, "p2": p2})


def wall_flow(degree=4, nx=128, ny=64, nz=2, seed=2, flux="HLLC", mach=6.0, T_inf=100.0,
              reynolds=1e5, steps=500, perturb=0.05, stretch=1.05):
    """Curvilinear viscous wall flow in [0,1]x[0,0.2]x[0,0.02]: y-stretched (ratio 1.05) mesh with
    seeded smooth 5% vertex displacements; Mach-6 uniform inflow at T=100 (rho=1, R=1), Re=1e5 per
    length; isothermal wall T_w = 3 T_inf at y=0; outflow at x=1 and y=0.2; periodic in z;
    wall blender limit 0."""
    g = 1.4
    p_inf = T_inf  # rho = 1, R = 1
    u_inf = mach * math.sqrt(g * p_inf)
    mesh = Mesh(degree, nx, ny, nz, (0, 0, 0), (1.0, 0.2, 0.02), (0, 0, 1),
                [0, 1, 2, 1, -1, -1], (-1, -1, 3), y_stretch_ratio=stretch, perturb_frac=perturb, seed=seed)
    ph = make_physics(mu0=u_inf / reynolds, T0=T_inf, S=SUTH_RATIO * T_inf, flux=flux, wall_blender_limit=0.0)
    bcs = [make_bc(L.BC_UNIFORM_INFLOW, p=p_inf, T=T_inf, vel=(u_inf, 0, 0)), make_bc(L.BC_OUTFLOW, p=p_inf),
           make_bc(L.BC_UNIFORM_TEMP_WALL, T=3 * T_inf, vel=(0, 0, 0)), make_bc(L.BC_PERIODIC)]
    vel = (C.c_double * 3)(u_inf, 0.0, 0.0)
    Q = np.zeros((5, mesh.n_dof))
    L.check(L.load().plens_ic_uniform(mesh.handle, g, 1.0, vel, p_inf, Q.ctypes.data_as(C.POINTER(C.c_double))),
            "ic")
    return Case("wall_flow", mesh, ph, bcs, Q, 0.4, steps,
                {"seed": seed, "cells": (nx, ny, nz), "degree": degree, "u_inf": u_inf})


def _structured_cells(X, side_bid):
    """Cells, neighbours and boundary indices of a structured lattice of vertex positions
    X[k][j][i][3] (i fastest): cell (i, j, k) -> index i + n0 (j + n1 k), reference directions
    (xi_0, xi_1, xi_2) along (i, j, k); side_bid[f] for the six sides."""
    n2, n1, n0 = X.shape[0] - 1, X.shape[1] - 1, X.shape[2] - 1
    vid = lambda i, j, k: i + (n0 + 1) * (j + (n1 + 1) * k)
    nc = n0 * n1 * n2
    cells = np.zeros((nc, 8), dtype=np.int32)
    nbr = -np.ones((nc, 6), dtype=np.int32)
    bid = -np.ones((nc, 6), dtype=np.int32)
    dims = (n0, n1, n2)
    for k in range(n2):
        for j in range(n1):
            for i in range(n0):
                c = i + n0 * (j + n1 * k)
                for v in range(8):
                    cells[c, v] = vid(i + (v & 1), j + ((v >> 1) & 1), k + ((v >> 2) & 1))
                for f in range(6):
                    I = [i, j, k]
                    d, side = f // 2, f % 2
                    I[d] += 1 if side else -1
                    if 0 <= I[d] < dims[d]:
                        nbr[c, f] = I[0] + n0 * (I[1] + n1 * I[2])
                    else:
                        bid[c, f] = side_bid[f]
    return X.reshape(-1, 3), cells, nbr, bid


def _smooth_state(mesh, g, rho0, vel, p0, amp, wave):
    """rho = rho0 (1 + amp sin(wave . x)), uniform velocity and pressure, at the mesh's nodes."""
    x = mesh.node_coords  # [nc][3][n]
    rho = rho0 * (1.0 + amp * np.sin(np.einsum("d,cdn->cn", np.asarray(wave, dtype=np.float64), x))).reshape(-1)
    v = np.asarray(vel, dtype=np.float64)
    Q = np.empty((5, mesh.n_dof))
    Q[0] = rho
    for d in range(3):
        Q[1 + d] = rho * v[d]
    Q[4] = p0 / (g - 1) + 0.5 * rho * (v @ v)
    return Q


def cylinder_flare(degree=3, mapping_degree=2, n_axial=12, n_normal=4, n_azim=2, mach=3.0, reynolds=1e4,
                   radius=1.0, thickness=0.3, length=2.0, flare_start=1.0, flare_deg=15.0, wedge_deg=10.0,
                   flux="HLLC", wall=True, steps=20, transpose_y_face_normals=True, perturb=0.0, seed=3):
    """Curved mesh on the "cylinder" manifold description (axis x through the origin): an
    azimuthal wedge of the layer over a cylinder of the given radius that turns into a conical
    flare at x = flare_start; layer thickness `thickness` in r, wedge angle `wedge_deg`, MappingQ of
    degree `mapping_degree`. Mach-`mach` uniform inflow (rho = 1, p = 1 / gamma) at x = 0, outflow
    at x = length and at the layer's outer surface, isothermal wall (T = 1.5 T_inf; `wall=False`:
    symmetry) on the body, symmetry on both wedge planes; the initial state adds a 5% smooth density
    wave to the free stream. (The wall ghost state 2 u_wall - u has density 2 p / (R T_w) - rho,
    which vanishes for the free stream at T_w = 2 T_inf, and the stage-3 viscous flux divides by it.)
    `perturb` > 0 moves the interior vertices by seeded smooth displacements of that fraction of the
    local spacing in r, phi and x."""
    g = 1.4
    p_inf, T_inf = 1.0 / g, 1.0 / g
    u_inf = mach
    xs = np.linspace(0.0, length, n_axial + 1)
    rb = radius + np.maximum(xs - flare_start, 0.0) * math.tan(math.radians(flare_deg))
    X = np.zeros((n_azim + 1, n_normal + 1, n_axial + 1, 3))
    rng = np.random.default_rng(seed)
    kw = rng.integers(1, 4, size=(3, 3))
    phase = rng.uniform(0, 2 * math.pi, 3)
    for k in range(n_azim + 1):
        for j in range(n_normal + 1):
            for i in range(n_axial + 1):
                r = rb[i] + thickness * j / n_normal
                ph = math.radians(wedge_deg) * k / n_azim
                x = xs[i]
                inner = 0 < i < n_axial and 0 < j < n_normal and 0 < k < n_azim
                if perturb > 0 and inner:
                    arg = 2 * math.pi * (kw @ np.array([i / n_axial, j / n_normal, k / n_azim])) + phase
                    r += perturb * thickness / n_normal * math.sin(arg[0])
                    ph += perturb * math.radians(wedge_deg) / n_azim * math.sin(arg[1])
                    x += perturb * length / n_axial * math.sin(arg[2])
                X[k, j, i] = (x, r * math.cos(ph), r * math.sin(ph))
    V, cells, nbr, bid = _structured_cells(X, [0, 1, 2, 1, 3, 3])
    md = L.ManifoldDesc()
    md.type = L.MANIFOLD_CYLINDER
    md.axis_dir[:] = [1.0, 0.0, 0.0]
    md.axis_point[:] = [0.0, 0.0, 0.0]
    mesh = Mesh.curved(degree, mapping_degree, V, cells, nbr, bid, md,
                       transpose_y_face_normals=transpose_y_face_normals)
    ph = make_physics(mu0=u_inf / reynolds, T0=T_inf, S=SUTH_RATIO * T_inf, flux=flux, wall_blender_limit=0.0)
    body = make_bc(L.BC_UNIFORM_TEMP_WALL, T=1.5 * T_inf, vel=(0, 0, 0)) if wall else make_bc(L.BC_SYMMETRY)
    bcs = [make_bc(L.BC_UNIFORM_INFLOW, p=p_inf, T=T_inf, vel=(u_inf, 0, 0)), make_bc(L.BC_OUTFLOW, p=p_inf),
           body, make_bc(L.BC_SYMMETRY)]
    Q = _smooth_state(mesh, g, 1.0, (u_inf, 0.0, 0.0), p_inf, 0.05, (2 * math.pi / length, 0.0, 0.0))
    return Case("cylinder_flare", mesh, ph, bcs, Q, 0.4, steps,
                {"degree": degree, "mapping_degree": mapping_degree, "cells": (n_axial, n_normal, n_azim)})


def nose_cone(degree=3, mapping_degree=2, n_sphere=4, n_cone=6, n_normal=3, n_azim=2, mach=3.0, reynolds=1e4,
              radius=1.0, thickness=0.25, cone_length=1.5, cone_deg=20.0, theta0_deg=30.0, wedge_deg=10.0,
              flux="HLLC", wall=True, steps=20, transpose_y_face_normals=True):
    """Curved mesh on the "nose cylinder" manifold description: an azimuthal wedge of the layer over
    a blunted cone (spherical nose of the given radius about the origin, axis pointing to -x out of
    the nose, cone of half-angle `cone_deg` tangent to the sphere, separation plane through the
    tangency circle). Along the body the mesh runs from polar angle `theta0_deg` (the tip is left
    out) over the sphere (n_sphere cells), then `cone_length` along the cone (n_cone cells); spherical
    manifold on the nose cells, cylindrical on the cone cells. Free stream Mach `mach` along +x
    (rho = 1, p = 1 / gamma) prescribed on the outer surface and at the cut near the tip, outflow at
    the cone's end, isothermal wall at 1.5 T_inf (symmetry with `wall=False`) on the body, symmetry
    on both wedge planes; a 5% smooth density wave is added."""
    g = 1.4
    p_inf, T_inf = 1.0 / g, 1.0 / g
    u_inf = mach
    a = np.array([-1.0, 0.0, 0.0])
    tt = math.radians(90.0 - cone_deg)
    th = np.linspace(math.radians(theta0_deg), tt, n_sphere + 1)
    ls = np.linspace(0.0, cone_length, n_cone + 1)[1:]
    tcone = -math.cos(math.radians(cone_deg)) * a
    n0 = n_sphere + n_cone
    X = np.zeros((n_azim + 1, n_normal + 1, n0 + 1, 3))
    for k in range(n_azim + 1):
        ph = math.radians(wedge_deg) * k / n_azim
        er = np.array([0.0, math.cos(ph), math.sin(ph)])
        t_dir = tcone + math.sin(math.radians(cone_deg)) * er
        for j in range(n_normal + 1):
            rr = radius + thickness * j / n_normal
            for i, t in enumerate(th):
                X[k, j, i] = rr * (math.cos(t) * a + math.sin(t) * er)
            base = rr * (math.cos(tt) * a + math.sin(tt) * er)
            for i, l in enumerate(ls):
                X[k, j, n_sphere + 1 + i] = base + l * t_dir
    V, cells, nbr, bid = _structured_cells(X, [0, 1, 2, 0, 3, 3])
    md = L.ManifoldDesc()
    md.type = L.MANIFOLD_NOSE_CYLINDER
    md.axis_dir[:] = list(a)
    md.axis_point[:] = list(radius * math.cos(tt) * a)
    md.nose_center[:] = [0.0, 0.0, 0.0]
    # the separation plane must put exactly the nose cells on the spherical side
    ctr = V[cells].mean(axis=1)
    sph = (ctr - np.array(md.axis_point[:])) @ a > 0
    expect = (np.arange(len(cells)) % n0) < n_sphere
    if not np.array_equal(sph, expect):
        raise ValueError("nose_cone: cone too short or layer too thick for the separation plane")
    mesh = Mesh.curved(degree, mapping_degree, V, cells, nbr, bid, md,
                       transpose_y_face_normals=transpose_y_face_normals)
    ph = make_physics(mu0=u_inf / reynolds, T0=T_inf, S=SUTH_RATIO * T_inf, flux=flux, wall_blender_limit=0.0)
    body = make_bc(L.BC_UNIFORM_TEMP_WALL, T=1.5 * T_inf, vel=(0, 0, 0)) if wall else make_bc(L.BC_SYMMETRY)
    bcs = [make_bc(L.BC_UNIFORM_INFLOW, p=p_inf, T=T_inf, vel=(u_inf, 0, 0)), make_bc(L.BC_OUTFLOW, p=p_inf),
           body, make_bc(L.BC_SYMMETRY)]
    Q = _smooth_state(mesh, g, 1.0, (u_inf, 0.0, 0.0), p_inf, 0.05, (2.0, 1.0, 0.5))
    return Case("nose_cone", mesh, ph, bcs, Q, 0.4, steps,
                {"degree": degree, "mapping_degree": mapping_degree, "cells": (n0, n_normal, n_azim)})


def varying_inflow_tube(degree=3, nx=12, ny=2, nz=2, perturb=0.1, seed=4, flux="HLLC", steps=20,
                        p_expr="1 + 0.1*sin(6.283185307179586*y)*cos(3*t)",
                        T_expr="(x < 1e-12 && z > 0.1) ? 1.05 : 1 + 0.02*t",
                        vel_expr="2*(1 + 0.05*cos(6.283185307179586*(y + z)) + 0.5*t^2); 0.1*sin(t); _pi/100"):
    """Box [0,1]x[0,1/4]^2 (perturbed vertices) with a varying inflow at x = 0 whose p, T and
    velocity depend on y, z and t (the defaults exercise functions, the ternary, the logical and
    power operators and a constant), outflow at x = 1, periodic in y and z; the initial state is
    the uniform flow rho = 1, p = 1, u = 2 (Re 1e3 per length)."""
    g = 1.4
    mesh = Mesh(degree, nx, ny, nz, (0, 0, 0), (1.0, 0.25, 0.25), (0, 1, 1), [0, 1, -1, -1, -1, -1], (-1, 2, 2),
                perturb_frac=perturb, seed=seed)
    ph = make_physics(mu0=2.0 / 1e3, T0=1.0, S=SUTH_RATIO, flux=flux)
    bcs = [make_bc(L.BC_VARYING_INFLOW, p_expr=p_expr, T_expr=T_expr, vel_expr=vel_expr),
           make_bc(L.BC_OUTFLOW, p=1.0), make_bc(L.BC_PERIODIC)]
    Q = _smooth_state(mesh, g, 1.0, (2.0, 0.0, 0.0), 1.0, 0.0, (0.0, 0.0, 0.0))
    return Case("varying_inflow_tube", mesh, ph, bcs, Q, 0.4, steps, {"degree": degree, "cells": (nx, ny, nz)})


def batch_instance(i, seed0=100, min_n=4, max_n=32, flux="HLLC"):
    """Instance i (seed seed0 + i) of the batched uneven-solve workload: a periodic density-wave box
    (as periodic_cube) or a normal-shock tube (as shock_tube), drawn with even probability; about
    n^3 cells with n log-uniform in [min_n, max_n] and even (box n x n x n; tube n^3/m^2 x m x m with
    m = max(1, n / 4)); degree uniform in {2, 3, 4, 5}; Reynolds number
    log-uniform in [1e3, 1e5]; Mach uniform in [0.3, 8] (box flow speed) or [1.5, 8] (tube shock).
    The step count of every instance is 50; the Courant number is the instance's own (below)."""
    rng = np.random.default_rng(seed0 + i)
    kind = "box" if rng.random() < 0.5 else "tube"
    n = int(np.exp(rng.uniform(np.log(min_n), np.log(max_n))) + 0.5)
    n = max(min_n, min(max_n, n + (n & 1)))
    degree = int(rng.integers(2, 6))
    reynolds = float(10.0 ** rng.uniform(3.0, 5.0))
    if kind == "box":
        mach = float(rng.uniform(0.3, 8.0))
        d = np.array([0.5, 0.3, 0.2]) / np.linalg.norm([0.5, 0.3, 0.2])
        u = d * mach * math.sqrt(1.4)  # p = rho = 1
        case = periodic_cube(degree=degree, n=n, seed=seed0 + i, flux=flux, mu0=float(np.linalg.norm(u)) / reynolds,
                             steps=50, velocity=tuple(u))
    else:
        mach = float(rng.uniform(1.5, 8.0))
        m = max(1, n // 4)
        case = shock_tube(degree=degree, nx=max(1, n ** 3 // (m * m)), ny=m, nz=m, seed=seed0 + i, flux=flux,
                          mach=mach, reynolds=reynolds, steps=50)
    # Courant number of the instance: 0.4, halved for shocks above Mach 6, where the blended scheme
    # loses positivity at 0.4 within the first steps (the oracle fails the same way)
    case.courant = 0.2 if (kind == "tube" and mach > 6.0) else 0.4
    case.meta.update({"instance": i, "kind": kind, "mach": mach, "reynolds": reynolds})
    return case


def with_gradient_variables(case, variables):
    """The same case with the BR1 lifting in "conservative" (CVAR_GRAD build) or "entropy"
    variables (the reference's default build)."""
    codes = {"conservative": L.GRAD_CONSERVATIVE, "entropy": L.GRAD_ENTROPY}
    ph = L.Physics.from_buffer_copy(case.physics)
    ph.gradient_variables = codes[variables]
    return dataclasses.replace(case, physics=ph, meta={**case.meta, "gradient_variables": variables})


def with_fluxes(case, surface=None, volume=None):
    """The same case with another inviscid surface flux and/or volume (two-point) flux scheme,
    named as the reference's "inviscid surface/volume flux scheme" parameters."""
    ph = L.Physics.from_buffer_copy(case.physics)
    meta = dict(case.meta)
    if surface is not None:
        ph.inv_surf_flux = L.FLUX[surface]
        meta["flux"] = surface
    if volume is not None:
        ph.inv_vol_flux = L.VOL_FLUX[volume]
        meta["vol_flux"] = volume
    return dataclasses.replace(case, physics=ph, meta=meta)


def with_precision(case, precision):
    """The same case run in another arithmetic variant: "fp64" (plain) or "fp32_bulk" (FP32 fluxes,
    avars and face terms; FP64 states, residual registers, residual accumulation, metrics, blender,
    time step and the near-equal log-mean branch)."""
    codes = {"fp64": L.PREC_FP64, "fp32_bulk": L.PREC_FP32_BULK}
    ph = L.Physics.from_buffer_copy(case.physics)
    ph.precision = codes[precision] if isinstance(precision, str) else int(precision)
    return dataclasses.replace(case, physics=ph, meta={**case.meta, "precision": precision})


CONFIGS = {"periodic_cube": periodic_cube, "shock_tube": shock_tube, "wall_flow": wall_flow}
