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
    """Owning wrapper of a host plens_mesh (structured box)."""

    def __init__(self, degree, nx, ny, nz, lo, hi, periodic, side_bid, periodic_bid,
                 y_stretch_ratio=1.0, perturb_frac=0.0, seed=0, transpose_y_face_normals=True):
        lib = L.load()
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
        self.n = (degree + 1) ** 3
        self.nf = (degree + 1) ** 2
        self.n_dof = self.n_cells * self.n
        self.transpose_y_face_normals = bool(transpose_y_face_normals)

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
        d.vertices = self.vertices.ctypes.data_as(dp) if with_vertices else None
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


def make_bc(type_, p=0.0, T=0.0, vel=(0.0, 0.0, 0.0)):
    b = L.BC()
    b.type, b.p, b.T = type_, p, T
    for d in range(3):
        b.vel[d] = vel[d]
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
