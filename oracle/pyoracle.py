"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (oracle/liboracle.so).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; the
product path never imports it.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
# the flop-counting build of the same source (single-threaded; oracle_flop_counts reads the counters)
FLOPS_LIB_PATH = os.path.join(HERE, "liboracle_flops.so")
FLOP_BUCKETS = ["positivity", "aux vars", "blender", "surface fluxes", "ho inv residual (volume, outside the pair fluxes)",
                "ho dif residual (volume, outside the viscous fluxes)", "lo inv residual", "residual blend",
                "time step", "RK update", "two-point volume fluxes", "dense change of basis",
                "stage-2 numerical fluxes", "viscous fluxes of the ho viscous volume term",
                "ho inv residual (surface part)", "ho dif residual (surface part)"]
_dp = C.POINTER(C.c_double)


class OPhys(C.Structure):
    _fields_ = [("gamma", C.c_double), ("molecular_weight", C.c_double), ("prandtl", C.c_double),
                ("mu0", C.c_double), ("T0", C.c_double), ("S", C.c_double), ("inv_surf_flux", C.c_int32),
                ("threshold_factor", C.c_double), ("threshold_exp_factor", C.c_double),
                ("sharpness_factor", C.c_double), ("alpha_min", C.c_double), ("alpha_max", C.c_double),
                ("wall_blender_limit", C.c_double), ("gradient_variables", C.c_int32),
                ("inv_vol_flux", C.c_int32)]


class OBC(C.Structure):
    _fields_ = [("type", C.c_int32), ("p", C.c_double), ("T", C.c_double), ("vel", C.c_double * 3),
                ("p_expr", C.c_char_p), ("T_expr", C.c_char_p), ("vel_expr", C.c_char_p)]


_libs = {}


def load(path=LIB_PATH):
    if path in _libs:
        return _libs[path]
    if not os.path.exists(path):
        from . import build
        build.build()
    lib = C.CDLL(path)
    vp = C.c_void_p
    ip = C.POINTER(C.c_int32)
    sig = {
        "oracle_create": (C.c_int, [C.c_int32, C.c_int32, _dp, ip, ip, C.POINTER(OPhys), C.POINTER(OBC), C.c_int32,
                                    C.c_int32, C.POINTER(vp)]),
        "oracle_create_mapped": (C.c_int, [C.c_int32, C.c_int32, C.c_int32, _dp, ip, ip, C.POINTER(OPhys),
                                           C.POINTER(OBC), C.c_int32, C.c_int32, C.POINTER(vp)]),
        "oracle_destroy": (None, [vp]),
        "oracle_output_fields": (C.c_int, [vp, C.c_double, C.c_double, _dp, _dp]),
        "oracle_set_time": (C.c_int, [vp, C.c_double]),
        "oracle_set_alpha": (C.c_int, [vp, _dp, _dp]),
        "oracle_set_raw_alpha": (C.c_int, [vp, _dp]),
        "oracle_set_mutation": (C.c_int, [C.c_int32]),
        "oracle_blended_cell_stages": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "oracle_set_expressions": (C.c_int, [vp, C.c_char_p, C.c_char_p, C.c_char_p]),
        "oracle_bc_ghost": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _dp, _dp, _dp, _dp, _dp]),
        "oracle_node_location": (C.c_int, [vp, C.c_int32, C.c_int32, _dp]),
        "oracle_eval_expression": (C.c_double, [C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_double]),
        "oracle_set_threads": (C.c_int, [C.c_int32]),
        "oracle_set_state": (C.c_int, [vp, _dp]),
        "oracle_get_state": (C.c_int, [vp, _dp]),
        "oracle_calc_rhs": (C.c_int, [vp, _dp, _dp, _dp, _dp]),
        "oracle_calc_dt": (C.c_int, [vp, C.c_double, _dp]),
        "oracle_rk4_step": (C.c_int, [vp, C.c_double, _dp]),
        "oracle_run": (C.c_int, [vp, C.c_int32, C.c_double, _dp]),
        "oracle_rk4_step_fixed_dt": (C.c_int, [vp, C.c_double, _dp]),
        "oracle_last_failure": (C.c_int, [vp, C.POINTER(C.c_int64), ip]),
        "oracle_section_times": (C.c_int, [vp, _dp, C.c_int32]),
        "oracle_flop_counts": (C.c_int, [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_int32, C.c_int32]),
        "oracle_cell_metrics": (C.c_int, [vp, C.c_int32, _dp, _dp, _dp]),
        "oracle_cell_subcell_normals": (C.c_int, [vp, C.c_int32, _dp]),
        "oracle_basis": (C.c_int, [C.c_int32, _dp, _dp, _dp]),
        "oracle_cob_matrix": (C.c_int, [C.c_int32, _dp]),
        "oracle_log_avg": (C.c_double, [C.c_double, C.c_double]),
        "oracle_inv_surf_flux": (None, [C.POINTER(OPhys), C.c_int32, C.c_double, _dp, _dp, _dp, _dp]),
        "oracle_inv_flux": (None, [C.POINTER(OPhys), _dp, _dp, _dp]),
        "oracle_vol_flux": (None, [C.POINTER(OPhys), _dp, _dp, _dp, _dp]),
        "oracle_dif_flux": (None, [_dp, _dp, _dp, _dp]),
        "oracle_blender_value": (C.c_double, [C.POINTER(OPhys), C.c_int32, _dp]),
        "oracle_rk4_coeffs": (None, [_dp, _dp, _dp]),
        "oracle_set_local_stepping": (C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_double, C.c_double]),
        "oracle_loc_dt": (C.c_int, [C.c_void_p, _dp]),
        "oracle_ss_error": (C.c_int, [C.c_void_p, _dp]),
    }
    for k, (r, a) in sig.items():
        f = getattr(lib, k)
        f.restype, f.argtypes = r, a
    _libs[path] = lib
    return lib


def p(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(_dp)


class mutation:
    """context manager planting oracle mistake `id` (oracle_set_mutation) for the duration of a block"""

    def __init__(self, mid):
        self.mid = mid

    def __enter__(self):
        load().oracle_set_mutation(self.mid)
        return self

    def __exit__(self, *exc):
        load().oracle_set_mutation(0)
        return False


def ophys_from(ph):
    """OPhys from a plens_amd Physics struct (same fields minus precision)."""
    o = OPhys()
    for name, _ in OPhys._fields_:
        setattr(o, name, getattr(ph, name))
    return o


class OraclePlens:
    """CPU restatement of PLENS::calc_rhs / calc_time_step / update (RK4)."""

    def __init__(self, case=None, mesh=None, physics=None, bcs=None, threads=0, lib_path=LIB_PATH):
        if case is not None:
            mesh, physics, bcs = case.mesh, case.physics, case.bcs
        self.lib = load(lib_path)
        if threads:
            self.lib.oracle_set_threads(threads)
        self.n_cells, self.n_dof, self.n = mesh.n_cells, mesh.n_dof, mesh.n
        self._ph = ophys_from(physics)
        arr = (OBC * max(len(bcs), 1))()
        for i, b in enumerate(bcs):
            arr[i].type, arr[i].p, arr[i].T = b.type, b.p, b.T
            arr[i].p_expr, arr[i].T_expr, arr[i].vel_expr = b.p_expr, b.T_expr, b.vel_expr
            for d in range(3):
                arr[i].vel[d] = b.vel[d]
        self._bcs = arr
        self._V = np.ascontiguousarray(mesh.vertices)
        self._nbr = np.ascontiguousarray(mesh.face_nbr, dtype=np.int32)
        self._bid = np.ascontiguousarray(mesh.face_bid, dtype=np.int32)
        self._h = C.c_void_p()
        ip = C.POINTER(C.c_int32)
        q = getattr(mesh, "mapping_degree", 1)
        if q > 1:
            self._S = np.ascontiguousarray(mesh.support_points, dtype=np.float64)
            rc = self.lib.oracle_create_mapped(mesh.n_cells, mesh.degree, q, self._S.ctypes.data_as(_dp),
                                               self._nbr.ctypes.data_as(ip), self._bid.ctypes.data_as(ip),
                                               C.byref(self._ph), arr, len(bcs), int(mesh.transpose_y_face_normals),
                                               C.byref(self._h))
        else:
            rc = self.lib.oracle_create(mesh.n_cells, mesh.degree, self._V.ctypes.data_as(_dp),
                                        self._nbr.ctypes.data_as(ip), self._bid.ctypes.data_as(ip),
                                        C.byref(self._ph), arr, len(bcs), int(mesh.transpose_y_face_normals),
                                        C.byref(self._h))
        if rc:
            raise RuntimeError(f"oracle_create failed {rc}")
        if case is not None:
            self.set_state(case.Q0)

    @staticmethod
    def set_threads(n):
        """OpenMP threads of the oracle (process-wide); returns the count in effect"""
        return load().oracle_set_threads(n)

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64)
        self.lib.oracle_set_state(self._h, Q.ctypes.data_as(_dp))

    def get_state(self):
        Q = np.empty((5, self.n_dof))
        self.lib.oracle_get_state(self._h, Q.ctypes.data_as(_dp))
        return Q

    def _fail(self, rc):
        dof, kind = C.c_int64(), C.c_int32()
        self.lib.oracle_last_failure(self._h, C.byref(dof), C.byref(kind))
        raise RuntimeError(f"oracle positivity failure rc={rc} dof={dof.value} kind={kind.value}")

    def calc_rhs(self, want_avars=False):
        rhs = np.empty((5, self.n_dof))
        a = np.empty(self.n_cells)
        ad = np.empty(self.n_cells)
        av = np.empty((9, self.n_dof)) if want_avars else None
        rc = self.lib.oracle_calc_rhs(self._h, rhs.ctypes.data_as(_dp), a.ctypes.data_as(_dp),
                                      ad.ctypes.data_as(_dp), av.ctypes.data_as(_dp) if av is not None else None)
        if rc:
            self._fail(rc)
        return (rhs, a, ad, av) if want_avars else (rhs, a, ad)

    def calc_time_step(self, courant):
        dt = C.c_double()
        self.lib.oracle_calc_dt(self._h, courant, C.byref(dt))
        return dt.value

    def update(self, courant, t0=0.0):
        """one RK4 step from solution time t0; returns the step"""
        t = C.c_double(t0)
        rc = self.lib.oracle_rk4_step(self._h, courant, C.byref(t))
        if rc:
            self._fail(rc)
        return t.value - t0

    def update_fixed_dt(self, dt):
        t = C.c_double(0.0)
        rc = self.lib.oracle_rk4_step_fixed_dt(self._h, dt, C.byref(t))
        if rc:
            self._fail(rc)
        return t.value

    def run(self, n_steps, courant, t0=0.0):
        """n_steps RK4 steps from solution time t0; returns the elapsed time"""
        t = C.c_double(t0)
        rc = self.lib.oracle_run(self._h, n_steps, courant, C.byref(t))
        if rc:
            self._fail(rc)
        return t.value - t0

    def set_local_stepping(self, request=True, start_time=0.0, limit=False, neighbor_factor=1.05,
                           global_factor=1.0):
        self.lib.oracle_set_local_stepping(self._h, int(request), start_time, int(limit), neighbor_factor,
                                           global_factor)

    def loc_dt(self):
        out = np.empty(self.n_cells)
        if self.lib.oracle_loc_dt(self._h, out.ctypes.data_as(_dp)):
            raise RuntimeError("no step taken yet")
        return out

    def set_time(self, t):
        self.lib.oracle_set_time(self._h, t)

    def set_alpha(self, alpha=None, alpha_d=None):
        """blender values (alpha, alpha_d per cell) for later RHS evaluations in place of the
        computed ones; no arguments restore the computed ones"""
        if alpha is None:
            self.lib.oracle_set_alpha(self._h, None, None)
            return
        self._alpha_given = (np.ascontiguousarray(alpha, dtype=np.float64),
                             np.ascontiguousarray(alpha_d, dtype=np.float64))
        self.lib.oracle_set_alpha(self._h, self._alpha_given[0].ctypes.data_as(_dp),
                                  self._alpha_given[1].ctypes.data_as(_dp))

    def set_raw_alpha(self, raw=None):
        """indicator values (per cell, before the wall override) for calc_blender's post-steps in
        place of get_blender's; None restores get_blender"""
        if raw is None:
            self.lib.oracle_set_raw_alpha(self._h, None)
            return
        self._raw_given = np.ascontiguousarray(raw, dtype=np.float64)
        self.lib.oracle_set_raw_alpha(self._h, self._raw_given.ctypes.data_as(_dp))

    def set_expressions(self, courant=None, wall_blender_limit=None, viscous_blending_region=None):
        """expression-valued Courant number, wall blender limit and viscous blending region"""
        enc = lambda x: x.encode() if x is not None else None
        self.lib.oracle_set_expressions(self._h, enc(courant), enc(wall_blender_limit), enc(viscous_blending_region))

    def blended_cell_stages(self):
        """(cell, RHS evaluation) pairs with alpha > 0 since creation"""
        n = C.c_int64()
        self.lib.oracle_blended_cell_stages(self._h, C.byref(n))
        return n.value

    def bc_ghost(self, stage, cell, face, q, u, av, nrm):
        """ghost state and avars of boundary face `face` of `cell` at face node q and stage 1..3"""
        g, ag = np.empty(5), np.empty(9)
        rc = self.lib.oracle_bc_ghost(self._h, stage, cell, face, q, p(u), p(av), p(nrm),
                                      g.ctypes.data_as(_dp), ag.ctypes.data_as(_dp))
        if rc:
            raise ValueError("not a boundary face")
        return g, ag

    def output_fields(self, courant, t=0.0):
        nodal = np.empty((15, self.n_dof))
        cell = np.empty((3, self.n_cells))
        rc = self.lib.oracle_output_fields(self._h, courant, t, nodal.ctypes.data_as(_dp), cell.ctypes.data_as(_dp))
        if rc:
            self._fail(rc)
        return nodal, cell

    def ss_error(self):
        e = C.c_double()
        if self.lib.oracle_ss_error(self._h, C.byref(e)):
            raise RuntimeError("no step taken yet")
        return e.value

    def flop_counts(self, reset=False):
        """(flops, calls) per bucket (FLOP_BUCKETS) of the counting build since the last reset"""
        f = (C.c_uint64 * 16)()
        k = (C.c_uint64 * 16)()
        if self.lib.oracle_flop_counts(f, k, 16, int(reset)) < 0:
            raise RuntimeError("not the flop-counting build of the oracle")
        return np.array(f[:], dtype=np.int64), np.array(k[:], dtype=np.int64)

    def section_times(self):
        t = np.zeros(8)
        self.lib.oracle_section_times(self._h, t.ctypes.data_as(_dp), 8)
        return t

    def cell_metrics(self, c, n, nf):
        Ja, dJ, nr = np.empty((9, n)), np.empty(n), np.empty((6, 3, nf))
        self.lib.oracle_cell_metrics(self._h, c, Ja.ctypes.data_as(_dp), dJ.ctypes.data_as(_dp),
                                     nr.ctypes.data_as(_dp))
        return Ja, dJ, nr

    def close(self):
        if self._h:
            self.lib.oracle_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
