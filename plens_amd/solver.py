"""Host-side mirror of the PLENS hot-path members (dgsem/plens.h:827-896) over the GPU C ABI.

    PLENS::calc_rhs()        -> GpuPlens.calc_rhs()        (dgsem/plens.cc:3688-3779)
    PLENS::calc_time_step()  -> GpuPlens.calc_time_step()  (dgsem/plens.cc:3810-3912, global stepping)
    PLENS::update()          -> GpuPlens.update()          (dgsem/plens.cc:4484-4518, RK order 4)
    PLENS::run() loop        -> GpuPlens.run(n_steps)      (dgsem/plens.cc:1547)

A positivity failure raises PositivityError carrying the lowest failing DoF and its kind, where
the reference prints the location and calls std::exit (dgsem/plens.cc:2540-2544).
"""
import ctypes as C

import numpy as np

from . import _lib as L


class PositivityError(L.PlensError):
    def __init__(self, code, dof, kind):
        self.dof = dof
        self.kind = {1: "density", 2: "energy"}.get(kind, "unknown")
        super().__init__(code, f"positivity failed at dof {dof} ({self.kind})")


def _dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def write_output_files(mesh, directory, base, counter, time, nodal, cell, record_times, wall_time=0.0):
    """plens_output_write: vtu piece, pvtu record, pvd collection and .times log of one output"""
    nodal = np.ascontiguousarray(nodal, dtype=np.float64)
    cell = np.ascontiguousarray(cell, dtype=np.float64)
    rt = np.ascontiguousarray(record_times, dtype=np.float64)
    L.check(L.load().plens_output_write(mesh.handle, str(directory).encode(), str(base).encode(), counter, time,
                                        wall_time, _dptr(nodal), _dptr(cell), _dptr(rt), len(rt)),
            "plens_output_write")


class GpuPlens:
    def __init__(self, case=None, mesh=None, physics=None, bcs=None, cells_per_block=0, metrics="map"):
        """metrics: "map" derives the metric terms in the kernels from each cell's trilinear map
        (cell vertices handed over); "stored" reads the host-computed Ja and detJ arrays."""
        if case is not None:
            mesh, physics, bcs = case.mesh, case.physics, case.bcs
        self.lib = L.load()
        self.mesh = mesh
        self.n_cells, self.n_dof = mesh.n_cells, mesh.n_dof
        self._desc = mesh.desc(with_vertices={"map": True, "stored": False}[metrics])
        arr = (L.BC * max(len(bcs), 1))()
        for i, b in enumerate(bcs):
            arr[i] = b
        self._bcs = arr
        self._phys = physics
        self._h = C.c_void_p()
        L.check(self.lib.plens_gpu_create(C.byref(self._desc), C.byref(self._phys), arr, len(bcs),
                                          C.byref(self._h)), "plens_gpu_create")
        if cells_per_block:
            L.check(self.lib.plens_gpu_set_tuning(self._h, cells_per_block), "plens_gpu_set_tuning")
        if case is not None:
            self.set_state(case.Q0)

    def _check(self, code, what):
        if code == L.PLENS_E_POSITIVITY:
            dof, kind = C.c_int64(), C.c_int32()
            self.lib.plens_gpu_last_failure(self._h, C.byref(dof), C.byref(kind))
            raise PositivityError(code, dof.value, kind.value)
        L.check(code, what)

    def set_state(self, Q):
        Q = np.ascontiguousarray(Q, dtype=np.float64).reshape(5, self.n_dof)
        self._check(self.lib.plens_gpu_set_state(self._h, _dptr(Q)), "set_state")

    def get_state(self):
        Q = np.empty((5, self.n_dof))
        self._check(self.lib.plens_gpu_get_state(self._h, _dptr(Q)), "get_state")
        return Q

    def calc_rhs(self):
        rhs = np.empty((5, self.n_dof))
        a = np.empty(self.n_cells)
        ad = np.empty(self.n_cells)
        self._check(self.lib.plens_gpu_calc_rhs(self._h, _dptr(rhs), _dptr(a), _dptr(ad)), "calc_rhs")
        return rhs, a, ad

    def calc_time_step(self, courant, cell_steps=False):
        """global step (calc_time_step, global stepping); with cell_steps also every cell's own step"""
        dt = C.c_double()
        dc = np.empty(self.n_cells) if cell_steps else None
        self._check(self.lib.plens_gpu_calc_dt(self._h, courant, C.byref(dt), _dptr(dc) if cell_steps else None),
                    "calc_dt")
        return (dt.value, dc) if cell_steps else dt.value

    def update_with_step(self, dt, t0=0.0):
        """one RK4 step of the given global step dt from solution time t0"""
        self._check(self.lib.plens_gpu_rk4_step_dt(self._h, t0, dt), "rk4_step_dt")

    def update(self, courant, t0=0.0):
        """one RK4 step from solution time t0; returns the step"""
        t = C.c_double(t0)
        self._check(self.lib.plens_gpu_rk4_step(self._h, courant, C.byref(t)), "rk4_step")
        return t.value - t0

    def run(self, n_steps, courant, t0=0.0):
        """n_steps RK4 steps from solution time t0; returns the elapsed time"""
        t = C.c_double(t0)
        self._check(self.lib.plens_gpu_run(self._h, n_steps, courant, C.byref(t)), "run")
        return t.value - t0

    def blended_cell_stages(self, reset=False):
        """(cell, stage) pairs with alpha > 0 over the RHS evaluations since creation or the last reset"""
        n = C.c_int64()
        self._check(self.lib.plens_gpu_blended_cell_stages(self._h, C.byref(n), int(reset)), "blended_cell_stages")
        return n.value

    def last_run_ms(self):
        ms = C.c_double()
        self._check(self.lib.plens_gpu_last_run_ms(self._h, C.byref(ms)), "last_run_ms")
        return ms.value

    def run_profiled(self, n_steps, courant):
        t = C.c_double(0.0)
        self._check(self.lib.plens_gpu_run_profiled(self._h, n_steps, courant, C.byref(t)), "run_profiled")
        times = np.zeros(5)
        self.lib.plens_gpu_kernel_times(self._h, _dptr(times), 5)
        return t.value, times

    def set_local_stepping(self, request=True, start_time=0.0, limit=False, neighbor_factor=1.05,
                           global_factor=1.0):
        self._check(self.lib.plens_gpu_set_local_stepping(self._h, int(request), start_time, int(limit),
                                                          neighbor_factor, global_factor), "set_local_stepping")

    def ss_error(self):
        e = C.c_double()
        self._check(self.lib.plens_gpu_ss_error(self._h, C.byref(e)), "ss_error")
        return e.value

    def set_expressions(self, courant=None, wall_blender_limit=None, viscous_blending_region=None):
        """expression-valued Courant number, wall blender limit and viscous blending region (prm
        strings in x, y, z, t; None keeps the numeric parameter / every cell)"""
        enc = lambda x: x.encode() if x is not None else None
        self._check(self.lib.plens_gpu_set_expressions(self._h, enc(courant), enc(wall_blender_limit),
                                                       enc(viscous_blending_region)), "set_expressions")

    def set_time(self, t):
        """solution time of the handle (the BCs' time of calc_rhs)"""
        self._check(self.lib.plens_gpu_set_time(self._h, t), "set_time")

    def output_fields(self, courant, t=0.0):
        """fields of PLENS::write() for the current state: nodal [15, n_dof] on each cell's
        equispaced lattice (txx..qz, rho, p, T, u, v, w), cell [3, n_cells] (alpha, alpha_d, loc_dt)"""
        nodal = np.empty((15, self.n_dof))
        cell = np.empty((3, self.n_cells))
        self._check(self.lib.plens_gpu_output_fields(self._h, courant, t, _dptr(nodal), _dptr(cell)), "output_fields")
        return nodal, cell

    def write_checkpoint(self, path, time, step=0, output_counter=0):
        self._check(self.lib.plens_gpu_write_checkpoint(self._h, str(path).encode(), time, step, output_counter),
                    "write_checkpoint")

    def read_checkpoint(self, path):
        """restore the state from a checkpoint; returns (time, step, output_counter)"""
        t, st, oc = C.c_double(), C.c_int64(), C.c_int64()
        self._check(self.lib.plens_gpu_read_checkpoint(self._h, str(path).encode(), C.byref(t), C.byref(st),
                                                       C.byref(oc)), "read_checkpoint")
        return t.value, st.value, oc.value

    def write_output(self, directory, base, counter, courant, time, record_times, wall_time=0.0):
        """PLENS::write(): the fields of the current state and the vtu/pvtu/pvd/times files;
        record_times = solution times of outputs 0..counter"""
        nodal, cell = self.output_fields(courant, time)
        write_output_files(self.mesh, directory, base, counter, time, nodal, cell, record_times, wall_time)
        return nodal, cell

    def time_kernel(self, kernel, reps=20):
        """average duration (ms) of one launch of kernel class `kernel` ("grad", "face", "lo" or "vol"),
        launched `reps` times back to back after an RHS evaluation; the state is unchanged"""
        k = {"grad": 0, "face": 2, "lo": 3, "vol": 4}[kernel]
        ms = C.c_double()
        self._check(self.lib.plens_gpu_time_kernel(self._h, k, reps, C.byref(ms)), "time_kernel")
        return ms.value

    def set_tuning(self, cells_per_block):
        self._check(self.lib.plens_gpu_set_tuning(self._h, cells_per_block), "set_tuning")

    def close(self):
        if self._h:
            self.lib.plens_gpu_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
