"""Batched uneven solves through the C ABI of include/plens_gpu.h (plens_gpu_batch_*): many
independent GpuPlens handles advanced together, one HIP graph launch per step for all of them,
with launch shapes from the autotuner's tuning file."""
import ctypes as C

import numpy as np

from . import _lib as L
from .solver import PositivityError


class GpuBatch:
    def __init__(self, solvers, tuning_db=None):
        self.lib = L.load()
        self.solvers = list(solvers)
        self.n = len(self.solvers)
        arr = (C.c_void_p * self.n)(*[s._h.value for s in self.solvers])
        self._h = C.c_void_p()
        db = tuning_db.encode() if tuning_db else None
        L.check(self.lib.plens_gpu_batch_create(arr, self.n, db, C.byref(self._h)), "plens_gpu_batch_create")

    def _check(self, code, what):
        if code == L.PLENS_E_POSITIVITY:
            inst, dof, kind = C.c_int32(), C.c_int64(), C.c_int32()
            self.lib.plens_gpu_batch_last_failure(self._h, C.byref(inst), C.byref(dof), C.byref(kind))
            err = PositivityError(code, dof.value, kind.value)
            err.instance = inst.value
            raise err
        L.check(code, what)

    def run(self, n_steps, courant):
        """advance every instance n_steps at its Courant number (a scalar or one per instance);
        returns each instance's elapsed time"""
        cfl = np.ascontiguousarray(np.broadcast_to(np.asarray(courant, np.float64), (self.n,)))
        t = np.zeros(self.n)
        self._check(self.lib.plens_gpu_batch_run(self._h, n_steps, cfl.ctypes.data_as(C.POINTER(C.c_double)),
                                                 t.ctypes.data_as(C.POINTER(C.c_double))), "batch_run")
        return t

    def get_state(self, i):
        Q = np.empty((5, self.solvers[i].n_dof))
        self._check(self.lib.plens_gpu_batch_get_state(self._h, i, Q.ctypes.data_as(C.POINTER(C.c_double))),
                    "batch_get_state")
        return Q

    def last_run_ms(self):
        ms = C.c_double()
        self._check(self.lib.plens_gpu_batch_last_run_ms(self._h, C.byref(ms)), "batch_last_run_ms")
        return ms.value

    def groups(self):
        out = []
        for g in range(self.lib.plens_gpu_batch_n_groups(self._h)):
            deg, ni, cpb, ft = C.c_int32(), C.c_int32(), C.c_int32(), C.c_int32()
            nd = C.c_int64()
            self._check(self.lib.plens_gpu_batch_group_info(self._h, g, C.byref(deg), C.byref(ni), C.byref(nd),
                                                            C.byref(cpb), C.byref(ft)), "batch_group_info")
            out.append({"degree": deg.value, "instances": ni.value, "n_dof": nd.value,
                        "cells_per_workgroup": cpb.value, "face_workgroup": ft.value})
        return out

    def close(self):
        if self._h:
            self.lib.plens_gpu_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
