"""Ensembles of independent cases advanced together on one GPU (C ABI
pt_ensemble_*): each case is a full context with its own stream; a pool of
host threads steps the cases so that their kernels overlap on the device.

No reference counterpart: the reference runs one case per process
(PS/main.f90:245-300).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from .simulation import PtError, Simulation


class Ensemble:
    def __init__(self, params_list, lib=None):
        self.lib = lib or _abi.load_library()
        self.params = list(params_list)
        arr = (_abi.PtParams * len(self.params))(*self.params)
        self.ptr = C.c_void_p()
        rc = self.lib.pt_ensemble_create(len(self.params), arr, C.byref(self.ptr))
        if rc != _abi.PT_OK:
            raise PtError(rc, "pt_ensemble_create", self.lib.pt_last_error().decode())
        self.cases = []
        for i, p in enumerate(self.params):
            ctx = C.c_void_p()
            self.lib.pt_ensemble_case(self.ptr, i, C.byref(ctx))
            self.cases.append(Simulation.view(ctx, p, self.lib))

    def __len__(self):
        return len(self.cases)

    def run_steps(self, nsteps: int, workers: int = 0) -> np.ndarray:
        """Advance every case by nsteps time steps; workers <= 0 tunes the
        host thread pool on the first steps.  Returns each case's status."""
        st = np.zeros(len(self.cases), dtype=np.int32)
        rc = self.lib.pt_ensemble_run_steps(self.ptr, int(nsteps), int(workers),
                                            st.ctypes.data_as(C.POINTER(C.c_int32)))
        if rc != _abi.PT_OK:
            raise PtError(rc, "pt_ensemble_run_steps", self.lib.pt_last_error().decode())
        return st

    def workers(self) -> int:
        w = C.c_int32()
        self.lib.pt_ensemble_workers(self.ptr, C.byref(w))
        return w.value

    def close(self):
        if self.ptr:
            for s in self.cases:
                s.close()
            self.lib.pt_ensemble_destroy(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
