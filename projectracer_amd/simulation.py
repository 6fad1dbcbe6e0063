"""Host-side mirror of the reference's hot-path subroutines over the C ABI.

`Simulation` owns one device context.  Its methods carry the names of the
reference subroutines they replace (PS/iterative_solver.f90 `multigrid`,
PS/momentum.f90 `compute_velocities`/`update_velocities`, PS/energy.f90
`energy_time`, PS/composition.f90 `tracers_to_corners`), take no state
arguments, and raise `PtError` with the reference's own message where the
reference would `stop`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class PtError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        msg = _abi.ERROR_TEXT.get(code, f"error {code}")
        super().__init__(f"{where}: {msg}" + (f" ({detail})" if detail else ""))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Simulation:
    def __init__(self, params: _abi.PtParams, lib=None):
        self.lib = lib or _abi.load_library()
        self.ctx = C.c_void_p()
        self._check(self.lib.pt_create(C.byref(params), C.byref(self.ctx)), "pt_create")
        nx, nz, g = C.c_int32(), C.c_int32(), C.c_int32()
        ntr = C.c_int64()
        self._check(self.lib.pt_get_dims(self.ctx, C.byref(nx), C.byref(nz), C.byref(g), C.byref(ntr)),
                    "pt_get_dims")
        self.nx, self.nz, self.grids, self.ntr = nx.value, nz.value, g.value, ntr.value
        self.params = params

    @classmethod
    def view(cls, ctx: C.c_void_p, params: _abi.PtParams, lib):
        """A Simulation over a context it does not own (an ensemble's case):
        close() leaves the context to its owner."""
        self = cls.__new__(cls)
        self.lib = lib
        self.ctx = ctx
        self._owned = False
        nx, nz, g = C.c_int32(), C.c_int32(), C.c_int32()
        ntr = C.c_int64()
        self._check(self.lib.pt_get_dims(self.ctx, C.byref(nx), C.byref(nz), C.byref(g), C.byref(ntr)),
                    "pt_get_dims")
        self.nx, self.nz, self.grids, self.ntr = nx.value, nz.value, g.value, ntr.value
        self.params = params
        return self

    # -- errors ------------------------------------------------------------
    def _check(self, rc: int, where: str):
        if rc != _abi.PT_OK:
            detail = self.lib.pt_last_error()
            raise PtError(rc, where, detail.decode() if detail else "")

    def close(self):
        if self.ctx and getattr(self, "_owned", True):
            self.lib.pt_destroy(self.ctx)
        self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- state -------------------------------------------------------------
    def load_state(self, case):
        T = np.ascontiguousarray(case.T_nodes, dtype=np.float64)
        arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (case.rtr, case.str_, case.Ttr, case.Ctr)]
        if any(a.size != self.ntr for a in arrs):
            raise ValueError(f"case has {[a.size for a in arrs]} tracer values, context expects {self.ntr} each")
        if T.size != (self.nx + 1) * (self.nz + 1):
            raise ValueError(f"case has {T.size} node temperatures, context expects {(self.nx + 1) * (self.nz + 1)}")
        self._check(self.lib.pt_load_state(self.ctx, _dp(T), case.Tbot_iso, case.Ttop_iso,
                                           *[_dp(a) for a in arrs]), "pt_load_state")

    # -- reference subroutines -----------------------------------------------
    def multigrid(self):
        self._check(self.lib.pt_multigrid(self.ctx), "multigrid")

    def compute_velocities(self):
        self._check(self.lib.pt_compute_velocities(self.ctx), "compute_velocities")

    def update_velocities(self):
        self._check(self.lib.pt_update_velocities(self.ctx), "update_velocities")

    def build_sf_matrix(self):
        """build_SF_matrix (PS/momentum.f90:321-488), iterate=0 contexts."""
        self._check(self.lib.pt_build_sf_matrix(self.ctx), "build_sf_matrix")

    def solve_sf_equation(self):
        """solve_SF_equation (PS/momentum.f90:621-653), iterate=0 contexts."""
        self._check(self.lib.pt_solve_sf_equation(self.ctx), "solve_sf_equation")

    def energy_time(self):
        self._check(self.lib.pt_energy_time(self.ctx), "energy_time")

    def tracers_to_corners(self):
        self._check(self.lib.pt_tracers_to_corners(self.ctx), "tracers_to_corners")

    def time_step(self):
        self._check(self.lib.pt_time_step(self.ctx), "time_step")

    def run_steps(self, n: int):
        self._check(self.lib.pt_run_steps(self.ctx, int(n)), "run_steps")

    def synchronize(self):
        self._check(self.lib.pt_synchronize(self.ctx), "synchronize")

    # -- exchange ------------------------------------------------------------
    def field(self, name: str) -> np.ndarray:
        """Field in Fortran layout: array[k + h, i + h] for the reference
        element (i, k) with halo h (RESIDUAL: array[k-1, i-1])."""
        fid = _abi.FIELD[name]
        n = self.lib.pt_field_extent(self.ctx, fid)
        out = np.empty(n, dtype=np.float64)
        self._check(self.lib.pt_get_field(self.ctx, fid, _dp(out)), f"get_field({name})")
        if fid == _abi.FIELD["RESIDUAL"]:
            return out.reshape(self.nz - 1, self.nx - 1)
        h = _abi.FIELD_HALO[fid]
        return out.reshape(self.nz + 1 + 2 * h, self.nx + 1 + 2 * h)

    def set_field(self, name: str, arr: np.ndarray):
        fid = _abi.FIELD[name]
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        if a.size != self.lib.pt_field_extent(self.ctx, fid):
            raise ValueError("field size mismatch")
        self._check(self.lib.pt_set_field(self.ctx, fid, _dp(a)), f"set_field({name})")

    def tracers(self):
        out = [np.empty(self.ntr, dtype=np.float64) for _ in range(4)]
        self._check(self.lib.pt_get_tracers(self.ctx, *[_dp(a) for a in out]), "get_tracers")
        return tuple(out)

    def nempty(self) -> np.ndarray:
        out = np.empty((self.nz + 1) * (self.nx + 1), dtype=np.int32)
        self._check(self.lib.pt_get_nempty(self.ctx, out.ctypes.data_as(C.POINTER(C.c_int32))),
                    "get_nempty")
        return out.reshape(self.nz + 1, self.nx + 1)

    def projection_counts(self) -> tuple:
        """(projections sorted by the windowed counting sort, by the radix sort)."""
        a, b = C.c_int64(), C.c_int64()
        self._check(self.lib.pt_projection_counts(self.ctx, C.byref(a), C.byref(b)), "projection_counts")
        return a.value, b.value

    def equalize_counts(self) -> tuple:
        """(redistributions made, tracers moved by the last one, tracers moved
        in total) of equalize_tracers_original (PS/composition.f90:84-263)."""
        a, b, c = C.c_int64(), C.c_int64(), C.c_int64()
        self._check(self.lib.pt_equalize_counts(self.ctx, C.byref(a), C.byref(b), C.byref(c)), "equalize_counts")
        return a.value, b.value, c.value

    def rng_state(self) -> tuple:
        """mzran stream 0 (seed_i, seed_j, seed_k, seed_n), the stream the
        redistribution draws from."""
        st = (C.c_int32 * 4)()
        self._check(self.lib.pt_get_rng_state(self.ctx, st), "get_rng_state")
        return tuple(int(v) for v in st)

    def set_rng_state(self, state) -> None:
        st = (C.c_int32 * 4)(*[int(v) for v in state])
        self._check(self.lib.pt_set_rng_state(self.ctx, st), "set_rng_state")

    def info(self) -> dict:
        inf = _abi.PtSolveInfo()
        self._check(self.lib.pt_get_info(self.ctx, C.byref(inf)), "get_info")
        return {name: getattr(inf, name) for name, _ in _abi.PtSolveInfo._fields_}

    def stats(self) -> dict:
        """stats (PS/basics.f90:363-462): the stats.dat and
        viscosity_smoother.dat values of the current state."""
        row = _abi.PtStatsRow()
        self._check(self.lib.pt_stats(self.ctx, C.byref(row)), "stats")
        return {name: getattr(row, name) for name, _ in _abi.PtStatsRow._fields_}

    # -- checkpoint / resume ---------------------------------------------------
    def print_restart(self, directory):
        """print_restart (PS/main.f90:457-496): Trestart and restart_tracers
        in `directory`, in the reference's unformatted layout."""
        self._check(self.lib.pt_write_restart(self.ctx, str(directory).encode()), "print_restart")

    def snapshots(self, directory, fcount: int = 0):
        """snapshots (PS/main.f90:377-424): the ASCII files T<10000+fcount>
        and P<10000+fcount> in `directory`."""
        self._check(self.lib.pt_write_snapshot(self.ctx, str(directory).encode(), fcount), "snapshots")

    def read_restart(self, directory, Tbot_iso: float, Ttop_iso: float):
        """The restart=1 start (PS/grid.f90:55-103) from the files in
        `directory`; the wall temperatures come from the caller."""
        self._check(self.lib.pt_read_restart(self.ctx, str(directory).encode(), Tbot_iso, Ttop_iso),
                    "read_restart")

    # -- timing ----------------------------------------------------------------
    def enable_kernel_timing(self, on: bool = True):
        self._check(self.lib.pt_enable_kernel_timing(self.ctx, 1 if on else 0), "enable_kernel_timing")

    def kernel_work(self, which: int) -> int:
        n = C.c_int64()
        self._check(self.lib.pt_kernel_work(self.ctx, which, C.byref(n)), "kernel_work")
        return n.value

    def kernel_time(self, which: int):
        ms = C.c_double()
        n = C.c_int64()
        self._check(self.lib.pt_kernel_time(self.ctx, which, C.byref(ms), C.byref(n)), "kernel_time")
        return ms.value, n.value
