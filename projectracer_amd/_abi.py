"""ctypes view of the C ABI declared in include/projectracer_gpu.h.

This module only describes the ABI (structs, constants, signatures) and
loads the HIP library that implements it.  Loading fails loudly when the
library has not been built: there is no CPU fallback for the hot path.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
LIB_PATH = PKG_DIR / "lib" / "libprojectracer_gpu.so"

PT_OK = 0
PT_ERR_TRACER_OUTSIDE = 1
PT_ERR_MAX_CYCLES = 2
PT_ERR_HIGH_RESIDUAL = 3
PT_ERR_SINGULAR_COARSE = 4
PT_ERR_HIP = 5
PT_ERR_INVALID = 6
PT_ERR_UNSUPPORTED = 7
PT_ERR_NO_DEVICE = 8

PT_PREC_FP64 = 0
PT_PREC_MIXED = 1
PT_GUESS_WARM = 0
PT_GUESS_LEARNED = 1

FIELD = dict(SF=0, T=1, TVIS=2, TBUOY=3, U=4, W=5, CNEW=6, CVIS=7, CBUOY=8,
             CONDUCT=9, VIS=10, RESIDUAL=11, SOURCE=12)
# halo width of each exchanged field (Fortran lower bound is -halo, except
# RESIDUAL which is (1:nx-1,1:nz-1))
FIELD_HALO = {0: 3, 1: 3, 2: 3, 3: 3, 4: 3, 5: 3, 6: 5, 7: 5, 8: 5, 9: 5, 10: 3, 12: 3}

PT_K_RKC_STAGE_L1 = 0
PT_K_TRACER_STAGE = 1
PT_K_PROJECT = 2
PT_K_TRACER_RK2 = 3
PT_K_RKC_STAGE_COARSE = 4
PT_K_FILTER = 5
PT_K_SOLVE = 6
PT_K_EQUALIZE = 7

ERROR_TEXT = {
    PT_ERR_TRACER_OUTSIDE: "Tracer advected outside of domain - stopping",
    PT_ERR_MAX_CYCLES: "Multigrid Convergence Problem: Too Many Multigrid Cycles Needed",
    PT_ERR_HIGH_RESIDUAL: "Multigrid Convergence Problem: Abnormally High Residual Detected -- stopping",
    PT_ERR_SINGULAR_COARSE: "singular coarse-grid matrix",
    PT_ERR_HIP: "HIP runtime error",
    PT_ERR_INVALID: "invalid argument",
    PT_ERR_UNSUPPORTED: "option outside the supported subset",
    PT_ERR_NO_DEVICE: "no gfx950 device visible",
}


class PtParams(C.Structure):
    """pt_params (namelist values of PS/basics.f90:54-66)."""
    _fields_ = [
        ("aspect", C.c_double), ("nx", C.c_int32), ("nz", C.c_int32),
        ("grids", C.c_int32), ("mingrid", C.c_int32),
        ("tolerance", C.c_double), ("max_cycles", C.c_int32), ("Nmulti", C.c_int32),
        ("eig_ratio", C.c_double), ("vis_iter", C.c_int32), ("nstageS", C.c_int32),
        ("dampingS", C.c_double), ("c_2", C.c_double), ("c_1", C.c_double),
        ("courant", C.c_double), ("courantT", C.c_double), ("courantC", C.c_double),
        ("courant_stream", C.c_double),
        ("tpc", C.c_int32), ("Tbot", C.c_int32), ("Vbc", C.c_int32),
        ("RaT", C.c_double), ("RaC", C.c_double),
        ("H", C.c_double), ("conduct_factor", C.c_double), ("Htype", C.c_double),
        ("ivis", C.c_int32),
        ("visP", C.c_double), ("visT", C.c_double), ("delta_visT", C.c_double),
        ("eig_ratio_min", C.c_double), ("visC", C.c_double), ("delta_visC", C.c_double),
        ("precision", C.c_int32), ("guess", C.c_int32),
        ("dent", C.c_double), ("iterate", C.c_int32), ("Database", C.c_int32),
        ("iC_continuous", C.c_int32), ("equalize", C.c_int32), ("original", C.c_int32),
        ("buffer", C.c_int32), ("dist_factor", C.c_double),
        ("RKC1", C.c_int32),
    ]


class PtSolveInfo(C.Structure):
    _fields_ = [
        ("q", C.c_int32), ("residual", C.c_double), ("res_ref", C.c_double),
        ("dt", C.c_double), ("time", C.c_double), ("solves", C.c_int64),
        ("cycles", C.c_int64), ("n_eta", C.c_int32), ("n_temp", C.c_int32),
        ("n_comp", C.c_int32), ("eig_fine", C.c_double), ("cycle_graphs", C.c_int32),
    ]


class PtStatsRow(C.Structure):
    """pt_stats_row: the stats.dat / viscosity_smoother.dat values of
    PS/basics.f90:363-462."""
    _fields_ = [
        ("time", C.c_double), ("Tavg", C.c_double), ("vrms", C.c_double), ("Nu", C.c_double),
        ("Nu_bot", C.c_double), ("RaT_avg", C.c_double), ("mobility", C.c_double),
        ("strain_max", C.c_double), ("q1", C.c_double), ("q2", C.c_double), ("q3", C.c_double),
        ("q4", C.c_double), ("empty_cells", C.c_int32), ("tracer_min", C.c_int32),
        ("tracer_max", C.c_int32), ("Cavg", C.c_double), ("ent", C.c_double),
        ("vis_min", C.c_double), ("vis_max", C.c_double), ("n_eta", C.c_int32),
        ("n_temp", C.c_int32), ("n_comp", C.c_int32), ("eig_fine", C.c_double),
    ]


def default_params(**overrides) -> PtParams:
    """Reference defaults for the supported path (inputs.nml values; 6
    levels, RKC1 35 stages, tolerance 1e-8), with keyword overrides."""
    p = PtParams(
        aspect=1.0, nx=64, nz=64, grids=6, mingrid=4,
        tolerance=1e-8, max_cycles=1000, Nmulti=1, eig_ratio=1.0, vis_iter=12,
        nstageS=35, dampingS=1000.0, c_2=0.95, c_1=0.9,
        courant=0.5, courantT=0.99, courantC=0.99, courant_stream=0.99,
        tpc=20, Tbot=0, Vbc=0, RaT=1e5, RaC=0.0, H=0.0, conduct_factor=1.0, Htype=0.0,
        ivis=0, visP=1.0, visT=1e3, delta_visT=2.0, eig_ratio_min=2.0, visC=10.0,
        delta_visC=2.0, precision=PT_PREC_FP64, guess=PT_GUESS_WARM, dent=0.5,
        iterate=1, Database=2, iC_continuous=1, equalize=0, original=1, buffer=10, dist_factor=1e-4, RKC1=1)
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def params_dict(p: PtParams) -> dict:
    return {name: getattr(p, name) for name, _ in PtParams._fields_}


_lib = None


def load_library(path: os.PathLike | None = None):
    """Load libprojectracer_gpu.so (HIP, gfx950).  Raises if it is missing:
    the hot path has no CPU fallback."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    lp = Path(path) if path else LIB_PATH
    if not lp.exists():
        raise RuntimeError(
            f"{lp} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(str(lp))
    vp = C.c_void_p
    dp = C.POINTER(C.c_double)
    ip = C.POINTER(C.c_int32)
    sig = {
        "pt_default_params": (None, [C.POINTER(PtParams)]),
        "pt_create": (C.c_int, [C.POINTER(PtParams), C.POINTER(vp)]),
        "pt_destroy": (C.c_int, [vp]),
        "pt_get_dims": (C.c_int, [vp, ip, ip, ip, C.POINTER(C.c_int64)]),
        "pt_load_state": (C.c_int, [vp, dp, C.c_double, C.c_double, dp, dp, dp, dp]),
        "pt_multigrid": (C.c_int, [vp]),
        "pt_compute_velocities": (C.c_int, [vp]),
        "pt_update_velocities": (C.c_int, [vp]),
        "pt_build_sf_matrix": (C.c_int, [vp]),
        "pt_solve_sf_equation": (C.c_int, [vp]),
        "pt_energy_time": (C.c_int, [vp]),
        "pt_tracers_to_corners": (C.c_int, [vp]),
        "pt_time_step": (C.c_int, [vp]),
        "pt_run_steps": (C.c_int, [vp, C.c_int32]),
        "pt_field_extent": (C.c_int64, [vp, C.c_int32]),
        "pt_get_field": (C.c_int, [vp, C.c_int32, dp]),
        "pt_set_field": (C.c_int, [vp, C.c_int32, dp]),
        "pt_get_tracers": (C.c_int, [vp, dp, dp, dp, dp]),
        "pt_get_nempty": (C.c_int, [vp, ip]),
        "pt_get_info": (C.c_int, [vp, C.POINTER(PtSolveInfo)]),
        "pt_stats": (C.c_int, [vp, C.POINTER(PtStatsRow)]),
        "pt_stats_noalloc": (C.c_int, [vp, C.POINTER(PtStatsRow)]),
        "pt_write_restart": (C.c_int, [vp, C.c_char_p]),
        "pt_write_snapshot": (C.c_int, [vp, C.c_char_p, C.c_int32]),
        "pt_read_restart": (C.c_int, [vp, C.c_char_p, C.c_double, C.c_double]),
        "pt_synchronize": (C.c_int, [vp]),
        "pt_enable_kernel_timing": (C.c_int, [vp, C.c_int32]),
        "pt_kernel_time": (C.c_int, [vp, C.c_int32, dp, C.POINTER(C.c_int64)]),
        "pt_kernel_work": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_int64)]),
        "pt_last_error": (C.c_char_p, []),
        "pt_ensemble_create": (C.c_int, [C.c_int32, C.POINTER(PtParams), C.POINTER(vp)]),
        "pt_ensemble_destroy": (C.c_int, [vp]),
        "pt_ensemble_case": (C.c_int, [vp, C.c_int32, C.POINTER(vp)]),
        "pt_ensemble_run_steps": (C.c_int, [vp, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]),
        "pt_ensemble_workers": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "pt_net_create": (C.c_int, [C.c_int32, C.POINTER(vp)]),
        "pt_net_destroy": (C.c_int, [vp]),
        "pt_net_num_params": (C.c_int, [vp, C.POINTER(C.c_int64)]),
        "pt_net_get_params": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "pt_net_set_params": (C.c_int, [vp, C.POINTER(C.c_float)]),
        "pt_net_forward": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                     C.POINTER(C.c_float)]),
        "pt_net_loss_grad": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_float),
                                       C.POINTER(C.c_float), C.POINTER(C.c_float), dp, C.POINTER(C.c_float)]),
        "pt_net_adamw_step": (C.c_int, [vp, C.c_double, C.c_double]),
        "pt_net_dataset": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_int32]),
        "pt_net_dataset_size": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "pt_net_dataset_get": (C.c_int, [vp, C.c_int32, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
        "pt_net_train": (C.c_int, [vp, C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_uint64, dp]),
        "pt_net_out_scale": (C.c_int, [vp, dp]),
        "pt_net_set_loss": (C.c_int, [vp, C.c_double]),
        "pt_set_guess_net": (C.c_int, [vp, vp, C.c_int32, C.c_int32]),
        "pt_guess_counts": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "pt_guess_last": (C.c_int, [vp, dp, dp, dp]),
        "pt_projection_counts": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "pt_equalize_counts": (C.c_int, [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int64)]),
        "pt_get_rng_state": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "pt_set_rng_state": (C.c_int, [vp, C.POINTER(C.c_int32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


EXPORTED_SYMBOLS = [
    "pt_default_params", "pt_create", "pt_destroy", "pt_get_dims", "pt_load_state",
    "pt_multigrid", "pt_compute_velocities", "pt_update_velocities", "pt_build_sf_matrix",
    "pt_solve_sf_equation", "pt_energy_time",
    "pt_tracers_to_corners", "pt_time_step", "pt_run_steps", "pt_field_extent",
    "pt_get_field", "pt_set_field", "pt_get_tracers", "pt_get_nempty", "pt_get_info",
    "pt_stats", "pt_stats_noalloc", "pt_write_restart", "pt_write_snapshot", "pt_read_restart", "pt_synchronize", "pt_enable_kernel_timing",
    "pt_kernel_time", "pt_kernel_work", "pt_last_error",
    "pt_ensemble_create", "pt_ensemble_destroy", "pt_ensemble_case", "pt_ensemble_run_steps",
    "pt_ensemble_workers", "pt_net_create", "pt_net_destroy", "pt_net_num_params", "pt_net_get_params", "pt_net_set_params",
    "pt_net_forward", "pt_net_loss_grad", "pt_net_adamw_step", "pt_net_dataset", "pt_net_dataset_size", "pt_net_dataset_get",
    "pt_net_train", "pt_net_out_scale", "pt_net_set_loss", "pt_set_guess_net", "pt_guess_counts", "pt_guess_last",
    "pt_projection_counts", "pt_equalize_counts", "pt_get_rng_state", "pt_set_rng_state",
]
