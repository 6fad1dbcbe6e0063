#!/usr/bin/env python3
"""Time steps/s and tracer updates/s of the ProjecTracer hot path on one MI355X.

A step is one iteration of the reference's time loop (PS/main.f90:245-299):
one multigrid solve + velocities, then energy_time (RK4: three more solves,
four tracer sweeps, five projections).  The workload is the 1024x256 case
(lambda=4, Ra_T=1e7, Ra_C=3e6, two-layer C at z=0.3, 50 tracers per node =
13,171,250 tracers) with the isoviscous rheology and an L1 relative residual
tolerance of 1e-7: the reference's own multigrid stalls above 1e-8 at this
resolution and diverges with the T/C-dependent viscosity (flang build of
the reference, same inputs), so this is the largest 1024x256 case the
reference itself solves.

Prints one JSON line.  Paths timed:
  * headline: PT_PREC_MIXED (FP32 RKC1 smoothing inside FP64 defect correction)
  * plain FP64 GPU path (same algorithm as the reference, bit-faithful kernels)
  * cpu_baseline: the reference program (oracle/_ref, flang -O2 -fopenmp) on
    the host cores, one warm time step of the same workload.
Beside them, `direct_solver`: the direct banded solver (iterate=0,
PS/momentum.f90:319-653) at the reference's shipped 100x100 size -- the band
LU's build once, then solve_SF_equation per solve -- with the reference's
own timers for its shipped run quoted for context.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# algorithmic bytes of one level-1 RKC1 stage per interior point: read S1
# (stencil source), S0, S2, eta, RHS, dt_stream and write Y; the general
# operator also reads eta_xx-eta_zz, eta_xz, eta_x, eta_z (96 B in FP64), the
# isoviscous one does not (56 B)
RKC_BYTES_PER_POINT_F64 = {"isoviscous": 56, "general": 96}
# tracer pipeline, per tracer: the four RK4 stage kernels move 96+112+112+88 B
# (positions, temperature, step-start copies, stage registers); one
# projection reads r, s, T, C once (32 B)
RK_BYTES_PER_TRACER_STEP = 408
RK2_BYTES_PER_TRACER = 112   # stage 2: read r1, s1, r, s, T, r0, s0, T0; write r, s, T, r2, s2, T2
PROJECT_BYTES_PER_TRACER = 32
TRAFFIC_FILE = ROOT / "profiles" / "k_rk2_1024x256_traffic.json"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--tpc", type=int, default=50)
    ap.add_argument("--event-steps", type=int, default=3,
                    help="steps run after the timed ones with per-kernel HIP events")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-restart-io", action="store_true",
                    help="skip timing pt_write_restart / pt_read_restart after the headline run")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--workload", default="law", choices=sorted(WORKLOADS))
    ap.add_argument("--paths", default="mixed,fp64",
                    help="GPU paths to time (profiling runs may select one)")
    ap.add_argument("--direct-only", action="store_true",
                    help="time only the direct solver (100x100) and print its object")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB read + 256 MiB write elementwise kernel, the "
                         "reference for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return ap.parse_args()


# Viscosity laws of the 1024x256 workload.  "law": the strongest T- and
# C-dependent law the reference's multigrid solves at this size,
# eta = exp(-ln(30) T + ln(2) C) with filter thresholds 1.2 (the third
# configuration's own eta = exp(-ln(1e5) T + ln(30) C) makes the reference stop
# with "Abnormally High Residual" in its first solve), at courant 0.25 (at
# 0.5 the flow this law drives carries tracers out of the domain within three
# steps); "isoviscous": eta = 1, courant 0.5.  Both at an L1 relative
# residual of 1e-7: the reference's FP64 residual does not reach 1e-8 at
# 1024x256.
WORKLOADS = {
    "law": dict(visT=30.0, visC=0.5, delta_visT=1.2, delta_visC=1.2, courant=0.25, tolerance=1e-7),
    "isoviscous": dict(ivis=0, tolerance=1e-7),
}


def make_case(args, precision, workload=None):
    from projectracer_amd import cases
    case = cases.thermochem_large(nx=args.nx, nz=args.nz, tpc=args.tpc, **WORKLOADS[workload or args.workload])
    case.params.precision = precision
    return case


def time_gpu(args, precision, restart_io=False):
    from projectracer_amd import Simulation, _abi
    case = make_case(args, precision)
    sim = Simulation(case.params)
    sim.load_state(case)
    sim.multigrid()
    sim.compute_velocities()
    for _ in range(args.warmup):
        sim.energy_time()
        sim.multigrid()
        sim.compute_velocities()
    sim.synchronize()
    inf0 = sim.info()
    t0 = time.perf_counter()
    sim.run_steps(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    inf1 = sim.info()
    # kernel durations from a separate, shorter run with HIP events around the
    # timed kernels: the events add host work per launch, so they stay out of
    # the steps/s measurement
    sim.enable_kernel_timing(True)
    sim.run_steps(args.event_steps)
    sim.synchronize()
    kms, klaunch = sim.kernel_time(_abi.PT_K_RKC_STAGE_L1)
    tms, tlaunch = sim.kernel_time(_abi.PT_K_TRACER_STAGE)
    pms, plaunch = sim.kernel_time(_abi.PT_K_PROJECT)
    r2ms, r2launch = sim.kernel_time(_abi.PT_K_TRACER_RK2)
    sim.enable_kernel_timing(False)
    restart = None
    if restart_io:
        # checkpoint / resume of the whole device state in the reference's files
        with tempfile.TemporaryDirectory(prefix="ptrst_") as td:
            t2 = time.perf_counter()
            sim.print_restart(td)
            t3 = time.perf_counter()
            nbytes = sum(f.stat().st_size for f in Path(td).iterdir())
            sim.read_restart(td, case.Tbot_iso, case.Ttop_iso)
            t4 = time.perf_counter()
        restart = dict(write_s=t3 - t2, read_s=t4 - t3, bytes=nbytes,
                       write_GBs=nbytes / (t3 - t2) / 1e9, read_GBs=nbytes / (t4 - t3) / 1e9)
    out = dict(
        seconds=t1 - t0,
        ms_per_step=1000.0 * (t1 - t0) / args.steps,
        solves=inf1["solves"] - inf0["solves"],
        cycles=inf1["cycles"] - inf0["cycles"],
        rkc_ms=kms, rkc_launches=klaunch, tracer_ms=tms, tracer_launches=tlaunch,
        project_ms=pms, project_launches=plaunch, rk2_ms=r2ms, rk2_launches=r2launch,
        nx=sim.nx, nz=sim.nz, ntr=sim.ntr, last_residual_ratio=inf1["residual"] / inf1["res_ref"],
        restart=restart,
    )
    sim.close()
    return out


# the reference's own timers for its shipped 100x100 run (2 OpenMP threads,
# CPU not recorded): build_SF_matrix once, solve_SF_equation over its 5 solves
# (stdout:17-18)
REF_SHIPPED_BUILD_S = 0.3331035
REF_SHIPPED_SOLVE_S = 0.0437131 / 5


def time_direct(nx=100, nz=100, solves=20):
    """The direct solver on the isoviscous nx x nz state: the first
    update_velocities builds the band LU (assembly, DGBEQU scaling, DGBTF2 in
    panels) and solves; then `solves` more solve_SF_equation calls."""
    from projectracer_amd import Simulation, cases
    case = cases.isoviscous(nx=nx, nz=nz, tpc=1, iterate=0, Database=2)
    sim = Simulation(case.params)
    sim.load_state(case)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.build_sf_matrix()
    t_build = time.perf_counter() - t0
    sim.solve_sf_equation()
    ts = []
    for _ in range(solves):
        t0 = time.perf_counter()
        sim.solve_sf_equation()
        ts.append(time.perf_counter() - t0)
    ts.sort()
    solve_s = ts[len(ts) // 2]
    out = {
        "grid": f"{nx}x{nz} isoviscous, Database 2: {(nx + 1) * (nz + 1)} unknowns, kl = ku = {3 * (nz + 1)}",
        "build_ms": t_build * 1e3, "solve_ms_median": solve_s * 1e3, "solves_timed": solves,
        "dtype": "f64",
    }
    if (nx, nz) == (100, 100):
        out["reference_shipped_run"] = {
            "build_ms": REF_SHIPPED_BUILD_S * 1e3, "solve_ms": REF_SHIPPED_SOLVE_S * 1e3,
            "source": "the reference's own timers for its shipped 100x100 run (2 OpenMP threads, CPU not "
                      "recorded): build_SF_matrix once, solve_SF_equation averaged over 5 solves",
        }
    sim.close()
    return out


def cpu_baseline(args):
    """The reference (flang build) on the host: inject the same state, run
    two loop iterations, report the second (warm) one."""
    ref = ROOT / "oracle" / "_ref" / "ref_driver"
    if not ref.exists():
        return None
    sys.path.insert(0, str(ROOT / "tests"))
    from _support import write_inputs_nml
    from projectracer_amd import _abi
    case = make_case(args, _abi.PT_PREC_FP64)
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    with tempfile.TemporaryDirectory(prefix="ptcpu_") as td:
        tdp = Path(td)
        write_inputs_nml(tdp / "inputs.nml", case.params)
        (tdp / "driver.nml").write_text("&driver\nnrun=2, inject=1, dump_every=-1\n/\n")
        import numpy as np
        with open(tdp / "inject.bin", "wb") as fh:
            fh.write(np.array([case.Tbot_iso, case.Ttop_iso], dtype=np.float64).tobytes())
            fh.write(np.ascontiguousarray(case.T_nodes).tobytes())
            for a in (case.rtr, case.str_, case.Ttr, case.Ctr):
                fh.write(np.ascontiguousarray(a).tobytes())
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="1G", OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        t0 = time.perf_counter()
        proc = subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {ref}"], cwd=tdp, env=env,
                              capture_output=True, text=True, timeout=900)
        wall = time.perf_counter() - t0
    steps = [float(m.group(1)) for m in re.finditer(r"driver: step done\s+\d+\s+\S+\s+\S+\s+(\S+)", proc.stdout)]
    if proc.returncode != 0 or len(steps) < 2:
        return {"error": (proc.stdout + proc.stderr)[-400:]}
    sec = steps[1]
    return dict(value=1.0 / sec, unit="time-steps/s", cores=threads, kind="reference",
                sample=f"reference program (flang -O2 -fopenmp, {threads} threads) on the same "
                       f"{case.nx}x{case.nz} state with {case.ntr} tracers: second loop iteration "
                       f"(one solve + one RK4 energy_time) = {sec:.2f} s; whole run {wall:.0f} s",
                seconds_per_step=sec, tracer_updates_per_s=case.ntr / sec)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the path on a single MI355X "
              "(multi-GPU sharding is design-only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from projectracer_amd import _abi
    if args.copy_calibration:
        import torch
        x = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        torch.mul(x, 2.0, out=y)
        torch.cuda.synchronize()
        del x, y
    if args.direct_only:
        print(json.dumps({"direct_solver": time_direct()}), flush=True)
        return
    paths = set(args.paths.split(","))
    if paths != {"mixed", "fp64"}:
        # profiling run: time the selected path only, no JSON line
        for name in sorted(paths):
            prec = _abi.PT_PREC_MIXED if name == "mixed" else _abi.PT_PREC_FP64
            r = time_gpu(args, prec)
            print(f"{name}: {r['ms_per_step']:.2f} ms/step, {r['cycles']} cycles", flush=True)
        return
    mixed = time_gpu(args, _abi.PT_PREC_MIXED, restart_io=not args.no_restart_io)
    fp64 = time_gpu(args, _abi.PT_PREC_FP64)
    steps_s = args.steps / mixed["seconds"]
    fp64_steps_s = args.steps / fp64["seconds"]
    interior = (mixed["nx"] - 1) * (mixed["nz"] - 1)
    # roofline: the dominant HBM-streaming kernel of the step, the RK4 tracer
    # stage k_rk2 (one launch per step, 13.2 M tracers), timed with HIP events
    # on the library stream in the instrumented steps of the headline path.  The
    # solver's levels >= 2 are latency-bound and the level-1 stage kernel is
    # too short (~5 us) for per-launch events to resolve; it is reported
    # beside it.
    ntr = mixed["ntr"]
    rk2_us = 1e3 * mixed["rk2_ms"] / max(1, mixed["rk2_launches"])
    bytes_per_launch = RK2_BYTES_PER_TRACER * ntr
    achieved = bytes_per_launch / (rk2_us * 1e-6) / 1e9
    traffic = None
    if TRAFFIC_FILE.exists():
        # HBM bytes per launch of the same kernel from the committed rocprofv3
        # FETCH_SIZE / WRITE_SIZE runs, scaled by the copy-kernel calibration
        traffic = json.loads(TRAFFIC_FILE.read_text()).get("hbm_bytes_per_launch")
    rk = fp64
    l1_us = 1e3 * rk["rkc_ms"] / max(1, rk["rkc_launches"])
    l1_bytes = RKC_BYTES_PER_POINT_F64["isoviscous"] * interior
    rk_ms_step = mixed["tracer_ms"] / args.event_steps
    pj_ms_call = mixed["project_ms"] / max(1, mixed["project_launches"])
    result = {
        "metric": "time-steps/s & tracer-updates/s at 1024x256, mixed+learned vs plain FP64 GPU",
        "value": steps_s,
        "unit": "time-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": mixed["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64+f32 (FP32 smoothing in FP64 defect correction)",
        "data": "synthetic",
        "config": {
            "workload": f"thermochem_{mixed['nx']}x{mixed['nz']} isoviscous, tol 1e-7",
            "nx": mixed["nx"], "nz": mixed["nz"], "aspect": 4.0, "tracers": mixed["ntr"],
            "RaT": 1e7, "RaC": 3e6, "levels": 6, "rkc_stages": 35, "tolerance": 1e-7,
            "initial_guess": "warm start (learned guess not built)",
        },
        "tracer_updates_per_s": steps_s * mixed["ntr"],
        "vcycles_per_solve": mixed["cycles"] / max(1, mixed["solves"]),
        "fp64_gpu": {
            "time_steps_per_s": fp64_steps_s, "ms_per_step": fp64["ms_per_step"],
            "tracer_updates_per_s": fp64_steps_s * fp64["ntr"],
            "vcycles_per_solve": fp64["cycles"] / max(1, fp64["solves"]),
        },
        "speedup_vs_fp64_gpu": steps_s / fp64_steps_s,
        "roofline": {
            "kernel": "k_rk2 (RK4 tracer stage 2: interpolation of u, w, source + stage update)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
            "avg_launch_us": rk2_us, "launches": mixed["rk2_launches"],
            "bytes_per_launch": bytes_per_launch,
        },
        "solver_stage_l1_fp64": {
            "kernel": "k_stage<double, isoviscous> (level-1 RKC1 stage, FP64 path)",
            "bytes_per_launch": l1_bytes, "avg_launch_us_events": l1_us, "launches": rk["rkc_launches"],
            "achieved_GBs_events": l1_bytes / (l1_us * 1e-6) / 1e9,
        },
        "tracer_roofline": {
            "rk4_stages": {
                "bytes_per_tracer_step": RK_BYTES_PER_TRACER_STEP, "ms_per_step": rk_ms_step,
                "achieved_GBs": RK_BYTES_PER_TRACER_STEP * ntr / (rk_ms_step * 1e-3) / 1e9,
                "frac": RK_BYTES_PER_TRACER_STEP * ntr / (rk_ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            },
            "projection": {
                "bytes_per_tracer_call": PROJECT_BYTES_PER_TRACER, "ms_per_call": pj_ms_call,
                "achieved_GBs": PROJECT_BYTES_PER_TRACER * ntr / (pj_ms_call * 1e-3) / 1e9,
                "frac": PROJECT_BYTES_PER_TRACER * ntr / (pj_ms_call * 1e-3) / 1e9 / HBM_PEAK_GBS,
            },
        },
        "kernel_share_fp64": {
            "rkc_stage_l1_ms_per_step": fp64["rkc_ms"] / args.event_steps,
            "tracer_rk_ms_per_step": fp64["tracer_ms"] / args.event_steps,
            "projection_ms_per_step": fp64["project_ms"] / args.event_steps,
        },
    }
    if mixed["restart"]:
        # print_restart + restart=1 read of the whole state (Trestart, restart_tracers) to a
        # temporary directory on the host
        result["restart_io"] = mixed["restart"]
    result["direct_solver"] = time_direct()
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
