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
# algorithmic bytes of one level-1 RKC1 stage per interior point: read S0, S1,
# S2, eta, eta_xx-eta_zz, eta_xz, eta_x, eta_z, RHS, dt_stream, write Y
RKC_BYTES_PER_POINT = {"f64": 96, "f32": 48}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--tpc", type=int, default=50)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def make_case(args, precision):
    from projectracer_amd import cases, _abi
    case = cases.thermochem_large(nx=args.nx, nz=args.nz, tpc=args.tpc, ivis=0, tolerance=1e-7)
    case.params.precision = precision
    return case


def time_gpu(args, precision):
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
    sim.enable_kernel_timing(True)
    sim.synchronize()
    t0 = time.perf_counter()
    sim.run_steps(args.steps)
    sim.synchronize()
    t1 = time.perf_counter()
    inf1 = sim.info()
    kms, klaunch = sim.kernel_time(_abi.PT_K_RKC_STAGE_L1)
    tms, tlaunch = sim.kernel_time(_abi.PT_K_TRACER_STAGE)
    pms, plaunch = sim.kernel_time(_abi.PT_K_PROJECT)
    sim.enable_kernel_timing(False)
    out = dict(
        seconds=t1 - t0,
        ms_per_step=1000.0 * (t1 - t0) / args.steps,
        solves=inf1["solves"] - inf0["solves"],
        cycles=inf1["cycles"] - inf0["cycles"],
        rkc_ms=kms, rkc_launches=klaunch, tracer_ms=tms, tracer_launches=tlaunch,
        project_ms=pms, project_launches=plaunch,
        nx=sim.nx, nz=sim.nz, ntr=sim.ntr, last_residual_ratio=inf1["residual"] / inf1["res_ref"],
    )
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
    mixed = time_gpu(args, _abi.PT_PREC_MIXED)
    fp64 = time_gpu(args, _abi.PT_PREC_FP64)
    steps_s = args.steps / mixed["seconds"]
    fp64_steps_s = args.steps / fp64["seconds"]
    interior = (mixed["nx"] - 1) * (mixed["nz"] - 1)
    # roofline of the dominant solver kernel (level-1 RKC1 stage), measured
    # with HIP events on the library stream inside the timed region
    rk = fp64
    avg_ms = rk["rkc_ms"] / max(1, rk["rkc_launches"])
    bytes_per_launch = RKC_BYTES_PER_POINT["f64"] * interior
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
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
            "kernel": "k_stage<double> (level-1 RKC1 stage, FP64 path)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "avg_launch_us": avg_ms * 1e3, "launches": rk["rkc_launches"],
            "bytes_per_launch": bytes_per_launch,
        },
        "kernel_share_fp64": {
            "rkc_stage_l1_ms_per_step": fp64["rkc_ms"] / args.steps,
            "tracer_rk_ms_per_step": fp64["tracer_ms"] / args.steps,
            "projection_ms_per_step": fp64["project_ms"] / args.steps,
        },
    }
    if not args.no_cpu_baseline:
        result["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
