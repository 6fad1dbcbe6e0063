#!/usr/bin/env python3
"""Time steps/s and tracer updates/s of the ProjecTracer hot path on one MI355X.

A step is one iteration of the reference's time loop (PS/main.f90:245-299):
one multigrid solve + velocities, then energy_time (RK4: three more solves,
four tracer sweeps, five projections with their filters).

Workload (headline): the 1024x256 thermochemical case (lambda=4, Ra_T=1e7,
Ra_C=3e6, two-layer C at z=0.3, T perturbation seed 2, 50 tracers per node
= 13,171,250 tracers) with the T- and C-dependent viscosity
eta = exp(-ln(30) T + ln(2) C), filter thresholds delta_visT = delta_visC =
1.2, courant 0.25 and an L1 relative residual tolerance of 1e-7.  This is
the strongest eta(T,C) law the reference's own multigrid solves at this
size: with the third configuration's eta = exp(-ln(1e5) T + ln(30) C) it
stops with "Abnormally High Residual" in its first solve
(tests/golden/ref_divergence1024.npz), and its FP64 residual does not reach
1e-8 at 1024x256.  The isoviscous case (eta = 1, courant 0.5) is timed
beside it.

Prints one JSON line.  Paths timed:
  * headline: PT_PREC_MIXED -- FP32 RKC1 V-cycles inside FP64 defect
    correction, stopped by the reference's FP64 L1 test
  * plain FP64 GPU path (the reference's W-cycle, bit-faithful kernels)
  * cpu_baseline: the reference program (oracle/_ref, flang -O3 -fopenmp) on
    the host cores, a window of warm time steps of the same workload.
Beside them, `direct_solver`: the direct banded solver (iterate=0,
PS/momentum.f90:319-653) at the reference's shipped 100x100 size.
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
# Algorithmic bytes of one RKC1 stage per interior point with the general
# (variable-viscosity) operator: read S1 (the stencil source, once), S0, S2,
# eta, eta_xx-eta_zz, eta_xz, eta_x, eta_z, RHS, dt_stream and write Y: 11
# words.  FP32 (the mixed path's smoother) 44 B, FP64 88 B.  The isoviscous
# operator reads 4 words fewer.
STAGE_WORDS = {"general": 11, "isoviscous": 7}
# Tracer pipeline per tracer and step: the four RK4 stage kernels move
# 96+112+112+88 B; each of the five projections reads r, s, T, C (32 B).
RK_BYTES_PER_TRACER_STEP = 408
RK2_BYTES_PER_TRACER = 112
PROJECT_BYTES_PER_TRACER = 32
PROJECTIONS_PER_STEP = 5
# One filter pass over the node field (halo included) reads the field once
# and writes it once; its check reads it once: 24 B per node.
FILTER_BYTES_PER_NODE = 24

# Viscosity laws of the 1024x256 workload (see the module docstring).
WORKLOADS = {
    "law": dict(visT=30.0, visC=0.5, delta_visT=1.2, delta_visC=1.2, courant=0.25, tolerance=1e-7),
    "isoviscous": dict(ivis=0, tolerance=1e-7),
}
LAW_TEXT = {
    "law": "eta = exp(-ln(30) T + ln(2) C), filter thresholds 1.2, courant 0.25, L1 tol 1e-7",
    "isoviscous": "eta = 1, courant 0.5, L1 tol 1e-7",
}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nx", type=int, default=1024)
    ap.add_argument("--nz", type=int, default=256)
    ap.add_argument("--tpc", type=int, default=50)
    ap.add_argument("--event-steps", type=int, default=2,
                    help="steps run after the timed ones with per-kernel HIP events")
    ap.add_argument("--workload", default="law", choices=sorted(WORKLOADS))
    ap.add_argument("--no-secondary", action="store_true", help="skip the isoviscous line")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-guess measurement")
    ap.add_argument("--no-ensemble", action="store_true", help="skip the batched-ensemble measurement")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-steps", type=int, default=5,
                    help="warm reference steps timed after its first loop iteration")
    ap.add_argument("--no-restart-io", action="store_true",
                    help="skip timing pt_write_restart / pt_read_restart after the headline run")
    ap.add_argument("--paths", default="mixed,fp64",
                    help="GPU paths to time (profiling runs may select one)")
    ap.add_argument("--direct-only", action="store_true",
                    help="time only the direct solver (100x100) and print its object")
    ap.add_argument("--no-equalize", action="store_true", help="skip the tracer redistribution measurement")
    ap.add_argument("--equalize-only", action="store_true",
                    help="time only the tracer redistribution (equalize=1, discrete types) at nx x nz")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a 256 MiB read + 256 MiB write elementwise kernel, the "
                         "reference for scaling rocprofv3 FETCH_SIZE/WRITE_SIZE")
    return ap.parse_args()


def make_case(args, precision, workload=None):
    from projectracer_amd import cases
    case = cases.thermochem_large(nx=args.nx, nz=args.nz, tpc=args.tpc, **WORKLOADS[workload or args.workload])
    case.params.precision = precision
    return case


def _interior_sf(sim):
    import numpy as np
    return np.asarray(sim.field("SF"))[3:3 + sim.nz + 1, 3:3 + sim.nx + 1].copy()


def time_gpu(args, precision, workload=None, restart_io=False, events=True):
    from projectracer_amd import Simulation, _abi
    case = make_case(args, precision, workload)
    sim = Simulation(case.params)
    sim.load_state(case)
    sim.multigrid()
    sim.compute_velocities()
    sf_first = _interior_sf(sim)
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
    out = dict(seconds=t1 - t0, ms_per_step=1000.0 * (t1 - t0) / args.steps,
               solves=inf1["solves"] - inf0["solves"], cycles=inf1["cycles"] - inf0["cycles"],
               nx=sim.nx, nz=sim.nz, ntr=sim.ntr, last_residual_ratio=inf1["residual"] / inf1["res_ref"],
               n_comp=inf1["n_comp"], n_temp=inf1["n_temp"], sf_first=sf_first, events=None, restart=None)
    if events and args.event_steps > 0:
        # per-kernel durations from a separate run of a few more steps with
        # HIP events on the library stream around the timed launches: the
        # events add host work per launch (and keep the cycles eager), so they
        # stay out of the steps/s measurement
        i0 = sim.info()
        sim.enable_kernel_timing(True)
        sim.run_steps(args.event_steps)
        sim.synchronize()
        i1 = sim.info()
        ev = {"steps": args.event_steps, "cycles": i1["cycles"] - i0["cycles"]}
        for name, k in (("stage_l1", _abi.PT_K_RKC_STAGE_L1), ("stage_coarse", _abi.PT_K_RKC_STAGE_COARSE),
                        ("tracer_stages", _abi.PT_K_TRACER_STAGE), ("rk2", _abi.PT_K_TRACER_RK2),
                        ("projection", _abi.PT_K_PROJECT), ("filter", _abi.PT_K_FILTER),
                        ("solve", _abi.PT_K_SOLVE)):
            ms, n = sim.kernel_time(k)
            ev[name] = dict(ms=ms, launches=n, work=sim.kernel_work(k))
        sim.enable_kernel_timing(False)
        out["events"] = ev
    if restart_io:
        # checkpoint / resume of the whole device state in the reference's files
        with tempfile.TemporaryDirectory(prefix="ptrst_") as td:
            t2 = time.perf_counter()
            sim.print_restart(td)
            t3 = time.perf_counter()
            nbytes = sum(f.stat().st_size for f in Path(td).iterdir())
            sim.read_restart(td, case.Tbot_iso, case.Ttop_iso)
            t4 = time.perf_counter()
        out["restart"] = dict(write_s=t3 - t2, read_s=t4 - t3, bytes=nbytes,
                              write_GBs=nbytes / (t3 - t2) / 1e9, read_GBs=nbytes / (t4 - t3) / 1e9)
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


def time_equalize(args, steps=4):
    """Tracer redistribution (equalize_tracers_original, discrete types) on
    the 1024x256 state, isoviscous, mixed path: the first step finds tpc
    tracers at every node (nothing to move); then `steps` steps, each
    redistributing.  Reports the redistribution's own time per call (HIP
    events around it, host waits included), the tracers moved per call, and
    the step time beside it."""
    from projectracer_amd import Simulation, _abi
    case = cases_large(args, iC_continuous=0, equalize=1, ivis=0, tolerance=1e-7)
    case.params.precision = _abi.PT_PREC_MIXED
    sim = Simulation(case.params)
    sim.load_state(case)
    sim.time_step()
    sim.synchronize()
    sim.enable_kernel_timing(True)
    n0, _, m0 = sim.equalize_counts()
    t0 = time.perf_counter()
    sim.run_steps(steps)
    sim.synchronize()
    t1 = time.perf_counter()
    ms, calls = sim.kernel_time(_abi.PT_K_EQUALIZE)
    n1, last, m1 = sim.equalize_counts()
    out = {
        "grid": f"{sim.nx}x{sim.nz}, {sim.ntr} tracers, buffer {case.params.buffer}, isoviscous, mixed path",
        "calls": calls, "redistributions": n1 - n0, "moved_per_call": (m1 - m0) / max(1, n1 - n0),
        "equalize_ms_per_call": ms / max(1, calls), "ms_per_step_with_timers": 1e3 * (t1 - t0) / steps,
        "dtype": "int32 counts, f64 positions",
    }
    sim.close()
    return out


def cases_large(args, **over):
    from projectracer_amd import cases
    return cases.thermochem_large(nx=args.nx, nz=args.nz, tpc=args.tpc, **over)


def time_learned(args, capture_steps=3, train_iters=100, batch=4, timed_steps=4):
    """The mixed path with the learned initial guess on the headline
    workload.  Two contexts run the same first solve and `capture_steps`
    steps (bit-identical: the solve is deterministic), one of them storing
    every solve's (inputs, correction) pair for the network; the network
    trains on those pairs (`train_iters` AdamW steps, batches of `batch`);
    then both run `timed_steps` more, one with the network's proposals."""
    from projectracer_amd import Simulation, _abi
    from projectracer_amd.learned import GuessNet, attach, guess_counts
    case = make_case(args, _abi.PT_PREC_MIXED)
    net = GuessNet(seed=0)
    out = {}
    sims = {}
    for mode in ("warm", "learned"):
        sim = Simulation(case.params)
        sim.load_state(case)
        if mode == "learned":
            net.dataset(sim.nz, sim.nx, 4 * capture_steps + 4)
            attach(sim, net, guess=False, capture=True)
        sim.multigrid()
        sim.compute_velocities()
        for _ in range(capture_steps):
            sim.energy_time()
            sim.multigrid()
            sim.compute_velocities()
        sim.synchronize()
        sims[mode] = sim
    t0 = time.perf_counter()
    loss = net.train(train_iters, batch, lr=1e-3, seed=0)
    out["train_s"] = time.perf_counter() - t0
    out["train_loss"] = loss
    out["pairs"] = net.dataset_size()
    attach(sims["learned"], net, guess=True, capture=False)
    for mode, sim in sims.items():
        i0 = sim.info()
        t0 = time.perf_counter()
        sim.run_steps(timed_steps)
        sim.synchronize()
        t1 = time.perf_counter()
        i1 = sim.info()
        out[mode] = {"ms_per_step": 1e3 * (t1 - t0) / timed_steps,
                     "vcycles_per_solve": (i1["cycles"] - i0["cycles"]) / max(1, i1["solves"] - i0["solves"])}
        if mode == "learned":
            out[mode]["proposals_tried_kept"] = guess_counts(sim)
        sim.close()
    out["vcycles_saved_frac"] = 1.0 - out["learned"]["vcycles_per_solve"] / out["warm"]["vcycles_per_solve"]
    out["note"] = (f"network trained on {out['pairs']} pairs captured from this run's first "
                   f"{capture_steps} steps; both paths then time {timed_steps} steps from the same state; "
                   "a proposal is kept only when its FP64 L1 residual is below the warm start's")
    net.close()
    return out


def time_ensemble(args, n=64, tune_steps=5, timed_steps=4):
    """The batched ensemble: n independent cases (cases.ensemble_case, seeds
    100.., grids 128x64 to 384x128, 30 tracers per node, courant 0.01) on one
    GPU through pt_ensemble_run_steps, mixed path.

    First, per case, the first solve from the initial state in FP64 and in
    the mixed path: the mixed psi's max-norm error relative to the FP64 psi
    (the FP64 path is the one the parity tests pin to the oracle), or the
    error that stopped either solve.  Then the ensemble: the first steps
    tune the host thread pool, then `timed_steps` steps of every case are
    timed with the tuned pool and again with one thread (one case at a
    time).  The aggregate counts the steps of the cases that ran through."""
    import numpy as np
    from projectracer_amd import PtError, Simulation, _abi, cases
    from projectracer_amd.ensemble import Ensemble
    t0 = time.perf_counter()
    cs = [cases.ensemble_case(100 + i) for i in range(n)]
    setup_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    errs, first_fail = [], {"fp64": {}, "mixed": {}}
    for i, c in enumerate(cs):
        sf = {}
        for label, prec in (("fp64", _abi.PT_PREC_FP64), ("mixed", _abi.PT_PREC_MIXED)):
            c.params.precision = prec
            sim = Simulation(c.params)
            sim.load_state(c)
            try:
                sim.multigrid()
                sf[label] = sim.field("SF").copy()
            except PtError as e:
                first_fail[label][i] = e.code
            sim.close()
        if len(sf) == 2:
            errs.append(float(np.max(np.abs(sf["mixed"] - sf["fp64"])) / max(1e-300, np.max(np.abs(sf["fp64"])))))
    scan_s = time.perf_counter() - t0
    for c in cs:
        c.params.precision = _abi.PT_PREC_MIXED
    ens = Ensemble([c.params for c in cs])
    for sim, c in zip(ens.cases, cs):
        sim.load_state(c)
    ntr = [c.ntr for c in cs]
    del cs
    st = ens.run_steps(tune_steps, workers=0)
    best = ens.workers()
    out = {"cases": n, "setup_s": setup_s, "tuned_workers": best}
    for label, w in (("tuned", best), ("one_thread", 1)):
        alive = st == _abi.PT_OK
        t0 = time.perf_counter()
        st = ens.run_steps(timed_steps, workers=w)
        dt = time.perf_counter() - t0
        ok = alive & (st == _abi.PT_OK)
        out[label] = {"seconds": dt, "steps_per_case": timed_steps, "cases_ok": int(ok.sum()),
                      "aggregate_time_steps_per_s": timed_steps * int(ok.sum()) / dt,
                      "aggregate_tracer_updates_per_s": timed_steps * sum(t for t, k in zip(ntr, ok) if k) / dt}
    out["failed_cases"] = {int(i): _abi.ERROR_TEXT.get(int(v), str(int(v))) for i, v in enumerate(st)
                           if v != _abi.PT_OK}
    out["concurrency_gain"] = out["tuned"]["aggregate_time_steps_per_s"] / max(
        1e-30, out["one_thread"]["aggregate_time_steps_per_s"])
    e = np.array(errs)
    out["first_solve_mixed_vs_fp64"] = {
        "cases_compared": int(e.size), "median": float(np.median(e)) if e.size else None,
        "max": float(e.max()) if e.size else None,
        "fp64_failed": {int(k): _abi.ERROR_TEXT.get(v, str(v)) for k, v in first_fail["fp64"].items()},
        "mixed_failed": {int(k): _abi.ERROR_TEXT.get(v, str(v)) for k, v in first_fail["mixed"].items()},
        "seconds": scan_s}
    out["note"] = ("mixed path; each case a full context with its own stream, stepped by a pool of host threads "
                   "(pool size tuned on the first steps); a failing case stops alone. Cases that stop do so in the "
                   "reference's algorithm too (the FP64 path, pinned to the oracle, stops on them as well)")
    ens.close()
    return out


def host_cores():
    """Cores this process may run on, capped by OMP_NUM_THREADS when set
    (the share of the host given to one GPU)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(args):
    """The reference (flang -O3 build) on the host: inject the same state,
    run one loop iteration untimed, then time `--cpu-steps` more with its own
    per-iteration wall clock (omp_get_wtime)."""
    ref = ROOT / "oracle" / "_ref" / "fast" / "ref_driver"
    if not ref.exists():
        return None
    sys.path.insert(0, str(ROOT / "tests"))
    import numpy as np
    from _support import write_inputs_nml
    from projectracer_amd import _abi
    case = make_case(args, _abi.PT_PREC_FP64)
    threads = args.cpu_threads or host_cores()
    nrun = 1 + args.cpu_steps
    with tempfile.TemporaryDirectory(prefix="ptcpu_") as td:
        tdp = Path(td)
        write_inputs_nml(tdp / "inputs.nml", case.params)
        (tdp / "driver.nml").write_text(f"&driver\nnrun={nrun}, inject=1, dump_every=-1\n/\n")
        with open(tdp / "inject.bin", "wb") as fh:
            fh.write(np.array([case.Tbot_iso, case.Ttop_iso], dtype=np.float64).tobytes())
            fh.write(np.ascontiguousarray(case.T_nodes).tobytes())
            for a in (case.rtr, case.str_, case.Ttr, case.Ctr):
                fh.write(np.ascontiguousarray(a).tobytes())
        env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_STACKSIZE="1G", OMP_PROC_BIND="close",
                   OMP_PLACES="cores")
        t0 = time.perf_counter()
        proc = subprocess.run(["bash", "-c", f"ulimit -s unlimited && exec {ref}"], cwd=tdp, env=env,
                              capture_output=True, text=True, timeout=1500)
        wall = time.perf_counter() - t0
    steps = [float(m.group(1)) for m in re.finditer(r"driver: step done\s+\d+\s+\S+\s+\S+\s+(\S+)", proc.stdout)]
    if proc.returncode != 0 or len(steps) < nrun:
        return {"error": (proc.stdout + proc.stderr)[-400:]}
    window = steps[1:nrun]
    sec = sum(window) / len(window)
    return dict(value=1.0 / sec, unit="time-steps/s", cores=threads, kind="reference",
                sample=f"reference program (flang -O3 -fopenmp, {threads} threads) on the same "
                       f"{case.nx}x{case.nz} state ({LAW_TEXT[args.workload]}) with {case.ntr} tracers: "
                       f"{len(window)} warm loop iterations after the first (one solve + one RK4 "
                       f"energy_time each) = {', '.join(f'{s:.2f}' for s in window)} s; whole run {wall:.0f} s",
                seconds_per_step=sec, tracer_updates_per_s=case.ntr / sec)


# HBM bytes per launch of the mixed path's FP32 stage kernel, measured with
# rocprofv3 FETCH_SIZE / WRITE_SIZE on the 1024x256 "law" workload and
# calibrated on a known copy (the summary and its commands sit beside it)
STAGE_TRAFFIC = ROOT / "profiles" / "k_stage_f32_general_1024x256_traffic.json"


def stage_traffic(args, ev):
    """Measured HBM bytes per stage launch, weighted by this run's launch
    mix (level 1 against levels 2-4); None off the profiled workload."""
    if args.workload != "law" or (args.nx, args.nz, args.tpc) != (1024, 256, 50) or not STAGE_TRAFFIC.exists():
        return None
    t = json.loads(STAGE_TRAFFIC.read_text())
    l1, co = ev["stage_l1"]["launches"], ev["stage_coarse"]["launches"]
    if l1 + co == 0:
        return None
    return (l1 * t["level1_bytes_per_launch"] + co * t["coarse_bytes_per_launch"]) / (l1 + co)


def stage_roofline(ev, words, bytes_per_word, label):
    """Roofline of the RKC1 stage kernel from the event steps: all its
    launches (level 1 and the coarse levels swept a launch per stage), the
    algorithmic bytes of each launch being its interior points times the
    stage's words."""
    l1, co = ev["stage_l1"], ev["stage_coarse"]
    ms = l1["ms"] + co["ms"]
    n = l1["launches"] + co["launches"]
    bpp = words * bytes_per_word
    total_bytes = bpp * (l1["work"] + co["work"])
    if n == 0 or ms <= 0:
        return None
    avg_us = 1e3 * ms / n
    achieved = total_bytes / n / (avg_us * 1e-6) / 1e9
    l1_us = 1e3 * l1["ms"] / max(1, l1["launches"])
    co_us = 1e3 * co["ms"] / max(1, co["launches"])
    return {
        "kernel": label, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": achieved / HBM_PEAK_GBS, "traffic": None,
        "avg_launch_us": avg_us, "launches": n, "bytes_per_launch": total_bytes / n,
        "bytes_per_point": bpp, "ms_per_step": ms / ev["steps"],
        "level1": {"launches": l1["launches"], "avg_launch_us": l1_us,
                   "bytes_per_launch": bpp * l1["work"] / max(1, l1["launches"]),
                   "achieved_GBs": bpp * l1["work"] / max(1, l1["launches"]) / (l1_us * 1e-6) / 1e9},
        "coarse_levels": {"launches": co["launches"], "avg_launch_us": co_us,
                          "bytes_per_launch": bpp * co["work"] / max(1, co["launches"]),
                          "achieved_GBs": bpp * co["work"] / max(1, co["launches"]) / (co_us * 1e-6) / 1e9
                          if co["launches"] else None},
    }


def step_bytes(ev, r, words, bytes_per_word):
    """Algorithmic bytes of one step: every stage launch's points x words,
    the tracer pipeline, and the filter passes of the last projection
    (n_comp, plus the first nine C passes when it is nonzero, and n_temp)
    over the halo-3 node field, for each of the five projections."""
    solver = words * bytes_per_word * (ev["stage_l1"]["work"] + ev["stage_coarse"]["work"]) / ev["steps"]
    tracers = (RK_BYTES_PER_TRACER_STEP + PROJECTIONS_PER_STEP * PROJECT_BYTES_PER_TRACER) * r["ntr"]
    nodes = (r["nx"] + 7) * (r["nz"] + 7)
    passes = r["n_comp"] + 9 * (r["n_comp"] > 0) + r["n_temp"]
    filters = PROJECTIONS_PER_STEP * passes * FILTER_BYTES_PER_NODE * nodes
    return solver + tracers + filters, dict(solver=solver, tracers=tracers, filters=filters)


def note(msg):
    """Progress on stderr (stdout carries the one JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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
    if args.equalize_only:
        print(json.dumps({"equalize": time_equalize(args)}), flush=True)
        return
    paths = set(args.paths.split(","))
    if paths != {"mixed", "fp64"}:
        # profiling run: time the selected path only, no JSON line
        for name in sorted(paths):
            prec = _abi.PT_PREC_MIXED if name == "mixed" else _abi.PT_PREC_FP64
            r = time_gpu(args, prec, events=False)
            print(f"{name}: {r['ms_per_step']:.2f} ms/step, {r['cycles']} cycles", flush=True)
        return
    import numpy as np
    note(f"{args.workload}: mixed path, {args.steps} timed steps")
    mixed = time_gpu(args, _abi.PT_PREC_MIXED, restart_io=not args.no_restart_io)
    note(f"{args.workload}: mixed {mixed['ms_per_step']:.1f} ms/step; FP64 path")
    fp64 = time_gpu(args, _abi.PT_PREC_FP64)
    note(f"{args.workload}: FP64 {fp64['ms_per_step']:.1f} ms/step")
    steps_s = args.steps / mixed["seconds"]
    fp64_steps_s = args.steps / fp64["seconds"]
    # the mixed path's first solve against the FP64 path's first solve from
    # the same injected state (both stop at the same L1 relative residual)
    den = float(np.max(np.abs(fp64["sf_first"])))
    mixed_err = float(np.max(np.abs(mixed["sf_first"] - fp64["sf_first"]))) / den
    general = args.workload != "isoviscous"
    words = STAGE_WORDS["general" if general else "isoviscous"]
    ev = mixed["events"]
    roof = stage_roofline(ev, words, 4, "k_stage<float, %s> (RKC1 stage of the mixed path's FP32 V-cycle: "
                          "level 1 and the coarse levels swept one launch per stage)"
                          % ("general operator" if general else "isoviscous"))
    if roof is not None:
        roof["traffic"] = stage_traffic(args, ev)
        if roof["traffic"] is not None:
            roof["traffic_source"] = str(STAGE_TRAFFIC.relative_to(ROOT))
    sb, sb_parts = step_bytes(ev, mixed, words, 4)
    ntr = mixed["ntr"]
    rk_ms_step = ev["tracer_stages"]["ms"] / ev["steps"]
    rk2_us = 1e3 * ev["rk2"]["ms"] / max(1, ev["rk2"]["launches"])
    pj_ms_call = ev["projection"]["ms"] / max(1, ev["projection"]["launches"])
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
            "workload": f"thermochem_{mixed['nx']}x{mixed['nz']}, {LAW_TEXT[args.workload]}",
            "nx": mixed["nx"], "nz": mixed["nz"], "aspect": 4.0, "tracers": ntr,
            "RaT": 1e7, "RaC": 3e6, "levels": 6, "rkc_stages": 35, **WORKLOADS[args.workload],
            "initial_guess": "warm start from the previous psi",
            "mixed_cycle": "V-cycle, coarse corrections from zero at each descent, levels >= 2 smoothed on the way up only with 15 RKC1 stages",
        },
        "tracer_updates_per_s": steps_s * ntr,
        "vcycles_per_solve": mixed["cycles"] / max(1, mixed["solves"]),
        "mixed_fp64_oracle_error": {
            "psi_first_solve_rel_max": mixed_err,
            "note": "max|psi_mixed - psi_FP64| / max|psi_FP64| after the first solve from the same state; "
                    "both stop at the same FP64 L1 relative residual",
            "last_residual_ratio": mixed["last_residual_ratio"],
        },
        "fp64_gpu": {
            "time_steps_per_s": fp64_steps_s, "ms_per_step": fp64["ms_per_step"],
            "tracer_updates_per_s": fp64_steps_s * fp64["ntr"],
            "vcycles_per_solve": fp64["cycles"] / max(1, fp64["solves"]),
            "cycle": "W-cycle (the reference's)",
        },
        "speedup_vs_fp64_gpu": steps_s / fp64_steps_s,
        "roofline": roof,
        "step_roofline": {
            "bytes_per_step": sb, "parts": sb_parts, "ms_per_step": mixed["ms_per_step"],
            "achieved": sb / (mixed["ms_per_step"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": sb / (mixed["ms_per_step"] * 1e-3) / 1e9 / HBM_PEAK_GBS,
        },
        "breakdown_ms_per_step_events": {
            "solves": ev["solve"]["ms"] / ev["steps"],
            "stage_launches": (ev["stage_l1"]["ms"] + ev["stage_coarse"]["ms"]) / ev["steps"],
            "filters": ev["filter"]["ms"] / ev["steps"],
            "projections": ev["projection"]["ms"] / ev["steps"],
            "tracer_rk_stages": rk_ms_step,
            "cycles_per_step": ev["cycles"] / ev["steps"],
            "note": "HIP events with the cycles launched eagerly; each span includes its waits",
        },
        "tracer_roofline": {
            "rk4_stages": {
                "bytes_per_tracer_step": RK_BYTES_PER_TRACER_STEP, "ms_per_step": rk_ms_step,
                "achieved_GBs": RK_BYTES_PER_TRACER_STEP * ntr / (rk_ms_step * 1e-3) / 1e9,
                "frac": RK_BYTES_PER_TRACER_STEP * ntr / (rk_ms_step * 1e-3) / 1e9 / HBM_PEAK_GBS,
            },
            "k_rk2": {
                "bytes_per_launch": RK2_BYTES_PER_TRACER * ntr, "avg_launch_us": rk2_us,
                "achieved_GBs": RK2_BYTES_PER_TRACER * ntr / (rk2_us * 1e-6) / 1e9,
                "frac": RK2_BYTES_PER_TRACER * ntr / (rk2_us * 1e-6) / 1e9 / HBM_PEAK_GBS,
            },
            "projection": {
                "bytes_per_tracer_call": PROJECT_BYTES_PER_TRACER, "ms_per_call": pj_ms_call,
                "achieved_GBs": PROJECT_BYTES_PER_TRACER * ntr / (pj_ms_call * 1e-3) / 1e9,
                "frac": PROJECT_BYTES_PER_TRACER * ntr / (pj_ms_call * 1e-3) / 1e9 / HBM_PEAK_GBS,
            },
        },
    }
    if mixed["restart"]:
        # print_restart + restart=1 read of the whole state (Trestart, restart_tracers) to a
        # temporary directory on the host
        result["restart_io"] = mixed["restart"]
    if not args.no_secondary and args.workload != "isoviscous":
        note("isoviscous: mixed and FP64 paths")
        im = time_gpu(args, _abi.PT_PREC_MIXED, workload="isoviscous", events=False)
        i64 = time_gpu(args, _abi.PT_PREC_FP64, workload="isoviscous", events=False)
        result["isoviscous"] = {
            "workload": f"thermochem_{im['nx']}x{im['nz']}, {LAW_TEXT['isoviscous']}",
            "mixed_time_steps_per_s": args.steps / im["seconds"], "mixed_ms_per_step": im["ms_per_step"],
            "mixed_vcycles_per_solve": im["cycles"] / max(1, im["solves"]),
            "fp64_time_steps_per_s": args.steps / i64["seconds"], "fp64_ms_per_step": i64["ms_per_step"],
            "fp64_vcycles_per_solve": i64["cycles"] / max(1, i64["solves"]),
            "speedup_vs_fp64_gpu": i64["seconds"] / im["seconds"],
        }
    if not args.no_ensemble:
        note("ensemble of 64 cases")
        result["ensemble"] = time_ensemble(args)
    if not args.no_learned:
        note("learned guess")
        result["learned_guess"] = time_learned(args)
    note("direct solver")
    result["direct_solver"] = time_direct()
    if not args.no_equalize:
        note("tracer redistribution")
        result["equalize"] = time_equalize(args)
    if not args.no_cpu_baseline:
        note("CPU baseline (the reference program on the host cores)")
        result["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
