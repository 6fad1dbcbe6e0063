#!/usr/bin/env python3
"""Write the golden fixtures under tests/golden/ (run in a container that has
the reference sources; the fixtures are committed, this script is not needed
to run the tests).

Two sources:
  * the reference's own shipped output files (data, read as text):
      T10000  -- initial snapshot of the shipped 100x100, tpc=60, 2-thread run
                 (PS/main.f90:377-408 writes xg, zg, T, SF, u, w, stress, vis,
                 Tvis, nempty, Cnew, Cvis per node); columns 9-12 kept.
      Trestart -- print_restart's record after the run's last step
                 (PS/main.f90:457-470): xg, zg, T, Cnew, SF as exact FP64,
                 kept whole with the record length;
      T10001  -- the snapshot of the same step: T, SF and Cnew columns, to
                 cross-check which array is which;
      T10000, P10000 text -- every 7th line of T10000 and all of P10000,
                 verbatim, for the snapshot writer's layout and editing;
      stats.dat, stdout -- scalar values quoted as constants in the tests.
  * the reference program compiled from its Fortran sources by
    oracle/Makefile (`make -C oracle ref` -> oracle/_ref/ref_driver), run on
    small injected states:
      ref_iso64.npz     -- 64x64 isoviscous, tpc=20, two steps (states after
                           the solves of steps 0 and 2), 1 thread
      ref_thermo96.npz  -- 96x32 T/C-dependent viscosity, tpc=4, one step
      ref_cloud.npz     -- initial tracer cloud drawn by the reference's own
                           initialize_tracers with 2 and 3 OpenMP threads
      ref_config1_100.npz -- the 64x64 isoviscous Ra_T=1e5 configuration
                           (tpc=20, RK4, courant 0.5) after 100 time steps:
                           psi, u, w, T at the nodes, per-node tracer counts,
                           every tracer's cell, a tracer subsample, q of every
                           solve
      ref_direct64.npz  -- the direct solver (iterate=0, Database=2) on the
                           64x64 isoviscous state, tpc=4: states after the
                           solves of steps 0, 1 and 2
      ref_direct_db1.npz -- the direct solver with Database=1 on a 48x32
                           thermochemical state (Ra_C=5e4, isoviscous), tpc=4:
                           states after the solves of steps 0 and 1
      ref_variants.npz  -- 48x32 (aspect 1.5), tpc=4, two steps, for the branches
                           no shipped file covers: rigid top/bottom with an
                           insulating bottom (Vbc=1, Tbot=1) through multigrid
                           and through the direct solver, and the T- and
                           depth-dependent viscosity law ivis=1; q of every
                           solve, states after the solves of steps 0 and 2
      ref_equalize.npz  -- tracer redistribution (equalize=1, original=1,
                           discrete types) on the EQUALIZE_RUNS states, three
                           steps, 1 thread: after each energy_time nequalize
                           and mzran stream 0's state; after each step's solve
                           the node counts, every tracer's type, every 7th
                           tracer's r, s, T, and (steps 1 and 3) psi, T, C
      ref_rk1.npz       -- the RK1 fine grid smoother (RKC1=0, Nmulti=70) on the
                           48x32 isoviscous and ivis=2 states, tpc=4, two
                           steps: q of every solve, node counts, psi, u, w, T
                           after the solves of steps 0 and 2
Usage: python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
import tempfile
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parents[1]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "tests"))

REF_DIR = Path("/root/reference")

NODE_FIELDS3 = ("SF", "T", "U", "W", "TVIS", "TBUOY", "VIS")
NODE_FIELDS5 = ("CNEW", "CVIS", "CONDUCT")


def _node(st, name):
    h = 3 if name in NODE_FIELDS3 else 5
    nx, nz = st["nx"], st["nz"]
    return st[name][h:h + nz + 1, h:h + nx + 1].copy()


def _state_arrays(prefix, st, names, tracer_stride):
    out = {}
    for n in names:
        out[f"{prefix}_{n}"] = _node(st, n)
    out[f"{prefix}_nempty"] = st["nempty"].astype(np.int32)
    sel = np.arange(0, st["ntr"], tracer_stride)
    for n in ("rtr", "str", "Ttr", "Ctr"):
        out[f"{prefix}_{n}"] = st[n][sel].copy()
    out[f"{prefix}_tracer_index"] = sel.astype(np.int64)
    out[f"{prefix}_dt"] = np.float64(st["dt"])
    out[f"{prefix}_time"] = np.float64(st["time"])
    return out


def _solves(solves):
    return np.array([(t, q) for t, q, _, _ in solves], dtype=np.int64), \
        np.array([(r, rr) for _, _, r, rr in solves], dtype=np.float64)


def shipped_snapshot():
    rows = []
    for line in (REF_DIR / "T10000").read_text().splitlines():
        parts = line.split()
        if len(parts) >= 12:
            rows.append(parts[:12])
    a = np.array(rows)
    nx = nz = 100
    assert a.shape[0] == (nx + 1) * (nz + 1)
    # written with i outer, k inner (PS/main.f90:390-392): reshape to (i, k)
    def col(c, dt):
        return a[:, c].astype(dt).reshape(nx + 1, nz + 1).T.copy()
    np.savez_compressed(HERE / "shipped_T10000.npz",
                        x=col(0, np.float64)[0, :], z=col(1, np.float64)[:, 0],
                        Tvis=col(8, np.float64), nempty=col(9, np.int32),
                        Cnew=col(10, np.float64), Cvis=col(11, np.float64))
    print("wrote shipped_T10000.npz")


def shipped_restart():
    raw = (REF_DIR / "Trestart").read_bytes()
    head = int.from_bytes(raw[:4], "little")
    nx = nz = 100
    h, h3, h5 = 6, 3, 5
    nxg, n3, n5 = nx + 1 + 2 * h, (nx + 1 + 2 * h3) * (nz + 1 + 2 * h3), (nx + 1 + 2 * h5) * (nz + 1 + 2 * h5)
    d = np.frombuffer(raw[4:-4], dtype="<f8")
    assert head == d.nbytes == 8 * (2 * nxg + 2 * n3 + n5)
    xg, zg = d[:nxg], d[nxg:2 * nxg]
    T = d[2 * nxg:2 * nxg + n3]
    C = d[2 * nxg + n3:2 * nxg + n3 + n5]
    SF = d[2 * nxg + n3 + n5:]
    rows = [line.split()[:12] for line in (REF_DIR / "T10001").read_text().splitlines() if len(line.split()) >= 12]
    a = np.array(rows)
    def col(c):
        return a[:, c].astype(np.float64).reshape(nx + 1, nz + 1).T.copy()
    np.savez_compressed(HERE / "shipped_Trestart.npz", record_bytes=np.int64(head), xg=xg, zg=zg, T=T, Cnew=C, SF=SF,
                        snap_T=col(2), snap_SF=col(3), snap_Cnew=col(10))
    print("wrote shipped_Trestart.npz")


def shipped_snapshot_text():
    t = (REF_DIR / "T10000").read_bytes().split(b"\n")
    p = (REF_DIR / "P10000").read_bytes().split(b"\n")
    assert t[-1] == b"" and p[-1] == b""
    idx = np.arange(0, len(t) - 1, 7)
    np.savez_compressed(HERE / "shipped_snapshot_text.npz", t_count=np.int64(len(t) - 1), t_idx=idx,
                        t_lines=np.array([t[i] for i in idx], dtype="S160"), p_lines=np.array(p[:-1], dtype="S56"))
    print("wrote shipped_snapshot_text.npz")


def reference_runs():
    from projectracer_amd import cases
    from _support import run_reference, REF_BIN
    if not REF_BIN.exists():
        raise SystemExit("oracle/_ref/ref_driver missing: run `make -C oracle ref` first")
    tmp = Path(tempfile.mkdtemp(prefix="golden_"))

    iso = cases.isoviscous(nx=64, nz=64, tpc=20)
    states, solves, _, init = run_reference(iso, nrun=2, dump_every=2, threads=1, workdir=tmp / "iso")
    out = {}
    for step in (0, 2):
        out.update(_state_arrays(f"s{step}", states[step], ("SF", "T", "U", "W", "TVIS", "CNEW"), 37))
    out["solves_tq"], out["solves_res"] = _solves(solves)
    for k in ("D1", "D2", "D3", "D4", "D", "bases", "mewt", "mew", "v", "cc"):
        out[f"init_{k}"] = init[k]
    out["init_stability"] = np.float64(init["stability"])
    out["init_dts"] = np.float64(init["dts"])
    out["init_delta_T"] = np.float64(init["delta_T"])
    out["init_delta_C"] = np.float64(init["delta_C"])
    np.savez_compressed(HERE / "ref_iso64.npz", **out)
    print("wrote ref_iso64.npz", [(t, q) for t, q, _, _ in solves])

    th = cases.thermochem(nx=96, nz=32, tpc=4, RaT=1e5, RaC=5e4, courant=0.1)
    states, solves, _, _ = run_reference(th, nrun=1, dump_every=1, threads=1, workdir=tmp / "thermo")
    out = {}
    for step in (0, 1):
        out.update(_state_arrays(f"s{step}", states[step],
                                 ("SF", "T", "U", "W", "TVIS", "CNEW", "CVIS", "VIS", "CONDUCT"), 11))
    out["solves_tq"], out["solves_res"] = _solves(solves)
    np.savez_compressed(HERE / "ref_thermo96.npz", **out)
    print("wrote ref_thermo96.npz", [(t, q) for t, q, _, _ in solves])

    out = {}
    for threads in (2, 3):
        cl = cases.isoviscous(nx=22, nz=10, tpc=3)
        states, _, _, _ = run_reference(cl, nrun=0, dump_every=0, inject=False, threads=threads,
                                        workdir=tmp / f"cloud{threads}")
        st = states[0]
        out[f"t{threads}_rtr"] = st["rtr"].copy()
        out[f"t{threads}_str"] = st["str"].copy()
    np.savez_compressed(HERE / "ref_cloud.npz", **out)
    print("wrote ref_cloud.npz")
    reference_config1(tmp)
    reference_direct(tmp)
    reference_variants(tmp)
    reference_equalize(tmp)


def reference_equalize(tmp, nrun=3, stride=7):
    from _support import run_reference, EQUALIZE_RUNS, equalize_case, parse_equalize
    out = {}
    for name in EQUALIZE_RUNS:
        states, _, stdout, _ = run_reference(equalize_case(name), nrun=nrun, dump_every=1, threads=1,
                                             workdir=tmp / f"equalize_{name}")
        out[f"{name}_eq"] = parse_equalize(stdout)
        for step in range(nrun + 1):
            st = states[step]
            out[f"{name}_s{step}_nempty"] = st["nempty"].astype(np.int16)
            out[f"{name}_s{step}_type"] = st["Ctr"].astype(np.int8)
            sel = np.arange(0, st["ntr"], stride)
            for n in ("rtr", "str", "Ttr"):
                out[f"{name}_s{step}_{n}"] = st[n][sel].copy()
            if step in (1, nrun):
                for n in ("SF", "T", "CNEW"):
                    out[f"{name}_s{step}_{n}"] = _node(st, n)
        print(name, "equalize rows", out[f"{name}_eq"].tolist())
    out["stride"] = np.int64(stride)
    np.savez_compressed(HERE / "ref_equalize.npz", **out)
    print("wrote ref_equalize.npz")


def reference_direct(tmp):
    from projectracer_amd import cases
    from _support import run_reference
    iso = cases.isoviscous(nx=64, nz=64, tpc=4, iterate=0, Database=2)
    states, _, _, _ = run_reference(iso, nrun=2, dump_every=1, threads=1, workdir=tmp / "direct64")
    out = {}
    for step in (0, 1, 2):
        for n in ("SF", "U", "W", "T"):
            out[f"s{step}_{n}"] = _node(states[step], n)
        out[f"s{step}_nempty"] = states[step]["nempty"].astype(np.int32)
        out[f"s{step}_dt"] = np.float64(states[step]["dt"])
    np.savez_compressed(HERE / "ref_direct64.npz", **out)
    print("wrote ref_direct64.npz")
    th = cases.thermochem(nx=48, nz=32, tpc=4, aspect=1.5, RaT=1e5, RaC=5e4, ivis=0, iterate=0, Database=1)
    states, _, _, _ = run_reference(th, nrun=1, dump_every=1, threads=1, workdir=tmp / "direct_db1")
    out = {}
    for step in (0, 1):
        for n in ("SF", "U", "W", "T"):
            out[f"s{step}_{n}"] = _node(states[step], n)
        out[f"s{step}_nempty"] = states[step]["nempty"].astype(np.int32)
    np.savez_compressed(HERE / "ref_direct_db1.npz", **out)
    print("wrote ref_direct_db1.npz")


def reference_variants(tmp):
    from _support import run_reference, VARIANTS, variant_case
    out = {}
    for name in VARIANTS:
        states, solves, _, _ = run_reference(variant_case(name), nrun=2, dump_every=1, threads=1,
                                             workdir=tmp / f"variant_{name}")
        for step in (0, 1, 2):
            out[f"{name}_s{step}_nempty"] = states[step]["nempty"].astype(np.int32)
        for step in (0, 2):
            for n in ("SF", "U", "W", "T"):
                out[f"{name}_s{step}_{n}"] = _node(states[step], n)
        out[f"{name}_solves_tq"], _ = _solves(solves)
    np.savez_compressed(HERE / "ref_variants.npz", **out)
    print("wrote ref_variants.npz")


def reference_rk1(tmp):
    from _support import run_reference, RK1_VARIANTS, variant_case
    out = {}
    for name in RK1_VARIANTS:
        states, solves, _, _ = run_reference(variant_case(name), nrun=2, dump_every=1, threads=1,
                                             workdir=tmp / f"variant_{name}")
        for step in (0, 1, 2):
            out[f"{name}_s{step}_nempty"] = states[step]["nempty"].astype(np.int32)
        for step in (0, 2):
            for n in ("SF", "U", "W", "T"):
                out[f"{name}_s{step}_{n}"] = _node(states[step], n)
        out[f"{name}_solves_tq"], _ = _solves(solves)
    np.savez_compressed(HERE / "ref_rk1.npz", **out)
    print("wrote ref_rk1.npz")


def reference_config1(tmp, tolerance=1e-8, name="ref_config1_100.npz"):
    from projectracer_amd import cases
    from _support import run_reference
    iso = cases.isoviscous(nx=64, nz=64, tpc=20, tolerance=tolerance)
    states, solves, _, _ = run_reference(iso, nrun=100, dump_every=100, threads=1,
                                         workdir=tmp / f"config1_{tolerance:g}")
    st = states[100]
    out = {f"s100_{n}": _node(st, n) for n in ("SF", "U", "W", "T")}
    out["s100_nempty"] = st["nempty"].astype(np.int32)
    nx, nz = st["nx"], st["nz"]
    dr, ds = iso.params.aspect / nx, 1.0 / nz
    # the tracer's cell: the node nearest to it (nint, PS/composition.f90:1005-1006)
    i = np.floor(np.abs(st["rtr"] / dr) + 0.5) * np.sign(st["rtr"])
    k = np.floor(np.abs(st["str"] / ds) + 0.5) * np.sign(st["str"])
    out["s100_cell"] = (k * (nx + 1) + i).astype(np.int32)
    sel = np.arange(0, st["ntr"], 29)
    out["s100_sel"] = sel
    for n in ("rtr", "str", "Ttr"):
        out[f"s100_{n}"] = st[n][sel].copy()
    out["time"] = np.float64(st["time"])
    out["solves_tq"], out["solves_res"] = _solves(solves)
    out["tolerance"] = np.float64(tolerance)
    np.savez_compressed(HERE / name, **out)
    print("wrote", name, len(solves), "solves")


# configurations the reference solves at their own full grid size, with a
# fixed OpenMP thread count (8): the summation order of its reductions
# depends on it at the 1e-15 level
FIX_THREADS = 8


def _sub(st, name, step):
    """Node values (0:nx,0:nz) of a field at every `step`-th node in each
    direction, plus the field's sum and sum of squares over all nodes."""
    f = _node(st, name)
    return f[::step, ::step].copy(), np.array([f.sum(), (f * f).sum()])


def reference_config2(tmp):
    """The second configuration at its full size: lambda=3, 384x128, Ra_T=1e6,
    Ra_C=5e5, eta=exp(-ln(1e3)T+ln(10)C), tpc=40 (1,986,600 tracers),
    tolerance 1e-8: the state after the first solve and its cycle count.
    The reference's W-cycle converges slowly on this law (its first solve is
    still at an L1 ratio of 4.5e-7 after 1000 cycles), so max_cycles is
    raised to 5000; within the first time step one of its solves then stops
    the program, so only the first solve is kept."""
    from projectracer_amd import cases
    from _support import run_reference
    th = cases.thermochem(nx=384, nz=128, tpc=40, max_cycles=5000)
    states, solves, _, _ = run_reference(th, nrun=0, dump_every=1, threads=FIX_THREADS,
                                         workdir=tmp / "config2", timeout=7200)
    out = {}
    for n in ("SF", "T", "CNEW", "VIS", "U", "W"):
        out[f"s0_{n}"] = _node(states[0], n)
    out["s0_nempty"] = states[0]["nempty"].astype(np.int32)
    out["solves_tq"], out["solves_res"] = _solves(solves)
    np.savez_compressed(HERE / "ref_config2_384x128.npz", **out)
    print("wrote ref_config2_384x128.npz", [(t, q) for t, q, _, _ in solves])


# the strongest T- and C-dependent viscosity law the reference's multigrid
# solves at 1024x256 (the third configuration's own law diverges there):
# eta = exp(-ln(30) T + ln(2) C), filter thresholds delta_visT = delta_visC =
# 1.2, courant 0.25, L1 relative residual tolerance 1e-7
LAW1024 = dict(visT=30.0, visC=0.5, delta_visT=1.2, delta_visC=1.2, courant=0.25, tolerance=1e-7)


def reference_law1024(tmp, tpc=4):
    """The 1024x256 thermochemical state (lambda=4, Ra_T=1e7, Ra_C=3e6, layer
    at z=0.3, seed 2) with LAW1024 and tpc=4 (1,053,700 tracers): states after
    the first solve and one time step, node fields kept at every 4th node
    with whole-field sums; node counts whole."""
    from projectracer_amd import cases
    from _support import run_reference
    case = cases.thermochem_large(tpc=tpc, **LAW1024)
    states, solves, _, _ = run_reference(case, nrun=1, dump_every=1, threads=FIX_THREADS,
                                         workdir=tmp / "law1024", timeout=7200)
    out = {}
    for step in (0, 1):
        for n in ("SF", "T", "CNEW", "VIS", "U"):
            out[f"s{step}_{n}"], out[f"s{step}_{n}_sums"] = _sub(states[step], n, 4)
        out[f"s{step}_nempty"] = states[step]["nempty"].astype(np.int16)
        out[f"s{step}_dt"] = np.float64(states[step]["dt"])
    out["solves_tq"], out["solves_res"] = _solves(solves)
    np.savez_compressed(HERE / "ref_law1024.npz", **out)
    print("wrote ref_law1024.npz", [(t, q) for t, q, _, _ in solves])


def reference_law1024_order(tmp, tpc=4):
    """How much the reference's own first solve on the LAW1024 state depends
    on the order of its sums: the same run with 4 OpenMP threads instead of 8
    (psi at every 4th node), and the oracle's single-threaded C restatement
    (psi at every 4th node and the whole-field sums).  Both differ from the
    8-thread run by a few 1e-10 at this size."""
    from projectracer_amd import cases
    from _support import run_reference, OracleSim, interior
    case = cases.thermochem_large(tpc=tpc, **LAW1024)
    states, solves, _, _ = run_reference(case, nrun=0, dump_every=1, threads=4, workdir=tmp / "law1024_t4",
                                         timeout=7200)
    out = {}
    out["t4_s0_SF"], out["t4_s0_SF_sums"] = _sub(states[0], "SF", 4)
    out["t4_solves_tq"], _ = _solves(solves)
    o = OracleSim(case.params)
    o.load_state(case)
    o.multigrid()
    o.compute_velocities()
    for n in ("SF", "U"):
        f = interior(o.field(n), 3, o.nx, o.nz)
        out[f"oracle_s0_{n}"] = f[::4, ::4].copy()
        out[f"oracle_s0_{n}_sums"] = np.array([f.sum(), (f * f).sum()])
    out["oracle_q"] = np.int32(o.info()["q"])
    np.savez_compressed(HERE / "ref_law1024_order.npz", **out)
    print("wrote ref_law1024_order.npz", [(t, q) for t, q, _, _ in solves], "oracle q", o.info()["q"])


def reference_divergence(tmp, tpc=4):
    """The third configuration's own law at 1024x256 (eta = exp(-ln(1e5) T +
    ln(30) C), delta_visT = delta_visC = 2), tpc=4: the reference stops in
    its first solve with 'Abnormally High Residual'.  Kept: the cycle it stops
    at and the per-cycle level-1 residuals it printed."""
    from projectracer_amd import cases
    from _support import run_reference
    case = cases.thermochem_large(tpc=tpc, tolerance=1e-8)
    _, _, stdout, _ = run_reference(case, nrun=1, dump_every=-1, threads=FIX_THREADS, workdir=tmp / "divergence")
    cyc = []
    for line in stdout.splitlines():
        parts = line.split()
        if len(parts) == 4:
            try:
                q = int(parts[0])
                float(parts[1].replace("D", "E"))
            except ValueError:
                continue
            if "." in parts[1] or "E" in parts[1]:
                cyc.append((q, float(parts[1]), float(parts[2])))
    stopped = "Abnormally High Residual" in stdout
    np.savez_compressed(HERE / "ref_divergence1024.npz", cycles=np.array(cyc), stopped=np.bool_(stopped))
    print("wrote ref_divergence1024.npz", len(cyc), "cycles, stopped", stopped)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] in ("config2", "law1024", "law1024_order", "divergence"):
        fn = {"config2": reference_config2, "law1024": reference_law1024, "law1024_order": reference_law1024_order,
              "divergence": reference_divergence}
        fn[sys.argv[1]](Path(tempfile.mkdtemp(prefix="golden_")))
        raise SystemExit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "direct":
        reference_direct(Path(tempfile.mkdtemp(prefix="golden_")))
        raise SystemExit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "variants":
        reference_variants(Path(tempfile.mkdtemp(prefix="golden_")))
        raise SystemExit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "rk1":
        reference_rk1(Path(tempfile.mkdtemp(prefix="golden_")))
        raise SystemExit(0)
    if len(sys.argv) > 1 and sys.argv[1] == "equalize":
        reference_equalize(Path(tempfile.mkdtemp(prefix="golden_")))
        raise SystemExit(0)
    if REF_DIR.exists():
        shipped_snapshot()
        shipped_restart()
        shipped_snapshot_text()
    reference_runs()
