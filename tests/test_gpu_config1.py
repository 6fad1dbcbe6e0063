"""The 64x64 isoviscous configuration end to end: lambda=1, Ra_T=1e5,
T = (1-z) + 0.01 cos(pi x) sin(pi z), 20 tracers per node (84,500), RK4 with
courant 0.5, 100 time steps, against the flang-built reference program run
on the same injected state (tests/golden/ref_config1_100.npz).

FP64 path: the multigrid cycle counts of every solve group (the main solve
and the three energy_time solves of each step; 401 solves), the time, the
per-node tracer counts and every tracer's cell are exact; psi, u, w and T
after 100 steps agree to 1e-10 (max-norm relative), the tracer positions and
temperatures of a subsample to 1e-10.  The mixed path reaches the same
residual target and stays within 1e-6 of the reference."""
import numpy as np
import pytest

from projectracer_amd import Simulation, cases, _abi
from _support import GOLDEN, interior, rel_err

pytestmark = pytest.mark.gpu


def _node(sim, name):
    h = _abi.FIELD_HALO[_abi.FIELD[name]]
    return interior(sim.field(name), h, sim.nx, sim.nz).copy()


def _run(precision):
    """The reference loop (PS/main.f90:245-299) 100 times, then the solve of
    the final state (as the reference harness dumps it).  Returns the
    simulation, the cycle counts per solve group in the reference's order
    [main solve of step t, the three energy_time solves of step t], and the
    time advanced by the steps' dt."""
    case = cases.isoviscous(nx=64, nz=64, tpc=20)
    case.params.precision = precision
    sim = Simulation(case.params)
    sim.load_state(case)
    groups, time = [], 0.0
    for step in range(100):
        c0 = sim.info()["cycles"]
        sim.multigrid()
        sim.compute_velocities()
        c1 = sim.info()["cycles"]
        sim.energy_time()
        inf = sim.info()
        groups += [c1 - c0, inf["cycles"] - c1]
        time = time + inf["dt"]
    sim.multigrid()
    sim.compute_velocities()
    return sim, groups, time


def _ref_groups(g):
    q = g["solves_tq"][:, 1]
    out = []
    for t in range(100):
        out.append(int(q[4 * t]))
        out.append(int(q[4 * t + 1:4 * t + 4].sum()))
    return out


def test_config1_100_steps_fp64():
    print("[config1] 64x64 isoviscous, 100 RK4 steps, FP64", flush=True)
    g = np.load(GOLDEN / "ref_config1_100.npz")
    sim, groups, time = _run(_abi.PT_PREC_FP64)
    ref = _ref_groups(g)
    diff = [(j // 2, "main" if j % 2 == 0 else "stages", groups[j], ref[j]) for j in range(200) if groups[j] != ref[j]]
    ratio = g["solves_res"][:, 0] / g["solves_res"][:, 1]
    print(f"[config1] solve groups differing (step, group, gpu, reference): {diff}", flush=True)
    print(f"[config1] reference final L1 ratios in those steps: "
          f"{[float(x) for d in diff for x in ratio[4 * d[0]:4 * d[0] + 4]]}", flush=True)
    inf = sim.info()
    ref_cycles = int(g["solves_tq"][:, 1].sum())
    errs = {n: rel_err(_node(sim, n), g[f"s100_{n}"]) for n in ("SF", "U", "W", "T")}
    print(f"[config1] solves {inf['solves']} cycles {inf['cycles']} (reference {len(g['solves_tq'])}, {ref_cycles}); "
          f"time {time!r} vs {float(g['time'])!r}; field errors {errs}", flush=True)
    assert inf["solves"] == len(g["solves_tq"]) == 401
    assert not diff and inf["cycles"] == ref_cycles
    assert inf["q"] == int(g["solves_tq"][-1, 1])
    assert time == float(g["time"])
    assert max(errs.values()) <= 1e-10
    assert np.array_equal(sim.nempty(), g["s100_nempty"])
    r, s, T, _ = sim.tracers()
    nx, nz = sim.nx, sim.nz
    cell = (np.floor(s * nz + 0.5) * (nx + 1) + np.floor(r * nx / 1.0 + 0.5)).astype(np.int32)
    assert np.array_equal(cell, g["s100_cell"])
    sel = g["s100_sel"]
    assert rel_err(r[sel], g["s100_rtr"]) <= 1e-10 and rel_err(s[sel], g["s100_str"]) <= 1e-10
    assert rel_err(T[sel], g["s100_Ttr"]) <= 1e-10


def test_config1_100_steps_mixed():
    print("[config1] 64x64 isoviscous, 100 RK4 steps, mixed precision", flush=True)
    g = np.load(GOLDEN / "ref_config1_100.npz")
    sim, _, _ = _run(_abi.PT_PREC_MIXED)
    inf = sim.info()
    errs = {n: rel_err(_node(sim, n), g[f"s100_{n}"]) for n in ("SF", "U", "W", "T")}
    print(f"[config1] mixed: cycles {inf['cycles']} (reference {int(g['solves_tq'][:, 1].sum())}); "
          f"errors vs the FP64 reference {errs}", flush=True)
    assert inf["residual"] / inf["res_ref"] <= 1e-8
    assert max(errs.values()) <= 1e-6
