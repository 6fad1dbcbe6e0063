"""The T- and C-dependent viscosity law of the 1024x256 measurement,
eta = exp(-ln(30) T + ln(2) C) with filter thresholds delta_visT =
delta_visC = 1.2, on the same thermochemical state (lambda=4, Ra_T=1e7,
Ra_C=3e6, layer at z=0.3, seed 2) reduced to 256x64 with 8 tracers per node
and courant 0.02 (the time step scales with the squared spacing, so this
keeps the tracers' steps per cell at the full size's courant 0.25 order).

  * The projection's filter loops (smoother_time_T, smoother_time,
    PS/smoother.f90:582-672) run on the device with a device-side stop flag;
    against the same loops with a host round trip per check (PT_FILTER_HOST=1)
    the pass counts and filtered fields are bit-identical over three steps.
  * The FP64 path against the oracle over three steps: the cycle count of
    every solve group, the filter pass counts n_comp / n_temp and the
    viscosity filter count n_eta exact, node counts exact, fields within 1e-10.
  * The mixed path (ivis=2, its FP32 V-cycle) on the same state: the FP64
    stop test met, psi within 1e-5 of the oracle's FP64 solve (both stop at
    an L1 relative residual of 1e-7).
"""
import os

import numpy as np
import pytest

from projectracer_amd import Simulation, cases, _abi
from _support import OracleSim, rel_err, interior

pytestmark = pytest.mark.gpu

LAW = dict(visT=30.0, visC=0.5, delta_visT=1.2, delta_visC=1.2, courant=0.02, tolerance=1e-7)


def _case():
    return cases.thermochem_large(nx=256, nz=64, tpc=8, **LAW)


def _node(sim, name):
    h = _abi.FIELD_HALO[_abi.FIELD[name]]
    return interior(sim.field(name), h, sim.nx, sim.nz)


def _filters_run(host_loop):
    if host_loop:
        os.environ["PT_FILTER_HOST"] = "1"
    try:
        case = _case()
        sim = Simulation(case.params)
    finally:
        os.environ.pop("PT_FILTER_HOST", None)
    sim.load_state(case)
    out = []
    for step in range(3):
        sim.time_step()
        inf = sim.info()
        out.append(dict(n_comp=inf["n_comp"], n_temp=inf["n_temp"],
                        **{n: sim.field(n).copy() for n in ("CVIS", "TVIS", "CBUOY", "TBUOY", "SF")}))
    sim.close()
    return out


def test_device_filter_loops_match_host_loops():
    print("[law] 256x64: filter loops on the device vs host round trips", flush=True)
    dev = _filters_run(False)
    host = _filters_run(True)
    for step, (d, h) in enumerate(zip(dev, host)):
        print(f"[law] step {step}: n_comp {d['n_comp']} / {h['n_comp']}, n_temp {d['n_temp']} / {h['n_temp']}",
              flush=True)
        assert d["n_comp"] == h["n_comp"] and d["n_temp"] == h["n_temp"], step
        for n in ("CVIS", "TVIS", "CBUOY", "TBUOY", "SF"):
            assert np.array_equal(d[n], h[n]), (step, n)
    assert max(d["n_comp"] for d in dev) > 20


def test_fp64_law_three_steps_vs_oracle():
    print("[law] 256x64 FP64 vs oracle, three steps", flush=True)
    case = _case()
    g = Simulation(case.params)
    o = OracleSim(case.params)
    g.load_state(case)
    o.load_state(case)
    for step in range(3):
        cg, co = g.info()["cycles"], o.info()["cycles"]
        g.time_step()
        o.time_step()
        gi, oi = g.info(), o.info()
        errs = {n: rel_err(_node(g, n), _node(o, n)) for n in ("SF", "T", "U", "W", "CNEW", "CVIS", "TVIS", "VIS")}
        print(f"[law] step {step}: cycles {gi['cycles'] - cg} / {oi['cycles'] - co}, n_comp {gi['n_comp']} / "
              f"{oi['n_comp']}, n_temp {gi['n_temp']} / {oi['n_temp']}, n_eta {gi['n_eta']} / {oi['n_eta']}, "
              + " ".join(f"{n}={e:.1e}" for n, e in errs.items()), flush=True)
        assert gi["cycles"] - cg == oi["cycles"] - co, step
        assert (gi["n_comp"], gi["n_temp"], gi["n_eta"]) == (oi["n_comp"], oi["n_temp"], oi["n_eta"]), step
        assert np.array_equal(g.nempty(), o.nempty()), step
        assert max(errs.values()) <= 1e-10, (step, errs)


def test_mixed_law_solve_vs_oracle():
    print("[law] 256x64 mixed (FP32 V-cycle) first solve vs the oracle's FP64 solve", flush=True)
    case = _case()
    o = OracleSim(case.params)
    o.load_state(case)
    o.multigrid()
    p = _case().params
    p.precision = _abi.PT_PREC_MIXED
    g = Simulation(p)
    g.load_state(case)
    g.multigrid()
    gi, oi = g.info(), o.info()
    err = rel_err(_node(g, "SF"), _node(o, "SF"))
    print(f"[law] mixed q={gi['q']} (FP64 oracle q={oi['q']}), L1 ratio {gi['residual'] / gi['res_ref']:.2e}, "
          f"psi error vs the FP64 oracle {err:.2e}", flush=True)
    assert gi["residual"] / gi["res_ref"] <= p.tolerance
    assert err <= 1e-5
