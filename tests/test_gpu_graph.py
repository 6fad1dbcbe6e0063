"""The multigrid cycle replayed as a captured HIP graph against the same
cycle launched kernel by kernel (PT_GRAPH=0 at context creation): several
steps, so the graph is replayed across cycles and solves with other work
between the replays; every cycle count and field must agree bit for bit, in
FP64 and in the mixed path (whose strip sweeps rotate their buffers, so a
cycle may need one graph per starting arrangement).  With a T- and
C-dependent viscosity every solve rebuilds the viscosity fields and the
coarse operator in place, and the replayed graphs must read the new ones."""
import os

import numpy as np
import pytest

from projectracer_amd import Simulation, cases, _abi
from _support import interior

pytestmark = pytest.mark.gpu


def _node(sim, name):
    h = _abi.FIELD_HALO[_abi.FIELD[name]]
    return interior(sim.field(name), h, sim.nx, sim.nz).copy()


def _run(case, graph):
    old = os.environ.get("PT_GRAPH")
    os.environ["PT_GRAPH"] = "1" if graph else "0"
    try:
        sim = Simulation(case.params)
    finally:
        if old is None:
            del os.environ["PT_GRAPH"]
        else:
            os.environ["PT_GRAPH"] = old
    sim.load_state(case)
    cycles = []
    sim.multigrid()
    sim.compute_velocities()
    for _ in range(3):
        sim.energy_time()
        sim.multigrid()
        sim.compute_velocities()
        cycles.append(sim.info()["cycles"])
    out = dict(cycles=cycles, sf=_node(sim, "SF"), u=_node(sim, "U"), t=_node(sim, "T"), ne=sim.nempty().copy(),
               graphs=sim.info()["cycle_graphs"])
    sim.close()
    return out


@pytest.mark.parametrize("precision,nx,nz", [(_abi.PT_PREC_FP64, 512, 128), (_abi.PT_PREC_MIXED, 512, 128),
                                             (_abi.PT_PREC_MIXED, 1024, 256)],
                         ids=["fp64-512x128", "mixed-512x128", "mixed-1024x256"])
def test_graph_replay_matches_eager_launches(precision, nx, nz):
    print(f"[graph] {nx}x{nz} isoviscous, precision {precision}", flush=True)
    case = cases.isoviscous(nx=nx, nz=nz, tpc=1, aspect=4.0, tolerance=1e-7)
    case.params.precision = precision
    a = _run(case, True)
    b = _run(case, False)
    print(f"[graph] cycles graph={a['cycles']} eager={b['cycles']}; captured cycle graphs {a['graphs']}", flush=True)
    assert a["graphs"] >= 1 and b["graphs"] == 0
    assert a["cycles"] == b["cycles"]
    for key in ("sf", "u", "t", "ne"):
        assert np.array_equal(a[key], b[key]), key


@pytest.mark.parametrize("precision", [_abi.PT_PREC_FP64, _abi.PT_PREC_MIXED], ids=["fp64", "mixed"])
def test_graph_replay_variable_viscosity(precision):
    print(f"[graph] 256x64 eta(T,C), precision {precision}", flush=True)
    case = cases.thermochem_large(nx=256, nz=64, tpc=4, visT=30.0, visC=0.5, delta_visT=1.2, delta_visC=1.2,
                                  courant=0.02, tolerance=1e-7)
    case.params.precision = precision
    a = _run(case, True)
    b = _run(case, False)
    print(f"[graph] cycles graph={a['cycles']} eager={b['cycles']}; captured cycle graphs {a['graphs']}", flush=True)
    assert a["graphs"] >= 1 and b["graphs"] == 0
    assert a["cycles"] == b["cycles"]
    for key in ("sf", "u", "t", "ne"):
        assert np.array_equal(a[key], b[key]), key
