"""The learned initial guess inside the solve, on the learned-guess
configuration reduced for a test: 256x128 (lambda=2) thermochemical cases
with Ra_T log-uniform in [1e5, 1e7], buoyancy ratio uniform in [0, 1], layer
depth uniform in [0.1, 0.4], eta = exp(-gamma T) with gamma uniform in
[0, ln(1e2)], 4 tracers per node, the mixed path at an L1 tolerance of 1e-7.

Training pairs come from the method itself: three cases run three steps
with capture on (every solve stores its inputs and its correction on the
device); the network then takes 200 AdamW steps with the hand-written
kernels.  Two held-out cases then run the same two warm-up steps twice and
two more steps with and without the network attached.  The plain solver
judges every proposal (kept only when its FP64 L1 residual is below the
warm start's), so every solve still meets the stop test; the test reports
what the network saves and bounds what it may cost."""
import numpy as np
import pytest

from projectracer_amd import Simulation, cases, _abi
from projectracer_amd.learned import GuessNet, attach, guess_counts

pytestmark = pytest.mark.gpu


def case5(seed, tpc=4):
    rng = np.random.default_rng(seed)
    RaT = 10 ** rng.uniform(5, 7)
    B, layer, gam = rng.uniform(0, 1), rng.uniform(0.1, 0.4), rng.uniform(0, np.log(1e2))
    c = cases.thermochem(nx=256, nz=128, tpc=tpc, aspect=2.0, RaT=RaT, RaC=B * RaT, visT=float(np.exp(gam)),
                         visC=1.0, layer=layer, seed=seed, ivis=1, courant=0.05, tolerance=1e-7)
    c.params.precision = _abi.PT_PREC_MIXED
    return c


def _steps(sim, n):
    for _ in range(n):
        sim.energy_time()
        sim.multigrid()
        sim.compute_velocities()


def test_learned_guess_trained_on_solver_snapshots():
    print("[learned] capture 3 cases x 3 steps at 256x128, train, compare on 2 held-out cases", flush=True)
    net = GuessNet(seed=0)
    net.dataset(128, 256, 64)
    for s in range(3):
        c = case5(100 + s)
        sim = Simulation(c.params)
        sim.load_state(c)
        sim.multigrid()
        sim.compute_velocities()
        attach(sim, net, guess=False, capture=True)
        _steps(sim, 3)
        sim.synchronize()
        sim.close()
    assert net.dataset_size() == 36
    first = net.train(1, 8, lr=1e-3, seed=0)
    last = net.train(199, 8, lr=1e-3, seed=1)
    print(f"[learned] 36 pairs; loss {first:.4f} -> {last:.4f}; output scale {net.out_scale():.3e}", flush=True)
    assert last < first
    total = {"warm": 0, "learned": 0}
    for s in range(2):
        c = case5(200 + s)
        for mode in ("warm", "learned"):
            sim = Simulation(c.params)
            sim.load_state(c)
            sim.multigrid()
            sim.compute_velocities()
            _steps(sim, 2)
            if mode == "learned":
                attach(sim, net, guess=True)
            c0, s0 = sim.info()["cycles"], sim.info()["solves"]
            _steps(sim, 2)
            inf = sim.info()
            assert inf["residual"] / inf["res_ref"] <= c.params.tolerance
            total[mode] += inf["cycles"] - c0
            if mode == "learned":
                tried, used = guess_counts(sim)
                assert tried == inf["solves"] - s0 == 8
                print(f"[learned] case {200 + s}: proposals kept {used} of {tried}", flush=True)
            sim.close()
    saved = 1.0 - total["learned"] / total["warm"]
    print(f"[learned] cycles over 16 held-out solves: warm start {total['warm']}, learned guess "
          f"{total['learned']} (saved {100 * saved:.1f}%)", flush=True)
    # a kept proposal starts from a lower residual than the warm start; the
    # cycle count may still differ by about one per solve either way
    assert total["learned"] <= total["warm"] + 16
