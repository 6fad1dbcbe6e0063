"""GPU edge cases of the tracer pipeline and size-independent properties at
the full 1024x256 size, through the C ABI.

Edge cases against the oracle (bit-exact nempty, fields within 1e-13):
tracers exactly on the walls and on half-cell ties (nint rounds half away
from zero), nodes left without tracers (they keep the previous Eulerian
value), and a tracer outside the domain (the reference stops with
"Tracer advected outside of domain"; here both return the same code).
At full size: tracer conservation, domain containment, determinism of two
identical runs, and the mixed path's agreement with the FP64 path.
"""
import numpy as np
import pytest

from projectracer_amd import Simulation, PtError, cases, _abi
from _support import OracleSim, rel_err, interior

pytestmark = pytest.mark.gpu


def _node(sim, name):
    h = _abi.FIELD_HALO[_abi.FIELD[name]]
    return interior(sim.field(name), h, sim.nx, sim.nz)


def _with_tracers(base, r, s, Ttr=None, Ctr=None):
    n = r.size
    return cases.Case(base.name, base.params, base.T_nodes, base.Tbot_iso, base.Ttop_iso,
                      np.ascontiguousarray(r), np.ascontiguousarray(s),
                      np.ascontiguousarray(base.Ttr[:n] if Ttr is None else Ttr),
                      np.ascontiguousarray(base.Ctr[:n] if Ctr is None else Ctr))


def test_walls_ties_and_empty_nodes_match_oracle():
    base = cases.isoviscous(nx=32, nz=32, tpc=4)
    nx, nz = base.params.nx, base.params.nz
    dr, ds = 1.0 / nx, 1.0 / nz
    r, s = base.rtr.copy(), base.str_.copy()
    rng = np.random.default_rng(4)
    # a block of nodes without tracers: move their tracers to other nodes
    hole = (r > 10.5 * dr) & (r < 14.5 * dr) & (s > 20.5 * ds) & (s < 23.5 * ds)
    r[hole] = rng.uniform(0.6, 0.9, hole.sum())
    s[hole] = rng.uniform(0.1, 0.4, hole.sum())
    # exact walls and half-cell ties
    m = 200
    r[:m] = np.where(np.arange(m) % 2 == 0, 0.0, 1.0)
    s[m:2 * m] = np.where(np.arange(m) % 2 == 0, 0.0, 1.0)
    ties = np.arange(2 * m, 3 * m)
    r[ties] = (rng.integers(0, nx, m) + 0.5) * dr
    s[ties] = (rng.integers(0, nz, m) + 0.5) * ds
    T = rng.uniform(0.0, 1.0, r.size)
    C = (rng.uniform(0.0, 1.0, r.size) < 0.3).astype(np.float64)
    case = _with_tracers(base, r, s, T, C)
    g = Simulation(case.params)
    o = OracleSim(case.params)
    g.load_state(case)
    o.load_state(case)
    ne = g.nempty()
    assert (ne == 0).sum() >= 12, "the hole should leave empty nodes"
    assert np.array_equal(ne, o.nempty())
    for name in ("T", "CNEW", "TVIS", "CVIS"):
        assert rel_err(_node(g, name), _node(o, name)) <= 1e-13, name
    # one step through the solve and RK4 with the ragged cloud
    g.time_step()
    o.time_step()
    assert g.info()["q"] == o.info()["q"]
    assert np.array_equal(g.nempty(), o.nempty())
    for name in ("SF", "T", "CNEW"):
        assert rel_err(_node(g, name), _node(o, name)) <= 1e-10, name


def test_tracer_outside_domain_is_reported():
    base = cases.isoviscous(nx=16, nz=16, tpc=2)
    r, s = base.rtr.copy(), base.str_.copy()
    r[7] = 1.0 + 1e-9
    case = _with_tracers(base, r, s)
    g = Simulation(case.params)
    o = OracleSim(case.params)
    with pytest.raises(PtError) as eg:
        g.load_state(case)
    with pytest.raises(PtError) as eo:
        o.load_state(case)
    assert eg.value.code == eo.value.code == _abi.PT_ERR_TRACER_OUTSIDE


@pytest.fixture(scope="module")
def big():
    case = cases.thermochem_large(ivis=0, tolerance=1e-7)
    out = {}
    for prec in (_abi.PT_PREC_FP64, _abi.PT_PREC_MIXED):
        p = cases.thermochem_large(ivis=0, tolerance=1e-7).params
        p.precision = prec
        sim = Simulation(p)
        sim.load_state(case)
        sim.time_step()
        sim.multigrid()
        out[prec] = dict(sf=_node(sim, "SF").copy(), nempty=sim.nempty().copy(), tr=sim.tracers(),
                         info=sim.info(), T=_node(sim, "T").copy())
        sim.close()
    out["ntr"] = case.ntr
    return out


def test_full_size_conservation_and_containment(big):
    for prec in (_abi.PT_PREC_FP64, _abi.PT_PREC_MIXED):
        d = big[prec]
        assert d["nempty"].sum() == big["ntr"]
        r, s, T, C = d["tr"]
        assert np.all((r >= 0.0) & (r <= 4.0) & (s >= 0.0) & (s <= 1.0))
        assert np.all(np.isfinite(T)) and np.all((C == 0.0) | (C == 1.0))
        inf = d["info"]
        assert inf["residual"] / inf["res_ref"] <= 1e-7


def test_full_size_mixed_agrees_with_fp64(big):
    f, m = big[_abi.PT_PREC_FP64], big[_abi.PT_PREC_MIXED]
    err = rel_err(m["sf"], f["sf"])
    print(f"1024x256 one step: psi(mixed) vs psi(FP64) = {err:.3e}; nempty differ at "
          f"{int((m['nempty'] != f['nempty']).sum())} nodes")
    assert err <= 1e-5
    assert rel_err(m["T"], f["T"]) <= 1e-6


def test_bench_workload_determinism():
    """The measured workload itself (1024x256, tpc=50, eta = exp(-ln(30) T +
    ln(2) C), mixed path): two runs of one step from the same state agree
    bit for bit (fields, node counts, every tracer)."""
    case = cases.thermochem_large(visT=30.0, visC=0.5, delta_visT=1.2, delta_visC=1.2, courant=0.25,
                                  tolerance=1e-7)
    case.params.precision = _abi.PT_PREC_MIXED
    print(f"[determinism] {case.name} {case.ntr} tracers, mixed, one step twice", flush=True)
    runs = []
    for _ in range(2):
        sim = Simulation(case.params)
        sim.load_state(case)
        sim.time_step()
        runs.append((_node(sim, "SF").copy(), _node(sim, "T").copy(), sim.nempty().copy(), sim.tracers()))
        sim.close()
    (a_sf, a_t, a_ne, a_tr), (b_sf, b_t, b_ne, b_tr) = runs
    assert np.array_equal(a_sf, b_sf) and np.array_equal(a_t, b_t) and np.array_equal(a_ne, b_ne)
    for x, y in zip(a_tr, b_tr):
        assert np.array_equal(x, y)
