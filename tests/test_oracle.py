"""Pinning the CPU oracle (oracle/liboracle.so) before the HIP path is held
to it.  CPU only.

Each test names what fixes the expected values:
  * the reference's shipped output files (stdout, stats.dat, T10000 of the
    shipped 100x100, tpc=60, 2-thread run);
  * the reference program compiled from its own Fortran sources
    (tests/golden/ref_*.npz, written by tests/golden/make_golden.py);
  * closed forms: textbook Fornberg stencils, the damped Chebyshev stability
    polynomial of RKC1, the operator's symbol on sine modes, exactness of the
    Taylor interpolation on bicubics, the partition of unity of bilinear
    shape functions, LAPACK's band LU (scipy).
"""
import math

import numpy as np
import pytest
from numpy.polynomial import chebyshev as npcheb

from projectracer_amd import _abi, cases
from _support import GOLDEN, OracleSim, oracle_lib, rel_err, interior

C = pytest.importorskip("ctypes")


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _node(sim, name):
    h = _abi.FIELD_HALO[_abi.FIELD[name]]
    return interior(sim.field(name), h, sim.nx, sim.nz)


def _sim(**over):
    p = _abi.default_params(**{"aspect": 1.0, "nx": 32, "nz": 32, "tpc": 4, "grids": 3, **over})
    return OracleSim(p)


# ---------------------------------------------------------------------------
# finite-difference tables (PS/main.f90:499-642)
# ---------------------------------------------------------------------------
TEXTBOOK = {
    "D1": [0, 1 / 12, -2 / 3, 0, 2 / 3, -1 / 12, 0],
    "D2": [0, -1 / 12, 4 / 3, -5 / 2, 4 / 3, -1 / 12, 0],
    "D3": [1 / 8, -1, 13 / 8, 0, -13 / 8, 1, -1 / 8],
    "D4": [-1 / 6, 2, -13 / 2, 28 / 3, -13 / 2, 2, -1 / 6],
}


def _fd_tables(sim):
    D = [np.zeros(7) for _ in range(4)]
    Dtab = np.zeros(28)
    oracle_lib().pto_fd_tables(sim.ctx, *[_dp(a) for a in D], _dp(Dtab))
    return dict(zip(("D1", "D2", "D3", "D4"), D)), Dtab.reshape(4, 7)


def test_fd_stencils_are_fornberg_weights():
    tabs, _ = _fd_tables(_sim())
    for name, ref in TEXTBOOK.items():
        np.testing.assert_allclose(tabs[name], ref, rtol=0, atol=4e-15, err_msg=name)


def test_fd_tables_match_reference_build():
    g = np.load(GOLDEN / "ref_iso64.npz")
    tabs, Dtab = _fd_tables(_sim(nx=64, nz=64, grids=5))
    for name in ("D1", "D2", "D3", "D4"):
        np.testing.assert_array_equal(tabs[name], g[f"init_{name}"], err_msg=name)
    np.testing.assert_array_equal(Dtab, g["init_D"])


# ---------------------------------------------------------------------------
# RKC1 coefficients (PS/Chebyshev.f90:82-110)
# ---------------------------------------------------------------------------
def _rkc(sim, s=35):
    arr = [np.zeros(s + 1) for _ in range(4)]
    st = C.c_double()
    oracle_lib().pto_rkc_coefficients(sim.ctx, *[_dp(a) for a in arr], C.byref(st))
    return arr, st.value


def test_rkc1_stability_matches_shipped_stdout():
    # stdout:10 of the shipped run: RKC1_stability=  -83.852981245559235
    _, stab = _rkc(_sim())
    assert stab == pytest.approx(-83.852981245559235, rel=2e-16, abs=0)


def test_rkc1_coefficients_match_reference_build():
    g = np.load(GOLDEN / "ref_iso64.npz")
    (mewt, mew, v, cc), stab = _rkc(_sim(nx=64, nz=64, grids=5))
    np.testing.assert_array_equal(mewt[1:], g["init_mewt"])
    np.testing.assert_array_equal(mew[2:], g["init_mew"])
    np.testing.assert_array_equal(v[2:], g["init_v"])
    np.testing.assert_array_equal(cc[2:], g["init_cc"])
    assert stab == g["init_stability"]


def test_rkc1_recursion_is_damped_chebyshev_polynomial():
    """On y' = lambda y the stage recursion of iterate_stream
    (PS/iterative_solver.f90:1561-1597) must give R_s(z) =
    T_s(w0 + w1 z)/T_s(w0), w0 = 1 + damping/s^2, w1 = T_s(w0)/T_s'(w0)
    (first-order damped RKC, van der Houwen & Sommeijer 1980)."""
    s, damping = 35, 1000.0
    (mewt, mew, v, cc), stab = _rkc(_sim())
    w0 = 1.0 + damping / s**2
    Ts = npcheb.Chebyshev.basis(s)
    w1 = Ts(w0) / Ts.deriv()(w0)
    assert stab == pytest.approx(-2.0 * w0 / w1, rel=1e-12)
    for z in np.linspace(stab, 0.0, 41):
        r_prev2, r_prev = 1.0, 1.0 + mewt[1] * z
        for j in range(2, s + 1):
            r = cc[j] * 1.0 + mew[j] * r_prev + v[j] * r_prev2 + mewt[j] * z * r_prev
            r_prev2, r_prev = r_prev, r
        expect = Ts(w0 + w1 * z) / Ts(w0)
        assert r_prev == pytest.approx(expect, rel=1e-9, abs=1e-12), z
        assert abs(r_prev) <= 1.0 + 1e-12


# ---------------------------------------------------------------------------
# time step and counts against the shipped run (100x100, tpc=60)
# ---------------------------------------------------------------------------
def test_time_step_matches_shipped_stats():
    # stats.dat row 2, column 1: time after the first step = dt =
    # 0.259875000000000E-04 (courant 0.99, conductivity 1, PS/energy.f90:756-810)
    sim = _sim(nx=100, nz=100, tpc=60, grids=2, courant=0.99)
    out = np.zeros(10)
    oracle_lib().pto_eigen_bases(sim.ctx, _dp(out))
    assert out[9] == pytest.approx(0.259875000000000e-04, rel=5e-15)


def test_tracer_count_matches_shipped_stdout():
    # stdout:7 of the shipped run: ntr= 612060
    sim = _sim(nx=100, nz=100, tpc=60, grids=2)
    assert (sim.nx, sim.nz, sim.ntr) == (100, 100, 612060)


# ---------------------------------------------------------------------------
# initial tracer cloud: mzran streams and the per-thread column split
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("threads", [2, 3])
def test_initial_cloud_matches_reference_build(threads):
    g = np.load(GOLDEN / "ref_cloud.npz")
    sim = OracleSim(cases.isoviscous(nx=22, nz=10, tpc=3).params)
    assert sim.ntr == g[f"t{threads}_rtr"].size
    r, s = np.zeros(sim.ntr), np.zeros(sim.ntr)
    oracle_lib().pto_initial_positions(sim.ctx, threads, _dp(r), _dp(s))
    np.testing.assert_array_equal(r, g[f"t{threads}_rtr"])
    np.testing.assert_array_equal(s, g[f"t{threads}_str"])


@pytest.mark.parametrize("threads", [1, 2, 3])
def test_python_mzran_cloud_matches_reference_build(threads):
    """cases.mzran_stream / cases._positions_mzran (the Python restatement
    the workload builders use) against the same reference clouds: bit-exact
    for every thread count, and thread 0's columns are the one-thread
    stream's prefix."""
    g = np.load(GOLDEN / "ref_cloud.npz")
    r1, s1 = cases._positions_mzran(22, 10, 3, 1.0, threads=1)
    if threads == 1:
        # one thread: the same columns as the first thread of every split
        for t in (2, 3):
            k = (23 // t) * 11 * 3
            np.testing.assert_array_equal(r1[:k], g[f"t{t}_rtr"][:k])
            np.testing.assert_array_equal(s1[:k], g[f"t{t}_str"][:k])
        oracle_r, oracle_s = np.zeros(r1.size), np.zeros(r1.size)
        sim = OracleSim(cases.isoviscous(nx=22, nz=10, tpc=3).params)
        oracle_lib().pto_initial_positions(sim.ctx, 1, _dp(oracle_r), _dp(oracle_s))
        np.testing.assert_array_equal(r1, oracle_r)
        np.testing.assert_array_equal(s1, oracle_s)
        return
    r, s = cases._positions_mzran(22, 10, 3, 1.0, threads=threads)
    np.testing.assert_array_equal(r, g[f"t{threads}_rtr"])
    np.testing.assert_array_equal(s, g[f"t{threads}_str"])


def test_shipped_initial_projection():
    """The shipped run's t=0 snapshot T10000: tracers from 2 mzran streams,
    T and C from the shipped namelist's custom initial conditions
    (PS/initial_conditions.f90:42-43, 98-100 with f0=0, klayer=35,
    dlayer=0.5, RaT=1e5, RaC=5e4), projected (tracers_to_corners) and
    filtered (smoother_time, smoother_time_T).  Printed with 5 significant
    digits (g14.5)."""
    g = np.load(GOLDEN / "shipped_T10000.npz")
    p = _abi.default_params(aspect=1.0, nx=100, nz=100, tpc=60, grids=2, RaT=1e5, RaC=5e4, ivis=0,
                            courant=0.99, delta_visC=1.0)
    sim = OracleSim(p)
    r, s = np.zeros(sim.ntr), np.zeros(sim.ntr)
    oracle_lib().pto_initial_positions(sim.ctx, 2, _dp(r), _dp(s))
    RaT, RaC, kl, dl = 1e5, 5e4, 35.0, 0.5

    def comp(z):
        return 1.0 / (1.0 + np.exp(-2.0 * kl * (dl - z)))

    def temp(z):
        return (RaC * comp(z) + (RaT - RaC) * (1.0 - z)) / RaT

    zn = np.arange(101) / 100.0
    T_nodes = np.repeat(temp(zn)[:, None], 101, axis=1)
    case = cases.Case("shipped_100x100", p, T_nodes, float(temp(0.0)), float(temp(1.0)),
                      r, s, temp(s), comp(s))
    # stdout:11 of the shipped run: Tbot_iso,Ttop_iso
    assert case.Tbot_iso == pytest.approx(0.99999999999999967, abs=2e-16)
    assert case.Ttop_iso == pytest.approx(3.1525583800734926e-16, rel=1e-12)
    sim.load_state(case)
    assert np.array_equal(sim.nempty(), g["nempty"])
    for name, key in (("TVIS", "Tvis"), ("CNEW", "Cnew"), ("CVIS", "Cvis")):
        got = _node(sim, name)
        ref = g[key]
        tol = 5.0001e-5 * np.maximum(np.abs(ref), 1e-99) + 1e-300
        bad = np.abs(got - ref) > tol
        # values printed as exact zero by g14.5 (|v| < 1e-999) stay tiny
        bad &= ~((ref == 0.0) & (np.abs(got) < 1e-30))
        assert not bad.any(), f"{name}: {bad.sum()} nodes differ, worst {np.max(np.abs(got - ref))}"


# ---------------------------------------------------------------------------
# closed forms
# ---------------------------------------------------------------------------
def _sigma2(t):
    return -5.0 / 2.0 + (8.0 / 3.0) * np.cos(t) - (1.0 / 6.0) * np.cos(2 * t)


def _sigma4(t):
    return 28.0 / 3.0 - 13.0 * np.cos(t) + 4.0 * np.cos(2 * t) - (1.0 / 3.0) * np.cos(3 * t)


@pytest.mark.parametrize("a,b", [(1, 1), (3, 2), (7, 5), (15, 9)])
def test_operator_on_sine_modes(a, b):
    """Isoviscous free-slip operator (stream_derivatives + stream_space +
    enforceBCs_stream): sine modes are eigenvectors with the symbol
    sigma4(tx)/dr^4 + sigma4(tz)/ds^4 + 2 sigma2(tx) sigma2(tz)/(dr^2 ds^2)."""
    case = cases.isoviscous(nx=64, nz=32, tpc=1, grids=3, aspect=2.0)
    sim = OracleSim(case.params)
    sim.load_state(case)
    sim.multigrid()              # sets the isoviscous eta = 1 hierarchy
    for level in (1, 2):
        n, m = 64 >> (level - 1), 32 >> (level - 1)
        if a >= n or b >= m:
            continue
        dr, ds = 2.0 / n, 1.0 / m
        tx, tz = a * math.pi / n, b * math.pi / m
        i, k = np.arange(n + 1), np.arange(m + 1)
        psi = np.sin(tz * k)[:, None] * np.sin(tx * i)[None, :]
        out = np.zeros((m - 1) * (n - 1))
        rc = oracle_lib().pto_apply_operator(sim.ctx, level, _dp(np.ascontiguousarray(psi)), _dp(out))
        assert rc == 0
        lam = _sigma4(tx) / dr**4 + _sigma4(tz) / ds**4 + 2.0 * _sigma2(tx) * _sigma2(tz) / (dr**2 * ds**2)
        expect = lam * psi[1:m, 1:n]
        # rounding bound: the stencil sums cancel down to the symbol, so the
        # error scales with the sum of |weights| (26.67 for D4, 5.33 for D2)
        scale = 80.0 / 3.0 / dr**4 + 80.0 / 3.0 / ds**4 + 2.0 * (16.0 / 3.0) ** 2 / (dr**2 * ds**2)
        err = np.max(np.abs(out.reshape(m - 1, n - 1) - expect))
        assert err <= 64 * np.finfo(float).eps * scale, (level, err / scale)


def test_interpolation_is_exact_on_bicubics():
    """compute_derivatives + interpolate (PS/composition.f90:1359-1393,
    1457-1478) reproduce any bicubic sum a_pq r^p s^q."""
    case = cases.isoviscous(nx=32, nz=16, tpc=1, grids=2, aspect=2.0)
    sim = OracleSim(case.params)
    sim.load_state(case)
    rng = np.random.default_rng(5)
    coef = rng.standard_normal((4, 4))
    h = 3
    dr, ds = 2.0 / 32, 1.0 / 16

    def f(r, s):
        return sum(coef[p, q] * r**p * s**q for p in range(4) for q in range(4))

    ii = np.arange(-h, 32 + h + 1)
    kk = np.arange(-h, 16 + h + 1)
    U = f(dr * ii[None, :], ds * kk[:, None])
    sim.set_field("U", U)
    r = rng.uniform(0.0, 2.0, 500)
    s = rng.uniform(0.0, 1.0, 500)
    out = np.zeros(500)
    assert oracle_lib().pto_interpolate(sim.ctx, _abi.FIELD["U"], 500, _dp(r), _dp(s), _dp(out)) == 0
    assert rel_err(out, f(r, s)) <= 1e-12


def test_projection_constants_counts_and_brute_force():
    """tracers_to_corners (PS/composition.f90:942-1087): bilinear hat
    weights max(0,1-|r-ri|/dr) max(0,1-|s-sk|/ds) form a partition of unity,
    so constants are reproduced; every tracer is counted once at
    (nint(r/dr), nint(s/ds)); interior node values equal the brute-force
    weighted averages."""
    nx, nz, tpc = 12, 8, 3
    case = cases.isoviscous(nx=nx, nz=nz, tpc=tpc, grids=2)
    rng = np.random.default_rng(11)
    n = case.ntr
    r = rng.uniform(0.0, 1.0, n)
    s = rng.uniform(0.0, 1.0, n)
    r[:5] = [0.0, 1.0, 0.5, 0.5 / nx, 1.5 / nx]     # walls and exact half-cell ties
    s[:5] = [0.0, 1.0, 0.5, 0.5 / nz, 0.5]
    Ttr = rng.uniform(0.0, 1.0, n)
    const = cases.Case("c", case.params, case.T_nodes, case.Tbot_iso, case.Ttop_iso, r, s,
                       np.full(n, 0.7), np.full(n, 0.3))
    sim = OracleSim(case.params)
    sim.load_state(const)
    dr, ds = 1.0 / nx, 1.0 / nz
    T = _node(sim, "T")
    assert np.all(np.abs(T[1:nz, :] - 0.7) <= 1e-15)
    assert np.all(np.abs(_node(sim, "CNEW") - 0.3) <= 1e-15)
    ib = np.floor(r / dr + 0.5).astype(int)
    kb = np.floor(s / ds + 0.5).astype(int)
    cnt = np.zeros((nz + 1, nx + 1), dtype=np.int64)
    np.add.at(cnt, (kb, ib), 1)
    assert np.array_equal(sim.nempty(), cnt)
    assert sim.nempty().sum() == n

    var = cases.Case("v", case.params, case.T_nodes, case.Tbot_iso, case.Ttop_iso, r, s, Ttr, np.zeros(n))
    sim = OracleSim(case.params)
    sim.load_state(var)
    num = np.zeros((nz + 1, nx + 1))
    den = np.zeros((nz + 1, nx + 1))
    for di in (-1, 0, 1):
        for dk in (-1, 0, 1):
            i, k = ib + di, kb + dk
            ok = (i >= 0) & (i <= nx) & (k >= 0) & (k <= nz)
            w = np.maximum(0, 1 - np.abs(r - i * dr) / dr) * np.maximum(0, 1 - np.abs(s - k * ds) / ds)
            np.add.at(num, (k[ok], i[ok]), (w * Ttr)[ok])
            np.add.at(den, (k[ok], i[ok]), w[ok])
    got = _node(sim, "T")
    brute = np.where(den > 0, num / np.where(den > 0, den, 1), 0)
    inner = (slice(1, nz), slice(1, nx))
    assert np.max(np.abs(got[inner] - brute[inner])) <= 1e-14


def test_band_lu_matches_lapack():
    from scipy.linalg import lapack
    rng = np.random.default_rng(3)
    for n, kl, ku in ((40, 5, 5), (297, 27, 27), (60, 3, 7)):
        A = np.zeros((n, n))
        for j in range(n):
            lo, hi = max(0, j - ku), min(n, j + kl + 1)
            A[lo:hi, j] = rng.standard_normal(hi - lo)
        ld = 2 * kl + ku + 1
        ab = np.zeros((ld, n), order="F")
        for j in range(n):
            for i in range(max(0, j - ku), min(n, j + kl + 1)):
                ab[kl + ku + i - j, j] = A[i, j]
        b = rng.standard_normal(n)
        lu, piv, info = lapack.dgbtrf(ab.copy(order="F"), kl, ku)
        assert info == 0
        x_ref, info = lapack.dgbtrs(lu, kl, ku, b.copy(), piv)
        abo = ab.copy(order="F")
        ipiv = np.zeros(n, dtype=np.int32)
        bo = b.copy()
        rc = oracle_lib().pto_band_lu_solve(n, kl, ku, _dp(abo), ipiv.ctypes.data_as(C.POINTER(C.c_int32)), _dp(bo))
        assert rc == 0
        np.testing.assert_array_equal(ipiv, piv + 1)
        assert rel_err(bo, x_ref) <= 1e-11
        assert rel_err(bo, np.linalg.solve(A, b)) <= 1e-9


# ---------------------------------------------------------------------------
# trajectories of the reference build
# ---------------------------------------------------------------------------
def _check_state(sim, g, prefix, names, tol):
    for name in names:
        err = rel_err(_node(sim, name), g[f"{prefix}_{name}"])
        assert err <= tol, f"{prefix} {name}: {err:.3e}"
    assert np.array_equal(sim.nempty(), g[f"{prefix}_nempty"])
    idx = g[f"{prefix}_tracer_index"]
    for got, key in zip(sim.tracers(), ("rtr", "str", "Ttr", "Ctr")):
        err = rel_err(got[idx], g[f"{prefix}_{key}"])
        assert err <= tol, f"{prefix} tracers {key}: {err:.3e}"


def test_isoviscous_64_two_steps_match_reference_build():
    g = np.load(GOLDEN / "ref_iso64.npz")
    case = cases.isoviscous(nx=64, nz=64, tpc=20)
    sim = OracleSim(case.params)
    sim.load_state(case)
    # one solve per loop iteration plus three inside each energy_time
    ref_q = [int(q) for _, q in g["solves_tq"]]
    t = 0.0
    for step in range(3):
        sim.multigrid()
        assert sim.info()["q"] == ref_q[4 * step], f"main solve of step {step}"
        sim.compute_velocities()
        if step in (0, 2):
            _check_state(sim, g, f"s{step}", ("SF", "T", "U", "W", "TVIS", "CNEW"), 1e-10)
        if step == 2:
            break
        sim.energy_time()
        assert sim.info()["q"] == ref_q[4 * step + 3], f"last RK4 solve of step {step}"
        t += sim.info()["dt"]
    assert t == pytest.approx(float(g["s2_time"]), rel=1e-14)


def test_thermochem_96_one_step_matches_reference_build():
    g = np.load(GOLDEN / "ref_thermo96.npz")
    case = cases.thermochem(nx=96, nz=32, tpc=4, RaT=1e5, RaC=5e4, courant=0.1)
    sim = OracleSim(case.params)
    sim.load_state(case)
    sim.multigrid()
    assert sim.info()["q"] == g["solves_tq"][0][1]
    sim.compute_velocities()
    names = ("SF", "T", "U", "W", "TVIS", "CNEW", "CVIS", "VIS", "CONDUCT")
    _check_state(sim, g, "s0", names, 1e-10)
    sim.energy_time()
    sim.multigrid()
    assert sim.info()["q"] == g["solves_tq"][4][1]
    sim.compute_velocities()
    _check_state(sim, g, "s1", names, 1e-10)


# ---------------------------------------------------------------------------
# stats (PS/basics.f90:363-462) against the shipped stats.dat row 1 (t = 0)
# ---------------------------------------------------------------------------
# stats.dat row 1 of the shipped 100x100 run (g25.15): time, Tavg, vrms, Nu,
# Nu_bot, RaT_avg, mobility, strain_max, q1..q4, empty, min, max, Cavg, ent
SHIPPED_STATS_ROW1 = dict(
    time=0.0, Tavg=0.500000000000000, vrms=0.354798091580736e-03, Nu=0.500000000000018,
    Nu_bot=0.499999999999999, RaT_avg=100000.000000000, mobility=0.773460460515217,
    strain_max=0.423294330063474e-01, q1=0.500000000000020, q2=0.500000000000020,
    q3=0.499999999999998, q4=0.499999999999998, empty_cells=0, tracer_min=60, tracer_max=60,
    Cavg=0.499999323770751, ent=0.203858152927001e-01)


def shipped_initial_case(grids=2, tolerance=1e-12):
    """The shipped run's t=0 state: 100x100, tpc=60, tracers from 2 mzran
    streams, custom T and C (PS/initial_conditions.f90:42-43, 98-100 with
    f0=0, klayer=35, dlayer=0.5, RaT=1e5, RaC=5e4); T at the nodes is the
    analytic field (stats at t=0 integrate it, not the projection)."""
    p = _abi.default_params(aspect=1.0, nx=100, nz=100, tpc=60, grids=grids, RaT=1e5, RaC=5e4, ivis=0,
                            courant=0.99, delta_visC=1.0, tolerance=tolerance, dent=0.5)
    sim = OracleSim(p)
    r, s = np.zeros(sim.ntr), np.zeros(sim.ntr)
    oracle_lib().pto_initial_positions(sim.ctx, 2, _dp(r), _dp(s))
    RaT, RaC, kl, dl = 1e5, 5e4, 35.0, 0.5

    def comp(z):
        return 1.0 / (1.0 + np.exp(-2.0 * kl * (dl - z)))

    def temp(z):
        return (RaC * comp(z) + (RaT - RaC) * (1.0 - z)) / RaT

    zn = np.arange(101) / 100.0
    T_nodes = np.repeat(temp(zn)[:, None], 101, axis=1)
    case = cases.Case("shipped_100x100", p, T_nodes, float(temp(0.0)), float(temp(1.0)),
                      r, s, temp(s), comp(s))
    return p, case


def test_stats_of_the_shipped_initial_state():
    """The shipped run solved its t=0 state with the direct banded solver;
    the multigrid solve of the same discrete operator to a relative residual
    of 1e-12 gives the same psi to ~1e-10, so the solve-dependent values
    (vrms, mobility, strain_max) match to 1e-8, and the others (integrals of
    the analytic T, the projected C, the conductivity) to the printed 15
    digits."""
    p, case = shipped_initial_case()
    sim = OracleSim(p)
    sim.load_state(case)
    sim.multigrid()
    sim.compute_velocities()
    st = sim.stats()
    print({k: st[k] for k in SHIPPED_STATS_ROW1})
    for k, ref in SHIPPED_STATS_ROW1.items():
        got = st[k]
        if isinstance(ref, int):
            assert got == ref, k
        elif k in ("vrms", "mobility", "strain_max"):
            assert got == pytest.approx(ref, rel=1e-8), k
        else:
            assert got == pytest.approx(ref, rel=5e-14, abs=5e-15), k


def test_restart_record_of_the_shipped_Trestart():
    """print_restart's Trestart (PS/main.f90:457-470) of the shipped 100x100
    run: one record of xg(-6:106), zg(-6:106), T(-3:103)^2, Cnew(1,-5:105)^2,
    SF(-3:103)^2.  The oracle's build_grid coordinates equal the file's bit
    for bit; T, SF and Cnew at the nodes are the same step's T10001 columns
    (printed to 5 significant digits), which fixes which array is which."""
    g = np.load(GOLDEN / "shipped_Trestart.npz")
    nx = nz = 100
    n3, n5 = (nx + 7) * (nz + 7), (nx + 11) * (nz + 11)
    assert int(g["record_bytes"]) == 8 * (2 * (nx + 13) + 2 * n3 + n5) == 283560
    xg, zg = np.empty(nx + 13), np.empty(nz + 13)
    oracle_lib().pto_grid_coords(nx, nz, 1.0, _dp(xg), _dp(zg))
    assert np.array_equal(xg, g["xg"]) and np.array_equal(zg, g["zg"])
    # a wrong halo count or aspect moves the coordinates
    xg2 = np.empty(nx + 13)
    oracle_lib().pto_grid_coords(nx, nz, 1.0 + 1e-15, _dp(xg2), _dp(zg))
    assert not np.array_equal(xg2, g["xg"])

    def nodes(a, h):
        n = nx + 1 + 2 * h
        return a.reshape(n, n)[h:h + nz + 1, h:h + nx + 1]

    def close5(a, b):
        return np.all(np.abs(a - b) <= 6e-5 * np.abs(b) + 1e-14)

    T, SF, C = nodes(g["T"], 3), nodes(g["SF"], 3), nodes(g["Cnew"], 5)
    assert close5(T, g["snap_T"]) and close5(SF, g["snap_SF"]) and close5(C, g["snap_Cnew"])
    assert not close5(SF, g["snap_T"])
    # enforceBCs_stream held when it was written: walls zero, antisymmetric ghosts
    S = g["SF"].reshape(nz + 7, nx + 7)
    assert np.all(S[:, 3] == 0.0) and np.all(S[:, nx + 3] == 0.0) and np.all(S[3, :] == 0.0)
    assert np.array_equal(S[4:nz + 3, 2], -S[4:nz + 3, 4]) and np.array_equal(S[2, 4:nx + 3], -S[4, 4:nx + 3])


def test_load_state_refuses_sizes_that_do_not_fit_the_grid():
    """optimize_grid may widen the grid (30x20 becomes 32x20): a case built
    for the requested grid no longer fits the context, and loading it is
    refused instead of reading past the case's arrays."""
    case = cases.isoviscous(nx=30, nz=20, tpc=3, aspect=1.5)
    sim = OracleSim(case.params)
    assert (sim.nx, sim.nz) == (32, 20)
    with pytest.raises(ValueError):
        sim.load_state(case)
    ok = cases.isoviscous(nx=32, nz=20, tpc=3, aspect=1.6)
    fit = OracleSim(ok.params)
    fit.load_state(ok)
    assert fit.ntr == ok.ntr
