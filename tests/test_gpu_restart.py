"""Checkpoint and resume through the C ABI (pt_write_restart /
pt_read_restart) in the reference's own file format.

* The shipped Trestart read into a 100x100 context and written back: the
  record framing, the build_grid coordinates and SF come back bit for bit
  (SF passes through enforceBCs_stream unchanged, as the reference wrote it
  with those boundary conditions).
* A state written after one step and read into a fresh context: the tracers
  (ID order) and SF are bit-identical; T and Cnew are the tracers'
  re-projection (rounding-level differences from the changed summation order
  within cells); writing again reproduces both files' grid, SF and tracers.
* Resuming: the restarted context (one extra warm-started solve, as the
  reference's restart=1 run does) follows the continuous run.
"""
import numpy as np
import pytest

from projectracer_amd import Simulation, cases, _abi
from _support import GOLDEN, interior, rel_err
from test_oracle import shipped_initial_case

pytestmark = pytest.mark.gpu


def _record(path):
    raw = path.read_bytes()
    head, tail = int.from_bytes(raw[:4], "little"), int.from_bytes(raw[-4:], "little")
    assert head == tail == len(raw) - 8
    return np.frombuffer(raw[4:-4], dtype="<f8")


def _write_record(path, *arrays):
    data = b"".join(np.ascontiguousarray(a, dtype="<f8").tobytes() for a in arrays)
    n = len(data).to_bytes(4, "little")
    path.write_bytes(n + data + n)


def _node(sim, name):
    h = _abi.FIELD_HALO[_abi.FIELD[name]]
    return interior(sim.field(name), h, sim.nx, sim.nz).copy()


def test_shipped_Trestart_round_trip(tmp_path):
    print("[restart] shipped Trestart, 100x100, 612060 tracers", flush=True)
    g = np.load(GOLDEN / "shipped_Trestart.npz")
    p, case = shipped_initial_case()
    src, dst = tmp_path / "in", tmp_path / "out"
    src.mkdir()
    dst.mkdir()
    _write_record(src / "Trestart", g["xg"], g["zg"], g["T"], g["Cnew"], g["SF"])
    # restart_tracers of the shipped run is not in the reference: the t=0 cloud stands in
    _write_record(src / "restart_tracers", case.rtr, case.str_, case.Ctr, case.Ttr)
    sim = Simulation(p)
    sim.read_restart(src, case.Tbot_iso, case.Ttop_iso)
    sim.print_restart(dst)
    out = _record(dst / "Trestart")
    shipped = (src / "Trestart").read_bytes()
    assert (dst / "Trestart").stat().st_size == len(shipped) == 283568
    n = 113
    assert np.array_equal(out[:n], g["xg"]) and np.array_equal(out[n:2 * n], g["zg"])
    assert np.array_equal(out[-g["SF"].size:], g["SF"])
    assert (dst / "restart_tracers").read_bytes() == (src / "restart_tracers").read_bytes()
    # T and Cnew are the t=0 cloud's projection: the same as a plain start from that cloud
    ref = Simulation(p)
    ref.load_state(case)
    assert np.array_equal(sim.nempty(), ref.nempty())
    assert rel_err(_node(sim, "CNEW"), _node(ref, "CNEW")) <= 1e-14


def test_restart_after_a_step_and_resume(tmp_path):
    print("[restart] 64x64 isoviscous FP64: write after one step, read, resume", flush=True)
    case = cases.isoviscous(nx=64, nz=64, tpc=8)
    a = Simulation(case.params)
    a.load_state(case)
    a.multigrid(); a.compute_velocities()
    a.energy_time(); a.multigrid(); a.compute_velocities()
    d1, d2 = tmp_path / "a", tmp_path / "b"
    d1.mkdir()
    d2.mkdir()
    a.print_restart(d1)
    b = Simulation(case.params)
    b.read_restart(d1, case.Tbot_iso, case.Ttop_iso)
    for x, y in zip(a.tracers(), b.tracers()):
        assert np.array_equal(x, y)
    assert np.array_equal(_node(a, "SF"), _node(b, "SF"))
    assert np.array_equal(a.nempty(), b.nempty())
    assert rel_err(_node(b, "T"), _node(a, "T")) <= 1e-14
    assert rel_err(_node(b, "CNEW"), _node(a, "CNEW")) <= 1e-14
    b.print_restart(d2)
    ra, rb = _record(d1 / "Trestart"), _record(d2 / "Trestart")
    n3 = 71 * 71
    assert np.array_equal(ra[:2 * 77], rb[:2 * 77]) and np.array_equal(ra[-n3:], rb[-n3:])
    assert (d1 / "restart_tracers").read_bytes() == (d2 / "restart_tracers").read_bytes()
    # resume: the restarted run solves once more from the written psi, then steps on
    b.multigrid(); b.compute_velocities()
    for _ in range(2):
        a.energy_time(); a.multigrid(); a.compute_velocities()
        b.energy_time(); b.multigrid(); b.compute_velocities()
    ea = {k: rel_err(_node(b, k), _node(a, k)) for k in ("SF", "U", "W", "T", "CNEW")}
    print(f"[restart] resumed vs continuous after 2 steps: {ea}", flush=True)
    assert max(ea.values()) <= 1e-6
    assert np.array_equal(a.nempty(), b.nempty())
    for x, y in zip(a.tracers(), b.tracers()):
        assert rel_err(y, x) <= 1e-8


def test_restart_rejects_a_mismatched_grid(tmp_path):
    case = cases.isoviscous(nx=64, nz=64, tpc=2)
    a = Simulation(case.params)
    a.load_state(case)
    a.print_restart(tmp_path)
    other = cases.isoviscous(nx=64, nz=32, tpc=2)
    b = Simulation(other.params)
    with pytest.raises(Exception) as e:
        b.read_restart(tmp_path, 1.0, 0.0)
    assert "record" in str(e.value)
    with pytest.raises(Exception):
        b.read_restart(tmp_path / "missing", 1.0, 0.0)
