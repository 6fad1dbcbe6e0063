"""Synthetic initial states for the workloads this path is measured on.

Each builder returns a `Case`: the namelist values (`PtParams`) plus the
state the reference program would hold after `initial_temperature` and
`initialize_tracers` (PS/basics.f90:335-361, PS/composition.f90:776-858):
node temperatures T(0:nx,0:nz), the isothermal wall values, and tracers
(r, s, T, C) in ID order, `tpc` tracers per grid node placed uniformly in
each node's dual cell and clamped to the walls (boundary_tracers=1).

Workloads (nx x nz, tracers = (nx+1)(nz+1)tpc, tracers counted per node as
the reference does, PS/basics.f90:238):
  * isoviscous_64     lambda=1, 64x64, Ra_T=1e5, Ra_C=0, eta=1, tpc=20,
                      T=(1-z)+0.01 cos(pi x) sin(pi z); positions from the
                      reference's own generator (mzran stream 0,
                      PS/composition.f90:7-20, 786-847 with one thread).
  * thermochem_384    lambda=3, 384x128, Ra_T=1e6, Ra_C=5e5,
                      eta=exp(-ln(1e3)T+ln(10)C), tpc=40, dense layer
                      C=1 for z<0.2+0.01cos(2 pi x/3), T=(1-z)+U(-0.01,0.01)
                      noise (seed 1).
  * ensemble_case     one case of the batched ensemble: the grid drawn from
                      {128x64, 192x64, 256x128, 384x128} with lambda = nx/nz,
                      Ra_T log-uniform in [1e5, 1e7], Ra_C = B Ra_T with B
                      uniform in [0, 1], layer depth uniform in [0.1, 0.4],
                      eta = exp(-gamma T) with gamma uniform in [0, ln(1e4)],
                      30 tracers per node, T perturbed as thermochem_1024,
                      courant 0.01 (seeds 100..163 for the 64 cases).
  * thermochem_1024   lambda=4, 1024x256, Ra_T=1e7, Ra_C=3e6,
                      eta=exp(-ln(1e5)T+ln(30)C), tpc=50 (13,171,250
                      tracers), layer at z=0.3, T=(1-z)+0.01 sum_k a_k
                      cos(k pi x/4) sin(pi z), a_k~N(0,1) (seed 2).
Reduced-size variants keep the physics and shrink the grid/tpc so the CPU
oracle finishes in seconds.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from ._abi import PtParams, default_params

PII = 3.1415926535897932


@dataclass
class Case:
    name: str
    params: PtParams
    T_nodes: np.ndarray      # (nz+1, nx+1) C order == Fortran T(0:nx,0:nz)
    Tbot_iso: float
    Ttop_iso: float
    rtr: np.ndarray
    str_: np.ndarray
    Ttr: np.ndarray
    Ctr: np.ndarray

    @property
    def nx(self):
        return self.params.nx

    @property
    def nz(self):
        return self.params.nz

    @property
    def ntr(self):
        return self.rtr.size


def mzran_stream(n: int) -> np.ndarray:
    """mzran (PS/composition.f90:7-20) from the thread-0 seeds of
    PS/basics.f90:321-324: wrapping int32 lagged-subtract + LCG."""
    si, sj, sk, sn = 521288629, 362436069, 16163801, 1131199299
    out = np.empty(n, dtype=np.float64)
    for t in range(n):
        ran = si - sk
        if ran < 0:
            ran += 2147483579
        si, sj, sk = sj, sk, ran
        sn = (69069 * sn + 1013904243) & 0xFFFFFFFF
        r = (ran + sn) & 0xFFFFFFFF
        if r >= 0x80000000:
            r -= 0x100000000
        out[t] = 0.5 + 0.2328306e-9 * float(r)
    return out


def _positions_mzran(nx, nz, tpc, aspect, threads=1):
    """initialize_tracers' positions (PS/composition.f90:786-811) with
    `threads` OpenMP threads: thread t fills the node columns
    nx_start(t)..nx_finish(t), (nx+1)/threads of them (the last thread the
    rest), in ID order, from its own mzran stream.  Thread t's seeds are
    thread 0's advanced t steps (PS/basics.f90:325-331), so its stream is
    thread 0's shifted by t."""
    dr, ds = aspect / nx, 1.0 / nz
    per = (nx + 1) // threads
    starts = [t * per for t in range(threads)]
    finish = [s + per - 1 for s in starts[:-1]] + [nx]
    longest = max(f - s + 1 for s, f in zip(starts, finish))
    u = mzran_stream(threads - 1 + 2 * longest * (nz + 1) * tpc)
    rs, ss = [], []
    for t, (i0, i1) in enumerate(zip(starts, finish)):
        ncol = i1 - i0 + 1
        n = ncol * (nz + 1) * tpc
        ut = u[t:t + 2 * n]
        pr, ps = ut[0::2], ut[1::2]
        i = np.repeat(np.arange(i0, i1 + 1), (nz + 1) * tpc)
        k = np.tile(np.repeat(np.arange(nz + 1), tpc), ncol)
        rs.append(dr * i.astype(np.float64) + (pr - 0.5) * dr)
        ss.append(ds * k.astype(np.float64) + (ps - 0.5) * ds)
    r, s = np.concatenate(rs), np.concatenate(ss)
    return np.clip(r, 0.0, aspect), np.clip(s, 0.0, 1.0)


def _positions_seeded(nx, nz, tpc, aspect, seed):
    dr, ds = aspect / nx, 1.0 / nz
    rng = np.random.default_rng(seed)
    n = (nx + 1) * (nz + 1) * tpc
    pr = rng.random(n)
    ps = rng.random(n)
    # node-major ID order, as initialize_tracers (i outer, k, then p)
    i = np.repeat(np.arange(nx + 1), (nz + 1) * tpc)
    k = np.tile(np.repeat(np.arange(nz + 1), tpc), nx + 1)
    r = dr * i.astype(np.float64) + (pr - 0.5) * dr
    s = ds * k.astype(np.float64) + (ps - 0.5) * ds
    return np.clip(r, 0.0, aspect), np.clip(s, 0.0, 1.0)


def _bilinear(T_nodes, r, s, dr, ds):
    nz1, nx1 = T_nodes.shape
    fi = np.clip(r / dr, 0, nx1 - 1)
    fk = np.clip(s / ds, 0, nz1 - 1)
    i0 = np.minimum(np.floor(fi).astype(np.int64), nx1 - 2)
    k0 = np.minimum(np.floor(fk).astype(np.int64), nz1 - 2)
    a, b = fi - i0, fk - k0
    return ((1 - a) * (1 - b) * T_nodes[k0, i0] + a * (1 - b) * T_nodes[k0, i0 + 1]
            + (1 - a) * b * T_nodes[k0 + 1, i0] + a * b * T_nodes[k0 + 1, i0 + 1])


def _finish(name, p, T_nodes, r, s, Ttr, Ctr):
    return Case(name, p, np.ascontiguousarray(T_nodes), float(T_nodes[0, 0]),
                float(T_nodes[-1, 0]), np.ascontiguousarray(r), np.ascontiguousarray(s),
                np.ascontiguousarray(Ttr), np.ascontiguousarray(Ctr))


def isoviscous(nx=64, nz=64, tpc=20, **over) -> Case:
    """Config: lambda=1, Ra_T=1e5, eta=1; T=(1-z)+0.01cos(pi x)sin(pi z)
    (initial_temp=2, PS/initial_conditions.f90:60-61) evaluated with libm."""
    p = default_params(nx=nx, nz=nz, tpc=tpc, RaT=1e5, RaC=0.0,
                       ivis=0, **{"aspect": 1.0, "courant": 0.5, "grids": 6, **over})
    aspect = p.aspect
    dr, ds = aspect / nx, 1.0 / nz

    def temp(x, z):
        return (1.0 - z) + 0.01 * math.cos(PII * x / aspect) * math.sin(PII * z / aspect)

    T = np.empty((nz + 1, nx + 1))
    for k in range(nz + 1):
        for i in range(nx + 1):
            T[k, i] = temp(dr * i, ds * k)
    T[0, :] = T[0, 0]
    T[nz, :] = T[nz, 0]
    r, s = _positions_mzran(nx, nz, tpc, aspect)
    Ttr = np.array([temp(x, z) for x, z in zip(r.tolist(), s.tolist())])
    return _finish(f"isoviscous_{nx}x{nz}", p, T, r, s, Ttr, np.zeros_like(r))


def thermochem(nx=384, nz=128, tpc=40, aspect=3.0, RaT=1e6, RaC=5e5, visT=1e3, visC=0.1,
               layer=0.2, seed=1, **over) -> Case:
    """Config: dense layer C=1 for z<layer+0.01cos(2 pi x/lambda), T=(1-z)
    plus uniform noise of amplitude 0.01 (seed), eta(T,C) with ivis=2."""
    p = default_params(aspect=aspect, nx=nx, nz=nz, tpc=tpc, RaT=RaT, RaC=RaC, visT=visT, visC=visC,
                       **{"ivis": 2, "delta_visT": 2.0, "delta_visC": 2.0, "courant": 0.5, "grids": 6, **over})
    dr, ds = aspect / nx, 1.0 / nz
    rng = np.random.default_rng(seed)
    x = dr * np.arange(nx + 1)
    z = ds * np.arange(nz + 1)
    T = (1.0 - z)[:, None] + rng.uniform(-0.01, 0.01, size=(nz + 1, nx + 1))
    T[0, :] = 1.0
    T[nz, :] = 0.0
    r, s = _positions_seeded(nx, nz, tpc, aspect, seed + 1000)
    Ttr = _bilinear(T, r, s, dr, ds)
    Ctr = (s < layer + 0.01 * np.cos(2.0 * PII * r / aspect)).astype(np.float64)
    del x
    return _finish(f"thermochem_{nx}x{nz}", p, T, r, s, Ttr, Ctr)


def thermochem_large(nx=1024, nz=256, tpc=50, aspect=4.0, seed=2, **over) -> Case:
    """The 1024x256 workload: Ra_T=1e7, Ra_C=3e6, eta=exp(-ln(1e5)T+ln(30)C),
    two-layer C with interface at z=0.3, T=(1-z)+0.01 sum_{k=1..8} a_k
    cos(k pi x/lambda) sin(pi z), a_k ~ N(0,1) (seed)."""
    p = default_params(aspect=aspect, nx=nx, nz=nz, tpc=tpc,
                       **{"RaT": 1e7, "RaC": 3e6, "ivis": 2, "visT": 1e5, "visC": 1.0 / 30.0,
                          "delta_visT": 2.0, "delta_visC": 2.0, "courant": 0.5, "grids": 6, **over})
    dr, ds = aspect / nx, 1.0 / nz
    rng = np.random.default_rng(seed)
    a = rng.standard_normal(8)

    def temp(xv, zv):
        acc = np.zeros(np.broadcast(xv, zv).shape)
        for kk in range(1, 9):
            acc = acc + a[kk - 1] * np.cos(kk * PII * xv / aspect)
        return (1.0 - zv) + 0.01 * acc * np.sin(PII * zv)

    xg = dr * np.arange(nx + 1)
    zg = ds * np.arange(nz + 1)
    T = temp(xg[None, :], zg[:, None])
    T[0, :] = T[0, 0]
    T[nz, :] = T[nz, 0]
    r, s = _positions_seeded(nx, nz, tpc, aspect, seed + 1000)
    Ttr = temp(r, s)
    Ctr = (s < 0.3 + 0.01 * np.cos(2.0 * PII * r / aspect)).astype(np.float64)
    return _finish(f"thermochem_{nx}x{nz}", p, T, r, s, Ttr, Ctr)


ENSEMBLE_GRIDS = ((128, 64), (192, 64), (256, 128), (384, 128))


def ensemble_case(seed: int, tpc: int = 30, gamma_max: float = math.log(1e4), **over) -> Case:
    """One case of the batched ensemble, drawn from `seed` (see the module
    docstring); ivis=1 with visT = e^gamma, visP = 1 gives eta = exp(-gamma T)."""
    rng = np.random.default_rng(seed)
    nx, nz = ENSEMBLE_GRIDS[int(rng.integers(0, len(ENSEMBLE_GRIDS)))]
    RaT = 10.0 ** rng.uniform(5.0, 7.0)
    B = rng.uniform(0.0, 1.0)
    layer = rng.uniform(0.1, 0.4)
    gam = rng.uniform(0.0, gamma_max)
    aspect = nx / nz
    p = default_params(aspect=aspect, nx=nx, nz=nz, tpc=tpc,
                       **{"RaT": RaT, "RaC": B * RaT, "ivis": 1, "visT": float(math.exp(gam)), "visP": 1.0,
                          "delta_visT": 2.0, "delta_visC": 2.0, "courant": 0.01, "grids": 6, **over})
    dr, ds = aspect / nx, 1.0 / nz
    a = rng.standard_normal(8)

    def temp(xv, zv):
        acc = np.zeros(np.broadcast(xv, zv).shape)
        for kk in range(1, 9):
            acc = acc + a[kk - 1] * np.cos(kk * PII * xv / aspect)
        return (1.0 - zv) + 0.01 * acc * np.sin(PII * zv)

    xg = dr * np.arange(nx + 1)
    zg = ds * np.arange(nz + 1)
    T = temp(xg[None, :], zg[:, None])
    T[0, :] = T[0, 0]
    T[nz, :] = T[nz, 0]
    r, s = _positions_seeded(nx, nz, tpc, aspect, seed + 1000)
    Ttr = temp(r, s)
    Ctr = (s < layer + 0.01 * np.cos(2.0 * PII * r / aspect)).astype(np.float64)
    return _finish(f"ensemble_{seed}_{nx}x{nz}", p, T, r, s, Ttr, Ctr)
