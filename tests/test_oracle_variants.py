"""The oracle on the branches that no shipped output file covers, against
the reference program compiled from its Fortran sources
(tests/golden/ref_variants.npz): rigid top/bottom walls with an insulating
bottom (Vbc=1, Tbot=1; PS/iterative_solver.f90:1616-1664, PS/energy.f90:558-592)
through multigrid and through the direct solver (PS/momentum.f90:362-606), and
the T- and depth-dependent viscosity law ivis=1 (PS/rheology.f90:2-23); and
the RK1 fine grid smoother (RKC1=0, PS/iterative_solver.f90:1173-1186,
:1468-1485, :1503-1513) on an isoviscous and an ivis=2 state
(tests/golden/ref_rk1.npz).  Two steps: the cycle count of every solve group
exact, node counts exact, psi, u, w, T within 1e-10.  CPU only."""
import numpy as np
import pytest

from _support import GOLDEN, OracleSim, RK1_VARIANTS, VARIANTS, rel_err, run_variant, variant_case


@pytest.mark.parametrize("name", list(VARIANTS) + list(RK1_VARIANTS))
def test_variant_against_the_reference_program(name):
    g = np.load(GOLDEN / ("ref_rk1.npz" if name in RK1_VARIANTS else "ref_variants.npz"))
    case = variant_case(name)
    sim = OracleSim(case.params)
    sim.load_state(case)

    def check(step, f):
        errs = {n: rel_err(f[n], g[f"{name}_s{step}_{n}"]) for n in f}
        print(name, step, errs)
        assert max(errs.values()) <= 1e-10, (step, errs)

    print(name, "solve groups", run_variant(sim, g, name, check))
