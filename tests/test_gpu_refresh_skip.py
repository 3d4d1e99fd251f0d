"""A t % 100 step right after an evaluate (or pfc_set_fields) skips its
refresh: the evaluate has just recomputed the spectral state from the same
densities (src/pfc-relax.c:1126-1128), which is what the refresh
(:1178-1184) would compute again.  Checked against the same run with the
refresh forced (pfc_perturb with rel = 0 multiplies the state by exactly 1
and marks it changed), against the oracle, and that a perturbation in
between brings the refresh back.
"""
import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import FP64, MIXED, hetero_params
from pfc_amd.engine import Engine
from pfc_amd.inputs import hetero_patchwork

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("Nx,Ny", [(128, 128), (64, 4096), (4096, 64)])
def test_skip_matches_forced_refresh(gpu, Nx, Ny, precision):
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=32, seed=9)
    a = Engine(p, Nx, Ny, 0.7, 0.7, precision)
    b = Engine(p, Nx, Ny, 0.7, 0.7, precision)
    for g in (a, b):
        g.set_fields(f)
        g.step(90, 10)
    ea = a.evaluate(density=False)       # a: the step at 100 skips its refresh
    eb = b.evaluate(density=False)
    b.perturb(0.0, 1)                    # b: state unchanged, refresh forced
    assert ea[0] == eb[0] and np.array_equal(ea[1], eb[1])
    a.step(100, 5)
    b.step(100, 5)
    tol = 1e-14 if precision == FP64 else 1e-6
    assert rel_l2(a.get_fields(), b.get_fields()) < tol
    a.close()
    b.close()


def test_relax_cadence_against_oracle(gpu):
    """relax()'s order (evaluate, then step) over two print intervals, FP64,
    against the oracle's refreshing steps: 1e-12."""
    Nx, Ny = 128, 64
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=24, seed=3)
    g = Engine(p, Nx, Ny, 0.7, 0.7, FP64)
    o = Oracle(p, Nx, Ny, 0.7, 0.7)
    g.set_fields(f)
    o.set_fields(f)
    for t0 in (0, 100):
        fg, _ = g.evaluate(density=False)
        fo, _ = o.evaluate()
        assert abs(fg - fo) <= 1e-12 * abs(fo)
        g.step(t0, 100)
        assert o.steps(t0, 100) == 0
    assert rel_l2(g.get_fields(), o.get_fields()) < 1e-12
    g.close()


def test_perturb_clears_the_mark(gpu):
    """pfc_perturb changes the density's state and the spectral state with
    different noise, so after it the state is no longer the evaluate's: the
    next t % 100 step must refresh.  Against the same run without the
    evaluate (nothing marked, so the step refreshes): equal to round-off,
    where a skipped refresh would leave the inconsistent perturbed state."""
    Nx, Ny = 128, 128
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=32, seed=5)
    a = Engine(p, Nx, Ny, 0.7, 0.7, FP64)
    b = Engine(p, Nx, Ny, 0.7, 0.7, FP64)
    for g in (a, b):
        g.set_fields(f)
        g.step(98, 2)
    a.evaluate(density=False)            # marks a's state fresh
    for g in (a, b):
        g.perturb(1e-3, 7)
        g.step(100, 3)
    assert rel_l2(a.get_fields(), b.get_fields()) < 1e-13
    a.close()
    b.close()
