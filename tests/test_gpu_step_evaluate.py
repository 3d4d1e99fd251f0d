"""pfc_step_evaluate: an evaluate (pfc_evaluate_sums) followed by steps,
with the evaluate's sums taken by the first step's row and column passes
(src/pfc-relax.c:987-1130 then :1133-1247, as relax() prints and steps).

Checked against pfc_evaluate_sums + pfc_step on the same engine state
(fave, nave and the densities after the steps), on tuned shapes in both
forms and on a generic shape (where it is that sequence), at step indices
that are and are not multiples of 100, and against the oracle's relax()
order.
"""
import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import FP64, MIXED, graphene_params, hetero_params
from pfc_amd.engine import Engine
from pfc_amd.inputs import hetero_patchwork

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def state(Nx, Ny, U, seed):
    if U == 3:
        return hetero_params(), hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=24, seed=seed)
    rng = np.random.default_rng(seed)
    return graphene_params(), -0.2 + 0.05 * rng.standard_normal((1, Ny, Nx))


# (Nx, Ny, U): tuned rows/columns of several tilings, 4096- and 8192-point columns
# (per-field column pass on the plain steps), 4096- and 8192-point rows (the
# persistent FP64 row pass), and a generic shape
SHAPES = [(64, 64, 3), (128, 64, 1), (64, 4096, 3), (64, 8192, 3), (4096, 64, 3), (8192, 64, 3),
          (256, 256, 1), (100, 90, 3)]


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("Nx,Ny,U", SHAPES)
@pytest.mark.parametrize("t0,n", [(100, 5), (37, 4), (0, 1)])
def test_matches_evaluate_then_step(gpu, Nx, Ny, U, precision, t0, n):
    p, f = state(Nx, Ny, U, seed=Nx + Ny + U)
    a = Engine(p, Nx, Ny, 0.7, 0.7, precision)
    b = Engine(p, Nx, Ny, 0.7, 0.7, precision)
    for g in (a, b):
        g.set_fields(f)
        g.step(t0 - 7 if t0 >= 7 else 1, 7)     # move off the set_fields state
    fa, na = a.step_evaluate(t0, n)
    fb, nb = b.evaluate(density=False)
    b.step(t0, n)
    # FP64: the same sums over other block partitions; mixed: the refreshed
    # FP32 state from another transform kernel (FP32 round-off)
    tol = 1e-13 if precision == FP64 else 1e-6
    assert abs(fa - fb) <= tol * abs(fb), (fa, fb)
    assert np.allclose(na, nb, rtol=tol, atol=1e-15)
    assert rel_l2(a.get_fields(), b.get_fields()) < (1e-13 if precision == FP64 else 1e-5)
    with pytest.raises(RuntimeError):
        a.get_energy()                          # as after pfc_evaluate_sums
    a.close()
    b.close()


def test_headline_grid(gpu):
    """4096^2 with three fields (the benchmark's grid and kernels, FP64): a
    print at step 100 fused into the refresh step against pfc_evaluate_sums
    then pfc_step, to 1e-13 (the same sums over other block partitions)."""
    N = 4096
    p = hetero_params(dt=0.25, c=1.0, sigma=2.0 / 7.2552)
    f = hetero_patchwork(N, N, 0.7, 0.7, domain=256, seed=2024, noise=0.01)
    with Engine(p, N, N, 0.7, 0.7, FP64) as a, Engine(p, N, N, 0.7, 0.7, FP64) as b:
        for g in (a, b):
            g.set_fields(f)
            g.step(91, 9)
        fa, na = a.step_evaluate(100, 3)
        fb, nb = b.evaluate(density=False)
        b.step(100, 3)
        assert abs(fa - fb) <= 1e-13 * abs(fb), (fa, fb)
        assert np.allclose(na, nb, rtol=1e-13, atol=1e-15)
        assert rel_l2(a.get_fields(), b.get_fields()) < 1e-13
        assert a.nan_step() == -1 and b.nan_step() == -1


@pytest.mark.parametrize("Nx,Ny", [(128, 64), (64, 4096)])
def test_relax_prints_against_oracle(gpu, Nx, Ny):
    """relax()'s order with a print every 100 steps, FP64: every printed fave
    to 1e-12, nave to 1e-14, the densities after 200 steps to 1e-12."""
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=24, seed=3)
    g = Engine(p, Nx, Ny, 0.7, 0.7, FP64)
    o = Oracle(p, Nx, Ny, 0.7, 0.7)
    g.set_fields(f)
    o.set_fields(f)
    for t0 in (0, 100):
        fg, ng = g.step_evaluate(t0, 100)
        fo, no = o.evaluate()
        assert o.steps(t0, 100) == 0
        assert abs(fg - fo) <= 1e-12 * abs(fo)
        assert np.allclose(ng, no, rtol=0, atol=1e-14)
    assert rel_l2(g.get_fields(), o.get_fields()) < 1e-12
    g.close()


def test_nan_reported_after_the_print(gpu):
    """A NaN density: the evaluate's fave is NaN, and the step reports it
    (pfc-relax prints the line, then exits 39, as the reference's step()
    fails after evaluate() has printed)."""
    p = hetero_params()
    f = hetero_patchwork(64, 64, 0.7, 0.7, domain=24, seed=5)
    f[0, 0, 0] = np.nan
    g = Engine(p, 64, 64, 0.7, 0.7, FP64)
    g.set_fields(f)
    fave, _ = g.step_evaluate(0, 3)
    assert np.isnan(fave)
    assert g.nan_step() == 0
    g.close()
