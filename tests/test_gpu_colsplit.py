"""The per-field column pass (k_col_step_f, two state buffers) against the
in-place all-fields pass (PFC_COLSPLIT=0) and against the oracle, in both
forms (the mixed form also alternates its FP64 k = 0 state).

The per-field pass runs non-refresh steps in pairs (state Nh -> Nh2 -> Nh);
a step without a partner and every t % 100 step take the in-place pass.  Its
update performs the same operations in the same order as the in-place one
(kupdate's couplings by ascending field index, then A and B), so the two
agree to round-off on every column length, field count and step pattern:
odd and even step counts, calls that start or end between the two steps of
a pair, refresh steps inside a call.
"""
import os

import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import FP64, MIXED, graphene_params, hetero_params
from pfc_amd.engine import Engine
from pfc_amd.inputs import hetero_patchwork

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def engine(params, Nx, Ny, split, precision=FP64):
    old = os.environ.get("PFC_COLSPLIT")
    os.environ["PFC_COLSPLIT"] = "1" if split else "0"
    try:
        return Engine(params, Nx, Ny, 0.7, 0.7, precision)
    finally:
        if old is None:
            del os.environ["PFC_COLSPLIT"]
        else:
            os.environ["PFC_COLSPLIT"] = old


# (t0, nsteps) call sequences: pairs, leftovers, and refresh steps inside calls
CALLS = [(0, 3), (3, 4), (7, 1), (8, 97), (105, 2)]
# The mixed form's comparison stays short: a last-bit FP32 difference (where
# the compiler fuses a multiply-add in one pass and not in the other) grows
# with the relaxation's sensitivity, to ~3e-6 over 100 steps at 8192 rows.
CALLS_MIXED = [(0, 3), (3, 4), (7, 1), (97, 6)]


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("Nx,Ny,U", [(64, 4096, 3), (128, 2048, 1), (64, 8192, 3), (256, 256, 2),
                                     (128, 64, 3), (64, 512, 3)])
def test_split_matches_in_place(gpu, Nx, Ny, U, precision):
    if U == 3:
        p = hetero_params()
    elif U == 1:
        p = graphene_params()
    else:
        from test_gpu_parity import two_field_params
        p = two_field_params()
    if U == 3:
        f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=32, seed=Ny + U)
    else:
        rng = np.random.default_rng(Ny + U)
        f = 0.1 + 0.05 * rng.standard_normal((U, Ny, Nx))
    # the same operations in the same order: equal up to the compiler's
    # choice of fused multiply-adds
    tol = 1e-14 if precision == FP64 else 1e-6
    a = engine(p, Nx, Ny, True, precision)
    b = engine(p, Nx, Ny, False, precision)
    for g in (a, b):
        g.set_fields(f)
    for t0, n in (CALLS if precision == FP64 else CALLS_MIXED):
        a.step(t0, n)
        b.step(t0, n)
        fa, fb = a.get_fields(), b.get_fields()
        assert rel_l2(fa, fb) < tol, (t0, n, rel_l2(fa, fb))
    ea, eb = a.evaluate(), b.evaluate()
    assert abs(ea[0] - eb[0]) <= tol * abs(eb[0])
    assert np.allclose(ea[1], eb[1], rtol=tol, atol=0)
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_split_hetero_4096_columns_against_oracle(gpu, precision):
    """Columns of 4096 (the benchmark's length) through pairs of per-field
    steps, against the oracle: FP64 1e-12 in the densities and the energy,
    mixed 2e-6 and 1e-6; the means to 1e-15 (FP64) and 1e-7 relative (the
    mixed form's FP32 density around its FP64 k = 0 state)."""
    Nx, Ny = 64, 4096
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=40, seed=4)
    g = engine(p, Nx, Ny, True, precision)
    g.set_fields(f)
    o = Oracle(p, Nx, Ny, 0.7, 0.7)
    o.set_fields(f)
    g.step(1, 6)
    assert o.steps(1, 6) == 0
    tf, te = (1e-12, 1e-12) if precision == FP64 else (2e-6, 1e-6)
    assert rel_l2(g.get_fields(), o.get_fields()) < tf
    fg, ng = g.evaluate()
    fo, no = o.evaluate()
    assert abs(fg - fo) <= te * abs(fo)
    if precision == FP64:
        assert np.allclose(ng, no, rtol=0, atol=1e-15)
    else:
        assert np.allclose(ng, no, rtol=1e-7, atol=0)
    g.close()
