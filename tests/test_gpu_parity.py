"""GPU parity: libpfcgpu (through its C ABI) against the CPU oracle on the same
seeded inputs.  FP64 bar: relative L2 density error <= 1e-10 and free energy
per area to <= 1e-12 relative; mixed precision: the stated FP32-level bounds.
"""
import math

import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import FP64, MIXED, LATTICE_A, graphene_params, hetero_params, make_params
from pfc_amd.engine import Engine
from pfc_amd.inputs import crystal, hetero_patchwork

pytestmark = pytest.mark.gpu

DX = LATTICE_A / 8.0


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def graphene_state(Nx, Ny, seed=1, nbar=0.2, amp=-0.2, noise=0.01):
    return crystal(Nx, Ny, DX, DX, nbar, amp, LATTICE_A, sym=3, theta=0.3, noise=noise,
                   seed=seed)[None]


def two_field_params(c=0.0):
    return make_params(
        U=2, lam=[0.01, -0.02], alpha=[[-0.3, 0.05], [0, -0.2]], beta=[[1.0, 0.1], [0, 1.1]],
        gamma=[[0.8, 0.2], [0, 0.6]], delta=[1.0, 0.9], epsilon=[[0, -0.3], [0, 0]], sigma=0.25,
        nu=[[1.0], [1.05]], omega=[[1.0], [1.0]], phi=[[0.0], [0.0]], mu=[[0, 1.02], [0, 0]],
        c=c, dt=0.5)


def run_pair(params, fields, Nx, Ny, dx, dy, t0, nsteps, precision=FP64):
    o = Oracle(params, Nx, Ny, dx, dy)
    o.set_fields(fields)
    g = Engine(params, Nx, Ny, dx, dy, precision)
    g.set_fields(fields)
    assert o.steps(t0, nsteps) == 0
    g.step(t0, nsteps)
    return o, g


@pytest.mark.parametrize("Nx,Ny", [(64, 64), (128, 64), (64, 256), (256, 128), (512, 512),
                                   (1024, 64), (64, 1024), (2048, 64)])
def test_roundtrip_and_one_step_sizes(gpu, Nx, Ny):
    p = graphene_params()
    f = graphene_state(Nx, Ny)
    with Engine(p, Nx, Ny, DX, DX) as g:
        g.set_fields(f)
        back = g.get_fields()
        assert rel_l2(back, f) < 1e-14
        o = Oracle(p, Nx, Ny, DX, DX)
        o.set_fields(f)
        o.step(1)
        g.step(1, 1)
        assert rel_l2(g.get_fields(), o.get_fields()) < 1e-13


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("Nx,Ny", [(4096, 64), (64, 4096), (8192, 64), (64, 8192), (2048, 128)])
def test_wide_and_tall_grids(gpu, Nx, Ny, precision):
    """Every row/column kernel configuration of the large sizes, 3 fields."""
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=48, seed=11)
    o, g = run_pair(p, f, Nx, Ny, 0.7, 0.7, 0, 3, precision=precision)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < (1e-13 if precision == FP64 else 2e-6), err


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("U", [1, 2])
@pytest.mark.parametrize("Nx,Ny", [(4096, 64), (2048, 128)])
def test_long_rows_one_and_two_fields(gpu, Nx, Ny, U, precision):
    """The 4096- and 2048-point row kernels (2 or 4 FP32 rows per block, two
    blocks per CU) with one and two fields, across the refresh step t = 100."""
    if U == 1:
        p, f = graphene_params(), graphene_state(Nx, Ny, seed=21)
    else:
        p = two_field_params(c=1.0)
        f = np.concatenate([graphene_state(Nx, Ny, seed=22), graphene_state(Nx, Ny, seed=23)])
    o, g = run_pair(p, f, Nx, Ny, DX, DX, 98, 4, precision=precision)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < (1e-13 if precision == FP64 else 2e-6), err


@pytest.mark.parametrize("precision", [FP64, MIXED])
@pytest.mark.parametrize("Nx,Ny", [(384, 384), (768, 384), (384, 1536), (1536, 768), (768, 64),
                                   (64, 1536)])
def test_three_times_power_of_two_grids(gpu, Nx, Ny, precision):
    """FFT lengths 3 * 2^k (a radix-12 first stage) in the row and column
    passes, 3 fields, across the refresh step t = 100, then evaluate."""
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=48, seed=13)
    o, g = run_pair(p, f, Nx, Ny, 0.7, 0.7, 98, 4, precision=precision)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < (1e-13 if precision == FP64 else 2e-6), err
    fo, no = o.evaluate()
    fg, ng = g.evaluate()
    assert abs(fg - fo) <= (1e-12 if precision == FP64 else 1e-6) * abs(fo), (fg, fo)


def test_mixed_k0_mode_is_fp64(gpu):
    """The mixed form carries the k = 0 mode in FP64.  A uniform field of a
    linear non-conserved model (gamma = delta = lambda = 0) decays as
    n <- A0 n with A0 = exp(-(alpha + beta) dt) = exp(-0.001): after 1000 steps
    the mean must be 0.3 A0^1000 to the FP32 resolution of the output; an FP32
    recursion would drift by ~1000 roundings of A0 (~1e-5)."""
    p = graphene_params(alpha=-0.999, gamma=0.0, delta=0.0, c=0.0, dt=1.0)
    a = -0.999 + 1.0
    with Engine(p, 64, 64, DX, DX, MIXED) as g:
        g.set_fields(np.full((1, 64, 64), 0.3))
        g.step(1, 1000)
        mean = float(g.get_fields().mean())
    expected = 0.3 * math.exp(-a) ** 1000
    assert abs(mean - expected) <= 2e-7 * expected, (mean, expected)


@pytest.mark.parametrize("c", [0.0, 1.0])
def test_graphene_steps_fp64(gpu, c):
    Nx, Ny = 128, 128
    p = graphene_params(c=c)
    f = graphene_state(Nx, Ny)
    o, g = run_pair(p, f, Nx, Ny, DX, DX, 1, 50)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < 1e-12, err


def test_refresh_crossing_fp64(gpu):
    """Steps 95..205 cross the t%100 refresh points of the reference."""
    Nx, Ny = 64, 128
    p = graphene_params()
    f = graphene_state(Nx, Ny, seed=3)
    o, g = run_pair(p, f, Nx, Ny, DX, 0.9 * DX, 95, 111)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < 1e-11, err


def test_two_fields_fp64(gpu):
    Nx, Ny = 128, 64
    p = two_field_params()
    rng = np.random.default_rng(5)
    f = np.stack([graphene_state(Nx, Ny, seed=5)[0], 0.1 + 0.05 * rng.standard_normal((Ny, Nx))])
    o, g = run_pair(p, f, Nx, Ny, DX, DX, 0, 30)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < 1e-12, err


@pytest.mark.parametrize("c", [1.0, 0.0])
def test_hetero_three_fields_fp64(gpu, c):
    Nx, Ny = 128, 128
    p = hetero_params(c=c)
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=48, seed=2)
    o, g = run_pair(p, f, Nx, Ny, 0.7, 0.7, 0, 40)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < 1e-12, err
    if c:
        # conserved dynamics keep every mean density
        assert np.allclose(g.get_fields().mean(axis=(1, 2)), f.mean(axis=(1, 2)), rtol=0, atol=1e-13)


@pytest.mark.parametrize("U", [1, 2, 3])
def test_evaluate_fp64(gpu, U):
    Nx, Ny = 128, 64
    if U == 1:
        p = graphene_params()
        f = graphene_state(Nx, Ny)
    elif U == 2:
        p = two_field_params()
        rng = np.random.default_rng(8)
        f = np.stack([graphene_state(Nx, Ny, seed=8)[0], 0.1 + 0.05 * rng.standard_normal((Ny, Nx))])
    else:
        p = hetero_params()
        f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=40, seed=4)
    o, g = run_pair(p, f, Nx, Ny, 0.7, 0.75, 0, 7)
    fo, no = o.evaluate()
    fg, ng = g.evaluate()
    assert abs(fg - fo) <= 1e-12 * abs(fo), (fg, fo)
    assert np.allclose(ng, no, rtol=1e-12, atol=1e-15)
    ef = g.get_energy()
    assert rel_l2(ef, o.get_energy()) < 1e-12
    # evaluate re-synchronises N; stepping on must still agree
    o.steps(7, 5)
    g.step(7, 5)
    assert rel_l2(g.get_fields(), o.get_fields()) < 1e-12


def test_nan_detection(gpu):
    Nx, Ny = 64, 64
    p = graphene_params()
    f = graphene_state(Nx, Ny)
    f[0, 0, 0] = math.nan
    with Engine(p, Nx, Ny, DX, DX) as g:
        g.set_fields(f)
        assert g.nan_step() == -1
        g.step(10, 3)
        assert g.nan_step() == 10


def test_mixed_precision_error(gpu):
    Nx, Ny = 256, 256
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=64, seed=9)
    o, g = run_pair(p, f, Nx, Ny, 0.7, 0.7, 0, 50, precision=MIXED)
    err = rel_l2(g.get_fields(), o.get_fields())
    assert err < 1e-5, err
    fo, no = o.evaluate()
    fg, ng = g.evaluate()
    assert abs(fg - fo) <= 1e-6 * abs(fo), (fg, fo)
    # conserved: the k = 0 modes carry the mean densities
    assert np.allclose(ng, no, rtol=1e-6, atol=1e-7)
