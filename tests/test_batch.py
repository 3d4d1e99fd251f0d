"""Batch form (pfc_batch_*): many systems of uneven sizes, U = 1 and 3, stepped
together by one launch per chunk and pass, steps replayed from HIP graphs.
Each system must follow the CPU oracle as a one-system run does (FP64) or
stay within the mixed form's bound of the one-system engine (mixed); graph
replays must equal eager launches bit for bit."""
import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import FP64, MIXED, LATTICE_A, graphene_params, hetero_params
from pfc_amd.engine import BATCH_AUTOTUNE, BATCH_NO_GRAPH, Batch, Engine, PfcError
from pfc_amd.inputs import crystal, hetero_patchwork

pytestmark = pytest.mark.gpu

DX = LATTICE_A / 8.0

# (Nx, Ny, U): every batch length on some axis, both U, square and oblong
SYSTEMS = [(256, 256, 1), (384, 256, 3), (256, 768, 1), (512, 512, 3), (1536, 256, 1),
           (256, 2048, 3), (1024, 384, 1), (768, 768, 3), (2048, 256, 1)]


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def system(Nx, Ny, U, seed):
    if U == 1:
        p = graphene_params()
        f = crystal(Nx, Ny, DX, DX, 0.2, -0.2, LATTICE_A, sym=3, theta=0.1 * seed, noise=0.01,
                    seed=seed)[None]
        return p, f, DX, DX
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=64, seed=seed)
    return p, f, 0.7, 0.7


def engines(precision, systems=SYSTEMS):
    out = []
    for k, (Nx, Ny, U) in enumerate(systems):
        p, f, dx, dy = system(Nx, Ny, U, k + 1)
        e = Engine(p, Nx, Ny, dx, dy, precision)
        e.set_fields(f)
        out.append((e, p, f, dx, dy))
    return out


def test_fp64_batch_follows_oracle(gpu):
    es = engines(FP64)
    with Batch([e for e, *_ in es], BATCH_AUTOTUNE) as b:
        b.step(95, 8)          # crosses the refresh step t = 100
        for (e, p, f, dx, dy), (Nx, Ny, U) in zip(es, SYSTEMS):
            o = Oracle(p, Nx, Ny, dx, dy)
            o.set_fields(f)
            assert o.steps(95, 8) == 0
            err = rel_l2(e.get_fields(), o.get_fields())
            assert err < 1e-12, (Nx, Ny, U, err)
            fo, _ = o.evaluate()
            fg, _ = e.evaluate()
            assert abs(fg - fo) <= 1e-12 * abs(fo), (Nx, Ny, U)
        for Nx, Ny, U in SYSTEMS:
            assert b.shape(0, Nx, U) in candidates(Nx, U)
            assert b.shape(1, Ny, U) in candidates(Ny, U)


def candidates(n, U):
    """points per thread the autotuner chooses between"""
    if n & (n - 1) == 0:
        return (8, 16) if U == 1 else (8, 4)
    return (12, 6) if U == 1 else (6, 12)


@pytest.mark.parametrize("k", [0, 1, 2])
@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_every_kernel_shape_matches_one_system_engines(gpu, monkeypatch, k, precision):
    """Each candidate tile of the class table (two radices, and half the rows
    or columns per block), forced for every system, across the t = 100
    refresh, against pfc_step on each system."""
    monkeypatch.setenv("PFC_BATCH_CANDIDATE", str(k))
    es, ref = engines(precision), engines(precision)
    with Batch([e for e, *_ in es]) as b:
        b.step(95, 8)
        if k < 2:
            for Nx, Ny, U in SYSTEMS:
                assert b.shape(0, Nx, U) == candidates(Nx, U)[k]
                assert b.shape(1, Ny, U) == candidates(Ny, U)[k]
    for (e, *_), (r, *_), s in zip(es, ref, SYSTEMS):
        r.step(95, 8)
        err = rel_l2(e.get_fields(), r.get_fields())
        assert err < (1e-12 if precision == FP64 else 2e-6), (s, k, err)


def test_mixed_batch_matches_one_system_engines(gpu):
    es = engines(MIXED)
    ref = engines(MIXED)
    with Batch([e for e, *_ in es]) as b:
        b.step(0, 12)
    for (e, *_), (r, *_), s in zip(es, ref, SYSTEMS):
        r.step(0, 12)
        err = rel_l2(e.get_fields(), r.get_fields())
        assert err < 2e-6, (s, err)


@pytest.mark.parametrize("precision", [FP64, MIXED])
def test_graph_replays_equal_eager_launches(gpu, precision):
    """Replays of the captured steps, with evaluates (eager work) between them,
    give the same bits as launching each step's kernels directly."""
    sub = SYSTEMS[:5]
    a, b = engines(precision, sub), engines(precision, sub)
    with Batch([e for e, *_ in a]) as ga, Batch([e for e, *_ in b], BATCH_NO_GRAPH) as gb:
        # single steps, a 20-step block starting with the refresh at 100, a
        # plain block, and a single refresh step at 200
        for t, n in ((97, 3), (100, 40), (140, 3), (143, 60)):
            ga.step(t, n)
            gb.step(t, n)
            for (x, *_), (y, *_) in zip(a, b):
                assert np.array_equal(x.get_fields(), y.get_fields())
                assert x.evaluate()[0] == y.evaluate()[0]


def test_fp64_blocks_match_one_system_engines(gpu):
    """45 steps from t = 0: two captured 20-step blocks (the first starting
    with a refresh) and five single steps, against pfc_step on each system."""
    sub = SYSTEMS[:6]
    es, ref = engines(FP64, sub), engines(FP64, sub)
    with Batch([e for e, *_ in es]) as b:
        b.step(0, 45)
    for (e, *_), (r, *_), s in zip(es, ref, sub):
        r.step(0, 45)
        err = rel_l2(e.get_fields(), r.get_fields())
        assert err < 1e-12, (s, err)


def test_nan_is_reported_per_system(gpu):
    es = engines(FP64, SYSTEMS[:3])
    bad = es[1][2].copy()
    bad[0, 0, 0] = np.nan
    es[1][0].set_fields(bad)
    with Batch([e for e, *_ in es]) as b:
        b.step(10, 3)
    assert es[0][0].nan_step() == -1
    assert es[1][0].nan_step() == 10
    assert es[2][0].nan_step() == -1


def test_rejected_batches(gpu):
    p = graphene_params()
    with Engine(p, 64, 64, DX, DX) as small, Engine(p, 256, 256, DX, DX) as a, \
            Engine(p, 256, 256, DX, DX, MIXED) as m:
        with pytest.raises(PfcError, match="-4"):
            Batch([small])
        with pytest.raises(PfcError, match="-1"):
            Batch([a, m])
        with Batch([a]) as b:
            with pytest.raises(PfcError, match="-5"):
                b.step(0, 1)     # no fields yet
