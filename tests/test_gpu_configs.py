"""The measured configurations at their full sizes on the GPU: the 512^2
graphene crystal run end to end through the pfc-relax binary, the 2048^2
graphene polycrystal in FP64, and the 4096^2 and 8192^2 three-field
graphene/hBN heterostructures in both forms.  Each runs a few steps against the CPU oracle
(which is too slow for the full runs) and then longer through properties
that do not depend on the size: conservation of the mean, decrease of the
free energy, and the mixed form's distance from FP64."""
import math

import numpy as np
import pytest

from oracle_lib import Oracle, optimize_with
from pfc_amd import FP64, MIXED, LATTICE_A, graphene_params, hetero_params
from pfc_amd.engine import Engine
from pfc_amd.inputs import hetero_patchwork, polycrystal

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_polycrystal_2048_fp64(gpu):
    """2048^2, 64 seeded grains at random orientations with 2-site boundaries
    at the mean density, conserved dynamics, dt = 1."""
    N = 2048
    dx = LATTICE_A / 8.0
    p = graphene_params(c=1.0, dt=1.0)
    f = polycrystal(N, N, dx, dx, 64, 0.2, -0.2, LATTICE_A, sym=3, noise=0.01, boundary=2,
                    seed=64)[None]
    o = Oracle(p, N, N, dx, dx)
    o.set_fields(f)
    assert o.steps(0, 3) == 0
    with Engine(p, N, N, dx, dx, FP64) as g:
        g.set_fields(f)
        g.step(0, 3)
        assert rel_l2(g.get_fields(), o.get_fields()) < 1e-12
        fo, no = o.evaluate()
        fg, ng = g.evaluate()
        assert abs(fg - fo) <= 1e-12 * abs(fo)
        energies = [fg]
        for t in (3, 103, 203):
            g.step(t, 100)
            fe, ne = g.evaluate()
            energies.append(fe)
            assert abs(ne[0] - ng[0]) <= 1e-13 * abs(ng[0])      # the mean is conserved
        assert all(b < a for a, b in zip(energies, energies[1:])), energies
        assert g.nan_step() == -1


def test_heterostructure_4096(gpu):
    """4096^2, three fields: the benchmark's workload (hetero.in parameters,
    Gaussian width a/2, seeded 50/50 patchwork of ~256-point domains, noise
    +-0.01, conserved, dt = 0.25).  FP64 and mixed against the oracle over 2
    steps (the first a refresh step), evaluate against the oracle, then mixed
    against FP64 over 100 steps.  This covers the padded column pitch of the
    full grid and the row pass's XCD block remap at the benchmarked size."""
    N = 4096
    p = hetero_params(dt=0.25, c=1.0, sigma=2.0 / 7.2552)
    f = hetero_patchwork(N, N, 0.7, 0.7, domain=256, seed=2024, noise=0.01)
    o = Oracle(p, N, N, 0.7, 0.7)
    o.set_fields(f)
    assert o.steps(0, 2) == 0
    ref = o.get_fields()
    fo, no = o.evaluate()
    del o
    with Engine(p, N, N, 0.7, 0.7, FP64) as d, Engine(p, N, N, 0.7, 0.7, MIXED) as m:
        d.set_fields(f)
        m.set_fields(f)
        d.step(0, 2)
        m.step(0, 2)
        ed, em = rel_l2(d.get_fields(), ref), rel_l2(m.get_fields(), ref)
        fd, nd = d.evaluate()
        fm, nm = m.evaluate()
        print(f"4096^2 U=3 after 2 steps: density FP64 {ed:.2e} mixed {em:.2e}; "
              f"energy FP64 {abs(fd - fo) / abs(fo):.2e} mixed {abs(fm - fo) / abs(fo):.2e}")
        assert ed < 1e-12
        assert em < 2e-6
        assert abs(fd - fo) <= 1e-12 * abs(fo)
        assert np.allclose(nd, no, rtol=0, atol=1e-15)
        assert abs(fm - fo) <= 1e-6 * abs(fo)
        d.step(2, 98)
        m.step(2, 98)
        fd, nd = d.evaluate()
        fm, nm = m.evaluate()
        assert abs(fm - fd) <= 1e-6 * abs(fd), (fm, fd)
        assert np.allclose(nm, nd, rtol=1e-7, atol=0)
        assert rel_l2(m.get_fields(), d.get_fields()) < 1e-3
        assert d.nan_step() == -1 and m.nan_step() == -1


@pytest.fixture(scope="module")
def c4():
    """The 8192^2 heterostructure start (three fields, 128 seeded domains,
    dt = 0.25) and the oracle after its first 2 steps."""
    N = 8192
    p = hetero_params(dt=0.25, c=1.0, sigma=2.0 / 7.2552)
    f = hetero_patchwork(N, N, 0.7, 0.7, n_domains=128, seed=8192, noise=0.01)
    o = Oracle(p, N, N, 0.7, 0.7)
    o.set_fields(f)
    assert o.steps(0, 2) == 0
    ref = o.get_fields()
    fo, _ = o.evaluate()
    yield dict(N=N, p=p, f=f, o=o, ref=ref, fo=fo)
    o.close()


def test_heterostructure_8192(gpu, c4):
    """8192^2, three fields, 128 seeded domains, dt = 0.25: FP64 and mixed
    against the oracle for 2 steps, then mixed against FP64 over 100 steps
    (energy per area within 1e-6).  The hetero.in
    parameters diverge at dt = 0.5 (the CPU oracle reaches NaN within 20
    steps at 256^2), so the demo's dt = 0.25 is used."""
    N, p, f, ref, fo = c4["N"], c4["p"], c4["f"], c4["ref"], c4["fo"]
    with Engine(p, N, N, 0.7, 0.7, FP64) as d, Engine(p, N, N, 0.7, 0.7, MIXED) as m:
        d.set_fields(f)
        m.set_fields(f)
        d.step(0, 2)
        m.step(0, 2)
        assert rel_l2(d.get_fields(), ref) < 1e-12
        assert rel_l2(m.get_fields(), ref) < 2e-6
        fd, nd = d.evaluate()
        assert abs(fd - fo) <= 1e-12 * abs(fo)
        d.step(2, 98)
        m.step(2, 98)
        fd, nd = d.evaluate()
        fm, nm = m.evaluate()
        assert abs(fm - fd) <= 1e-6 * abs(fd), (fm, fd)
        assert np.allclose(nm, nd, rtol=1e-7, atol=0)
        assert rel_l2(m.get_fields(), d.get_fields()) < 1e-3
        assert d.nan_step() == -1 and m.nan_step() == -1


def test_c4_box_search_8192(gpu, c4):
    """The box search of optimize (src/pfc-relax.c:1250-1289) at 8192^2 after
    the first 2 steps, along x with the reference's first sampling step
    d = 1e-4: the energies at dx - d, dx and dx + d from the GPU (sums-only
    evaluates after pfc_set_box) against the oracle's to 1e-12, and the box
    the search picks from each set of energies.  The parabola's vertex
    divides the energies' differences by their second difference, so the
    boxes may differ by up to d * |energy difference| / |second difference|
    (a few times that), which the test states and checks."""
    N, p, f, o = c4["N"], c4["p"], c4["f"], c4["o"]
    d = 1e-4
    boxes = [(0.7 - d, 0.7), (0.7 + d, 0.7)]
    fo = {(0.7, 0.7): c4["fo"]}
    for b in boxes:
        o.set_box(*b)
        fo[b] = o.evaluate()[0]
    o.set_box(0.7, 0.7)
    with Engine(p, N, N, 0.7, 0.7, FP64) as g:
        g.set_fields(f)
        g.step(0, 2)
        fg = {}
        for b in [(0.7, 0.7)] + boxes:
            g.set_box(*b)
            fg[b] = g.evaluate(density=False)[0]
        g.set_box(0.7, 0.7)
    errs = {b: abs(fg[b] - fo[b]) / abs(fo[b]) for b in fo}
    assert max(errs.values()) <= 1e-12, errs
    lo, mid, hi = (0.7 - d, 0.7), (0.7, 0.7), (0.7 + d, 0.7)
    xo, _, _ = optimize_with(N, N, 0.7, 0.7, d, 1, 0, lambda a, b: fo[(a, b)])
    xg, _, _ = optimize_with(N, N, 0.7, 0.7, d, 1, 0, lambda a, b: fg[(a, b)])
    second = abs(fo[lo] - 2 * fo[mid] + fo[hi])
    bound = 4 * d * max(abs(fg[b] - fo[b]) for b in fo) / second + 1e-16
    print(f"8192^2 box search: energies {max(errs.values()):.2e}, dx oracle {xo!r} GPU {xg!r}, "
          f"difference {abs(xg - xo):.2e} (bound {bound:.2e}), moved {xo - 0.7:.3e}")
    assert xo != 0.7
    assert abs(xg - xo) <= bound, (xg, xo, bound)
