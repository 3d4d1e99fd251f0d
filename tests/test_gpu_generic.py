"""The generic path (pfc_amd/csrc/pfc_generic.hip): grid shapes outside the
tuned kernels' 2^k and 3 * 2^k lengths (the reference's FFTW plans take any
Nx and Ny, src/pfc-relax.c:166, :196-239) and field counts above four (any
U, :123-159).  Transforms: mixed radix 8/4/2/3/5/7/11/13 or Bluestein.

Checked against numpy's FP64 FFT (Gaussian smoothing, smooth.c:73-119) and
against the CPU oracle (step :1133-1247, evaluate :987-1130, relax
:1292-1312): FP64 to 1e-12 in the energy and 1e-10 in the densities, mixed to
FP32-level bounds."""
import math
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import FP64, MIXED, LATTICE_A, graphene_params, hetero_params, make_params
from pfc_amd.engine import BIN_PATH, Engine, smooth
from pfc_amd.inputs import crystal, hetero_patchwork

pytestmark = pytest.mark.gpu

DX = LATTICE_A / 8.0


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def numpy_smooth(f, dx, dy, sigma):
    """numpy's FP64 FFT with the reference's wave numbers: ky = j dky for
    j < Ny/2 (integer division), else (j - Ny) dky (src/smooth.c:93-94,
    src/pfc-relax.c:935-936).  For odd Ny that maps row (Ny - 1)/2 to
    -(Ny + 1)/2 dky, where numpy's fftfreq has +(Ny - 1)/2; the reference's
    rule is the one kept."""
    Ny, Nx = f.shape
    kx = 2 * np.pi * np.fft.rfftfreq(Nx, d=dx)
    j = np.arange(Ny)
    ky = 2 * np.pi / (Ny * dy) * np.where(j < Ny // 2, j, j - Ny)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    return np.fft.irfft2(np.fft.rfft2(f) * np.exp(-k2 / (2 * sigma ** 2)), s=f.shape)


# (Nx, Ny): radices 5/2, Bluestein in x (prime 97), Bluestein in y, both
# Bluestein, radix 3/5/7/2 mixes, a long radix-5 axis, odd small axes
SHAPES = [(100, 60), (97, 64), (64, 97), (101, 45), (105, 336), (320, 210), (4000, 8), (7, 9)]


@pytest.mark.parametrize("Nx,Ny", SHAPES)
def test_generic_smoothing_vs_numpy(gpu, Nx, Ny):
    rng = np.random.default_rng(Nx * 1000 + Ny)
    f = rng.standard_normal((Ny, Nx))
    dx, dy, sigma = 0.7, 0.9, 0.8
    ref = numpy_smooth(f, dx, dy, sigma)
    got = smooth(f, dx, dy, sigma, FP64)
    assert rel_l2(got, ref) < 1e-12
    got32 = smooth(f, dx, dy, sigma, MIXED)
    assert rel_l2(got32, ref) < 3e-6


def test_generic_graphene_100x90_vs_oracle(gpu):
    """One field on 100 x 90 (radices 5, 2; 5, 3, 2), conserved: 150 steps
    crossing the t = 0 and t = 100 refreshes, then evaluate."""
    Nx, Ny = 100, 90
    p = graphene_params(c=1.0, dt=0.5)
    f = crystal(Nx, Ny, DX, DX, -0.2, 0.2, LATTICE_A, sym=3, theta=0.2, noise=0.01, seed=3)[None]
    o = Oracle(p, Nx, Ny, DX, DX)
    o.set_fields(f)
    assert o.steps(0, 150) == 0
    with Engine(p, Nx, Ny, DX, DX, FP64) as g:
        g.set_fields(f)
        g.step(0, 150)
        assert rel_l2(g.get_fields(), o.get_fields()) < 1e-10
        fg, ng = g.evaluate()
        fo, no = o.evaluate()
        assert abs(fg - fo) <= 1e-12 * abs(fo)
        assert abs(ng[0] - no[0]) < 1e-14
        assert g.nan_step() == -1


@pytest.mark.parametrize("prec", [FP64, MIXED])
def test_generic_hetero_97x64_vs_oracle(gpu, prec):
    """Three fields on 97 x 64 (Bluestein along x): 40 steps from t = 0."""
    Nx, Ny = 97, 64
    p = hetero_params(dt=0.25, c=1.0, sigma=2.0 / 7.2552)
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=24, seed=9, noise=0.01)
    o = Oracle(p, Nx, Ny, 0.7, 0.7)
    o.set_fields(f)
    assert o.steps(0, 40) == 0
    with Engine(p, Nx, Ny, 0.7, 0.7, prec) as g:
        g.set_fields(f)
        g.step(0, 40)
        err = rel_l2(g.get_fields(), o.get_fields())
        fg, ng = g.evaluate()
        fo, no = o.evaluate()
        if prec == FP64:
            assert err < 1e-10
            assert abs(fg - fo) <= 1e-12 * abs(fo)
            assert np.allclose(ng, no, rtol=0, atol=1e-14)
        else:
            assert err < 2e-5
            assert abs(fg - fo) <= 2e-6 * abs(fo)
            # the means of FP32 densities (the k = 0 state itself is FP64)
            assert np.allclose(ng, no, rtol=0, atol=1e-7)


def many_field_params(U, c=1.0, dt=0.1):
    rng = np.random.default_rng(U)
    alpha = [[0.0] * U for _ in range(U)]
    beta = [[0.0] * U for _ in range(U)]
    gamma = [[0.0] * U for _ in range(U)]
    eps = [[0.0] * U for _ in range(U)]
    mu = [[0.0] * U for _ in range(U)]
    for u in range(U):
        alpha[u][u] = -0.3 + 0.05 * u
        beta[u][u] = 1.0 + 0.1 * u
        gamma[u][u] = 0.2 * (u % 2)
        for v in range(u + 1, U):
            alpha[u][v] = float(rng.uniform(-0.05, 0.05))
            beta[u][v] = float(rng.uniform(0.0, 0.03))
            gamma[u][v] = float(rng.uniform(-0.2, 0.2))
            eps[u][v] = float(rng.uniform(-0.3, 0.3))
            mu[u][v] = float(rng.uniform(0.95, 1.05))
    return make_params(U=U, lam=[0.01 * u for u in range(U)], alpha=alpha, beta=beta, gamma=gamma,
                       delta=[1.0] * U, epsilon=eps, sigma=0.3,
                       nu=[[1.0 + 0.01 * u] for u in range(U)], omega=[[1.0]] * U, phi=[[0.0]] * U,
                       mu=mu, c=c, dt=dt)


@pytest.mark.parametrize("U,Nx,Ny", [(5, 64, 64), (8, 48, 40)])
def test_generic_many_fields_vs_oracle(gpu, U, Nx, Ny):
    """More fields than the tuned kernels hold (U = 5 on a 64 x 64 grid the
    tuned kernels would otherwise take; U = 8, the largest accepted)."""
    p = many_field_params(U)
    rng = np.random.default_rng(U)
    f = 0.05 * rng.standard_normal((U, Ny, Nx)) + np.linspace(-0.2, 0.2, U)[:, None, None]
    o = Oracle(p, Nx, Ny, 0.8, 0.8)
    o.set_fields(f)
    assert o.steps(0, 30) == 0
    with Engine(p, Nx, Ny, 0.8, 0.8, FP64) as g:
        g.set_fields(f)
        g.step(0, 30)
        assert rel_l2(g.get_fields(), o.get_fields()) < 1e-11
        fg, ng = g.evaluate()
        fo, no = o.evaluate()
        assert abs(fg - fo) <= 1e-12 * abs(fo)
        assert np.allclose(ng, no, rtol=0, atol=1e-14)


def test_generic_nan_is_reported(gpu):
    Nx, Ny = 100, 60
    p = graphene_params()
    f = crystal(Nx, Ny, DX, DX, 0.2, -0.2, LATTICE_A, noise=0.01, seed=1)[None].copy()
    f[0, 0, 0] = np.nan
    with Engine(p, Nx, Ny, DX, DX, FP64) as g:
        g.set_fields(f)
        g.step(7, 3)
        assert g.nan_step() == 7


RUN = """I\tin.n
\tout-#.n
\t50\t[\t200\t]\t14\t17
M\t1
l\t0.0
a\t-0.25
b\t1.0
g\t1.0
d\t1.0
e\t
s\t0.2
n\t1.0
o\t1.0
p\t0.0
m\t
R\t1\t200\t{dx}\t{dy}\t1.0\t0\t1\t1
"""


def test_generic_cli_100x100(gpu, tmp_path):
    """pfc-relax on a 100 x 100 grid, as pfc-init can produce it."""
    Nx = Ny = 100
    f = crystal(Nx, Ny, DX, DX, -0.2, 0.2, LATTICE_A, sym=3, theta=0.1, noise=0.01, seed=5)
    with open(tmp_path / "in.n", "w") as fh:
        fh.write(f"{Nx} {Ny} {DX:.16e} {DX:.16e}\n")
        np.savetxt(fh, f.ravel(), fmt="%.17e")
    (tmp_path / "g.in").write_text(RUN.format(dx=repr(DX), dy=repr(DX)))
    r = subprocess.run([BIN_PATH, "g.in"], cwd=tmp_path, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [list(map(float, l.split())) for l in r.stdout.splitlines()]
    o = Oracle(graphene_params(c=1.0, dt=1.0), Nx, Ny, DX, DX)
    o.set_fields(f[None])
    olog, snaps = o.relax(200, 50, snap_steps=(200,))
    assert [int(x[0]) for x in rows] == [0, 50, 100, 150, 200]
    for row, orow in zip(rows, olog):
        assert abs(row[4] - orow[3]) <= 1e-12 * abs(orow[3])
    with open(tmp_path / "out-200.n") as fh:
        fh.readline()
        n = np.loadtxt(fh).reshape(Ny, Nx)
    assert rel_l2(n, snaps[0, 0]) < 1e-10
