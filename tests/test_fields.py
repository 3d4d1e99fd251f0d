"""Orientation and deformation fields (the reference's fields tool,
src/fields.c:104-424).  On the CPU the numpy restatement (oracle/fields_ref.py)
is pinned by closed forms and the command's contract is checked; on the GPU
pfc_fields is compared with the restatement on random densities and with the
closed forms directly."""
import cmath
import math
import os
import subprocess

import numpy as np
import pytest

from oracle import fields_ref as fr
from pfc_amd import engine


def plane_wave(W, H, mx, my, theta):
    """cos(k0 . r) with k0 = (2 pi mx / (W dx), 2 pi my / (H dy)) of length 1
    at angle theta: dx, dy are chosen to make it so."""
    dx = 2 * math.pi * mx / (W * math.cos(theta))
    dy = 2 * math.pi * my / (H * math.sin(theta)) if my else 0.7
    j, i = np.meshgrid(np.arange(H), np.arange(W), indexing="ij")
    ph = 2 * math.pi * (mx * i / W + my * j / H)
    return np.cos(ph), ph, dx, dy


def phi_closed_form(p, ph, fold, sk, sb, theta):
    """For n = cos(t), t = k0 . r, |k0| = 1: n - min = c0 + cos t; the kernel
    passes k = 0 with exp(-1/(2 sk^2)) and +-k0 with E = e^{i fold theta}
    (even fold), so the product is c0^2 e^a + E/2 + c0 (e^a + E) cos t
    + (E/2) cos 2t; smoothing scales mode m by exp(-m^2/(2 sb^2))."""
    c0 = -p.min()
    ea = math.exp(-0.5 / sk ** 2)
    E = cmath.exp(1j * fold * theta)
    g1, g2 = math.exp(-0.5 / sb ** 2), math.exp(-2.0 / sb ** 2)
    z = c0 * c0 * ea + E / 2 + c0 * (ea + E) * g1 * np.cos(ph) + (E / 2) * g2 * np.cos(2 * ph)
    return z / np.abs(z).max(), (c0 * (ea + E) * g1, E * g2)


def chi_closed_form(ph, B, D, scales):
    """power 2: |grad z|^2 = |B sin t + D sin 2t|^2 (|k0| = 1), a cosine series
    g0 + g1 cos t + ... + g4 cos 4t; scale n multiplies mode m by
    exp(-n^2 m^2 / 2); each scale is divided by its maximum over the grid."""
    bb, dd, bd = abs(B) ** 2, abs(D) ** 2, (B * D.conjugate()).real
    g = [bb / 2 + dd / 2, bd, -bb / 2, -bd, -dd / 2]
    chi = np.zeros_like(ph)
    for n in scales:
        lev = sum(g[m] * math.exp(-0.5 * n * n * m * m) * np.cos(m * ph) for m in range(5))
        chi += lev / lev.max()
    return chi


CASES = [(0.0, 16, 0), (math.pi / 6, 16, 8)]   # theta, mx, my on a 256^2 grid


@pytest.mark.parametrize("theta,mx,my", CASES)
@pytest.mark.parametrize("fold", [2, 6])
def test_oracle_closed_form(theta, mx, my, fold):
    W = H = 256
    p, ph, dx, dy = plane_wave(W, H, mx, my, theta)
    sk, sb = 0.3, 0.5
    phi, kept = fr.orientation(p, dx, dy, fold, sk, sb)
    ref, (B, D) = phi_closed_form(p, ph, fold, sk, sb, theta)
    assert np.max(np.abs(phi - ref)) < 1e-12
    sc = fr.scales(W, H, dx, dy)
    assert sc == [2, 4]
    # the kept spectrum is that of z before the division by max |z|
    chi = fr.deformation(kept, dx, dy, 2.0)
    assert np.max(np.abs(chi - chi_closed_form(ph, B, D, sc))) < 1e-10


def run(args, cwd, env=None):
    e = dict(os.environ)
    for k in ("PMI_RANK", "OMPI_COMM_WORLD_RANK", "PMIX_RANK"):
        e.pop(k, None)
    e.update(env or {})
    r = subprocess.run([engine.FIELDS_PATH, *args], cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, env=e, timeout=120)
    return r.returncode, r.stdout, r.stderr


def test_command_contract(tmp_path):
    rc, out, _ = run([], tmp_path)
    assert rc == 0 and "orientation field" in out
    rc, out, _ = run(["a.n", "b.phi", "6", "0.3"], tmp_path)
    assert (rc, out) == (1, "Error: Invalid syntax.\n")
    rc, out, _ = run(["a.n", "b.phi", "6", "0.3", "0.5", "1.0"], tmp_path)
    assert (rc, out) == (1, "Error: Invalid syntax.\n")
    rc, out, _ = run(["none.n", "b.phi", "6", "0.3", "0.5"], tmp_path)
    assert (rc, out) == (1, "Error: Input file not found.\n")
    (tmp_path / "bad.n").write_text("x y\n")
    rc, out, _ = run(["bad.n", "b.phi", "6", "0.3", "0.5"], tmp_path)
    assert (rc, out) == (1, "Error: Invalid data.\n")
    rc, out, _ = run(["none.n", "b.phi", "6", "0.3", "0.5"], tmp_path, env={"PMI_RANK": "2"})
    assert (rc, out) == (0, "")


def write_field(path, f, dx, dy):
    Ny, Nx = f.shape
    with open(path, "w") as fh:
        fh.write(f"{Nx} {Ny} {dx:.6f} {dy:.6f}\n")
        np.savetxt(fh, f.ravel(), fmt="%.17e")


def test_without_gpu_reaches_the_device(tmp_path):
    from conftest import gpu_available
    if gpu_available():
        pytest.skip("GPU present: the GPU tests below cover the run")
    write_field(tmp_path / "a.n", np.zeros((64, 64)), 0.7, 0.7)
    rc, _, err = run(["a.n", "b.phi", "6", "0.3", "0.5"], tmp_path)
    assert rc == 42 and "pfc_fields failed" in err


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.gpu
@pytest.mark.parametrize("W,H", [(64, 64), (128, 64), (64, 256), (512, 512), (1024, 2048), (384, 768),
                                 (1536, 384),
                                 # sides only the generic transforms take: radix 5/3, a
                                 # prime (Bluestein), 640 x 480, an odd side
                                 (100, 90), (97, 64), (640, 480), (1000, 125)])
@pytest.mark.parametrize("power", [1.0, 2.0])
def test_matches_oracle(gpu, W, H, power):
    rng = np.random.default_rng(W * 7 + H)
    # a lattice-like field with noise, wave number ~1
    x = np.arange(W)[None, :] * 0.9
    y = np.arange(H)[:, None] * 0.8
    p = np.cos(x) * np.cos(0.8 * y) + 0.2 * rng.standard_normal((H, W))
    phi, chi = engine.fields(p, 0.9, 0.8, 6, 0.3, 0.5, power)
    rphi, kept = fr.orientation(p, 0.9, 0.8, 6, 0.3, 0.5)
    assert rel(phi, rphi) < 1e-12
    rchi = fr.deformation(kept, 0.9, 0.8, power)
    if fr.scales(W, H, 0.9, 0.8):
        assert rel(chi, rchi) < 1e-11
    else:
        assert not chi.any()


@pytest.mark.gpu
@pytest.mark.parametrize("theta,mx,my", CASES)
def test_closed_form(gpu, theta, mx, my):
    W = H = 256
    p, ph, dx, dy = plane_wave(W, H, mx, my, theta)
    phi, chi = engine.fields(p, dx, dy, 6, 0.3, 0.5, 2.0)
    ref, (B, D) = phi_closed_form(p, ph, 6, 0.3, 0.5, theta)
    assert np.max(np.abs(phi - ref)) < 1e-12
    assert np.max(np.abs(chi - chi_closed_form(ph, B, D, [2, 4]))) < 1e-10


@pytest.mark.gpu
def test_command_files(gpu, tmp_path):
    W, H = 128, 64
    p, ph, dx, dy = plane_wave(W, H, 8, 0, 0.0)
    write_field(tmp_path / "a.n", p, dx, dy)
    rc, out, err = run(["a.n", "o.phi", "o.chi", "6", "0.3", "0.5", "1.0"], tmp_path)
    assert (rc, out) == (0, ""), err
    lines = (tmp_path / "o.phi").read_text().splitlines()
    assert lines[0] == f"{W} {H} {dx:.6f} {dy:.6f}"
    vals = np.array([[float(v) for v in l.split()] for l in lines[1:]])
    assert vals.shape == (W * H, 2)
    dxr, dyr = round(dx, 6), round(dy, 6)
    rphi, kept = fr.orientation(p, dxr, dyr, 6, 0.3, 0.5)
    assert np.max(np.abs(vals[:, 0] + 1j * vals[:, 1] - rphi.ravel())) < 1e-6
    chi = np.array([float(l) for l in (tmp_path / "o.chi").read_text().splitlines()[1:]])
    assert chi.size == W * H
    assert (tmp_path / "o.chi").read_text().splitlines()[1] == f"{chi[0]:e}"
