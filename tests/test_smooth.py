"""Gaussian smoothing (the reference's smooth tool, src/smooth.c:17-151):
the command's contract on the CPU, and on the GPU the field against numpy's
FP64 FFT and against the closed form of a smoothed cosine."""
import math
import os
import subprocess

import numpy as np
import pytest

from pfc_amd import FP64, MIXED
from pfc_amd import engine


def run(args, cwd, env=None):
    e = dict(os.environ)
    for k in ("PMI_RANK", "OMPI_COMM_WORLD_RANK", "PMIX_RANK"):
        e.pop(k, None)
    e.update(env or {})
    r = subprocess.run([engine.SMOOTH_PATH, *args], cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, env=e, timeout=120)
    return r.returncode, r.stdout, r.stderr


def smooth_numpy(f, dx, dy, sigma):
    """src/smooth.c:82-117 in numpy: rfft2, times exp(-k^2/(2 sigma^2)), irfft2
    (the division by sum(C2R(G)) = 1 is left out)."""
    Ny, Nx = f.shape
    kx = 2 * math.pi * np.fft.rfftfreq(Nx, d=dx)
    ky = 2 * math.pi * np.fft.fftfreq(Ny, d=dy)
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    return np.fft.irfft2(np.fft.rfft2(f) * np.exp(-0.5 / sigma / sigma * k2), s=f.shape)


def test_usage_and_errors(tmp_path):
    rc, out, _ = run([], tmp_path)
    assert rc == 0 and out.startswith("This is a tool for smoothing PFC data.")
    rc, out, _ = run(["a.n", "b.eta"], tmp_path)
    assert (rc, out) == (1, "Error: Invalid syntax.\n")
    rc, out, _ = run(["none.n", "b.eta", "0.2"], tmp_path)
    assert (rc, out) == (1, "Input file not found!\n")
    rc, out, _ = run(["none.n", "b.eta", "0.2"], tmp_path, env={"PMI_RANK": "1"})
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
    rc, _, err = run(["a.n", "b.eta", "0.2"], tmp_path)
    assert rc == 42 and "pfc_smooth failed" in err


@pytest.mark.gpu
@pytest.mark.parametrize("Nx,Ny", [(64, 64), (128, 64), (64, 256), (512, 512), (384, 768)])
def test_matches_numpy_fp64(gpu, Nx, Ny):
    rng = np.random.default_rng(Nx + Ny)
    f = rng.standard_normal((Ny, Nx))
    dx, dy, sigma = 0.7, 0.65, 0.2
    got = engine.smooth(f, dx, dy, sigma, FP64)
    ref = smooth_numpy(f, dx, dy, sigma)
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-13


@pytest.mark.gpu
@pytest.mark.parametrize("prec,tol", [(FP64, 1e-13), (MIXED, 3e-6)])
def test_cosine_closed_form(gpu, prec, tol):
    """f = 0.3 + cos(kx x) cos(ky y) with kx, ky on the grid's frequencies:
    smoothing keeps the mean and scales the mode by exp(-(kx^2+ky^2)/(2 sigma^2))."""
    Nx, Ny, dx, dy, sigma = 1024, 512, 0.7, 0.7, 0.2
    kx = 2 * math.pi * 3 / (Nx * dx)
    ky = 2 * math.pi * 5 / (Ny * dy)
    x = np.arange(Nx)[None, :] * dx
    y = np.arange(Ny)[:, None] * dy
    mode = np.cos(kx * x) * np.cos(ky * y)
    got = engine.smooth(0.3 + mode, dx, dy, sigma, prec)
    ref = 0.3 + math.exp(-(kx * kx + ky * ky) / (2 * sigma * sigma)) * mode
    assert np.max(np.abs(got - ref)) < tol


@pytest.mark.gpu
def test_command_output(gpu, tmp_path):
    rng = np.random.default_rng(7)
    f = rng.standard_normal((128, 256))
    write_field(tmp_path / "in.n", f, 0.7, 0.7)
    rc, out, err = run(["in.n", "out.eta", "0.2"], tmp_path)
    assert rc == 0, err
    lines = (tmp_path / "out.eta").read_text().splitlines()
    assert lines[0] == "256 128 0.700000 0.700000"
    assert all("e" in l and len(l.split("e")[0].split(".")[1]) == 12 for l in lines[1:10])
    got = np.array([float(v) for v in lines[1:]]).reshape(128, 256)
    ref = smooth_numpy(f, 0.7, 0.7, 0.2)
    assert np.max(np.abs(got - ref)) < 1e-12 * np.max(np.abs(ref)) + 1e-15
