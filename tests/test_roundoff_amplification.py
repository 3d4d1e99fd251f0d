"""How far the box search (optimize, src/pfc-relax.c:1250-1289) carries a
round-off-level difference, on the CPU: the oracle built twice from the same
source, once unfused (-ffp-contract=off, the shipped build) and once with FMA
contraction (-mfma -ffp-contract=fast), runs the graphene case of
tests/test_gpu_cli.py with and without box searches.

Without searches the two builds agree to ~1e-16 in fave and ~1e-14 in the
density.  With a search every 50 steps they differ by more than 1e-12 in the
box and in fave: each search fits a parabola through energies whose second
difference is ~1e-9 of the energy, so a round-off difference in the energies
moves the chosen box by ~1e-12 relative, and later steps carry it.  Two
correct FP64 builds therefore cannot meet a 1e-12 bar on this path; the GPU
test's tolerance for it (tests/test_gpu_cli.py) rests on this measurement.
"""
import os
import subprocess

import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import LATTICE_A, graphene_params
from pfc_amd.inputs import crystal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _has_fma():
    try:
        with open("/proc/cpuinfo") as f:
            return " fma " in f.read()
    except OSError:
        return False


@pytest.mark.skipif(not _has_fma(), reason="host CPU without FMA")
def test_box_search_amplifies_fma_roundoff(tmp_path):
    so = str(tmp_path / "liboracle_fma.so")
    subprocess.run(["gcc", "-O2", "-fopenmp", "-mfma", "-ffp-contract=fast", "-fPIC", "-std=c99",
                    "-shared", "-o", so, os.path.join(ROOT, "oracle", "pfc_oracle.c"), "-lm"],
                   check=True)
    Nx = Ny = 64
    dx = LATTICE_A / 8
    dy = dx * 1.01
    f = crystal(Nx, Ny, dx, dy, -0.2, 0.2, LATTICE_A, sym=3, theta=0.2, noise=0.01, seed=4)
    out = {}
    for c, topt in ((0, 0), (1, 50)):
        runs = []
        for lib in (None, so):
            o = Oracle(graphene_params(c=float(c)), Nx, Ny, dx, dy, so=lib)
            o.set_fields(f[None])
            runs.append(o.relax(250, 100, Topt=topt, xopt=1, yopt=1, snap_steps=(250,)))
            o.close()
        (l1, s1), (l2, s2) = runs
        out[topt] = (float(np.max(np.abs(l1[:, 1:3] - l2[:, 1:3]) / np.abs(l1[:, 1:3]))),
                     float(np.max(np.abs(l1[:, 3] - l2[:, 3]) / np.abs(l1[:, 3]))),
                     float(np.linalg.norm(s1 - s2) / np.linalg.norm(s1)))
    box0, fave0, dens0 = out[0]
    box1, fave1, dens1 = out[50]
    assert box0 == 0.0 and fave0 < 1e-15 and dens0 < 1e-13
    # the searches lift the same round-off difference past 1e-12
    assert box1 > 1e-12 and fave1 > 5e-13
    assert box1 < 1e-10 and fave1 < 1e-10 and dens1 < 1e-9
