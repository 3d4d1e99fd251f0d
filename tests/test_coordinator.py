"""The coordinator restatement (src/Coordinator.java:29-397) on the CPU:
exact honeycomb geometry of an ideal inverted-triangular density, an
independent brute-force restatement of the triplet rule on a distorted field,
the argv contract and file formats of the command, and Java's number format."""
import itertools
import math
import os
import subprocess

import numpy as np
import pytest

from pfc_amd import coordinator as co


def triangular_minima_field(Nx, Ny, dx, dy, a, theta=0.0, shift=(0.0, 0.0)):
    """-sum of three plane waves: minima on a triangular lattice of spacing a
    (the ring centres), maxima on the honeycomb between them."""
    x = (np.arange(Nx) * dx)[None, :] - shift[0]
    y = (np.arange(Ny) * dy)[:, None] - shift[1]
    q = 4 * math.pi / (math.sqrt(3) * a)
    f = np.zeros((Ny, Nx))
    for k in range(3):
        ang = theta + math.pi / 2 + 2 * math.pi * k / 3
        f += np.cos(q * (math.cos(ang) * x + math.sin(ang) * y))
    return -f


def test_java_double_format():
    cases = {58.0: "58.0", 0.1: "0.1", 1.0 / 3.0: "0.3333333333333333", 2.0 / 3.0: "0.6666666666666666",
             0.001: "0.001", 1.0e-4: "1.0E-4", 1.0e7: "1.0E7", 1.2345678e7: "1.2345678E7",
             123456.789: "123456.789", 9999999.0: "9999999.0", -2.5: "-2.5", 0.0: "0.0",
             1e21: "1.0E21", 1.5e-10: "1.5E-10", 100.0: "100.0"}
    for d, s in cases.items():
        assert co.java_double(d) == s, (d, co.java_double(d), s)


def test_ideal_honeycomb():
    """Ring spacing a = 8 dx along x; an Nx x Ny box holding whole cells.  Every
    atom has three bonds at a/sqrt(3), every ring six atoms, and there are two
    atoms per ring."""
    a = 7.2552
    nx_cells, ny_rows = 8, 8
    dx = a / 16
    dy = a * math.sqrt(3) / 2 / 14            # 14 grid rows per ring row
    Nx, Ny = 16 * nx_cells, 14 * ny_rows
    # rows of ring centres along x (theta = 0 puts one reciprocal vector along y),
    # shifted off the grid origin: a ring at x = -0 would wrap to x = w exactly
    # and fall outside the bins, where the reference throws
    psi = triangular_minima_field(Nx, Ny, dx, dy, a, theta=0.0, shift=(0.3 * dx, 0.4 * dy))
    xy, (w, h), nrings, flags = co.coordinate(psi, dx, dy, a, 1.0)
    assert nrings == nx_cells * ny_rows
    assert len(xy) == 2 * nrings
    assert flags == 0
    assert w == pytest.approx(Nx * dx / a) and h == pytest.approx(Ny * dy / a)
    # bond length a/sqrt(3) in lattice constants (sub-grid parabola error ~1e-2)
    from scipy.spatial import cKDTree
    d, _ = cKDTree(np.mod(xy, [w, h]), boxsize=[w, h]).query(np.mod(xy, [w, h]), k=4)
    assert np.allclose(d[:, 1:], 1 / math.sqrt(3), atol=0.02)


def brute_force_atoms(rings, w, h, la):
    """All triplets of rings pairwise within 2 la with no fourth ring strictly
    inside their circumcircle; the atom is the triplet's centroid (periodic)."""
    def dif(b, a):
        d = b - a
        d[0] -= w * np.round(d[0] / w)
        d[1] -= h * np.round(d[1] / h)
        return d
    out = []
    n = len(rings)
    for i, j, k in itertools.combinations(range(n), 3):
        ij, ik, jk = dif(rings[j], rings[i]), dif(rings[k], rings[i]), dif(rings[k], rings[j])
        if max(np.hypot(*ij), np.hypot(*ik), np.hypot(*jk)) > 2 * la:
            continue
        D = 2 * (ij[0] * ik[1] - ij[1] * ik[0])
        if D == 0:
            continue
        ux = (ik[1] * (ij @ ij) - ij[1] * (ik @ ik)) / D
        uy = (ij[0] * (ik @ ik) - ik[0] * (ij @ ij)) / D
        it = np.array([ux, uy])
        R2 = it @ it
        ok = True
        for l in range(n):
            if l in (i, j, k):
                continue
            dl = dif(rings[l], rings[i]) - it
            if dl @ dl < R2 * (1 - 1e-12):
                ok = False
                break
        if ok:
            out.append(np.mod(rings[i] + (ij + ik) / 3, [w, h]))
    return np.array(out)


def test_distorted_field_matches_brute_force():
    """A sheared, rotated and noisy field: the atoms equal those of the
    brute-force triplet rule applied to the same rings."""
    rng = np.random.default_rng(3)
    a, dx, dy = 7.2552, 0.9, 0.85
    Nx = Ny = 64
    psi = triangular_minima_field(Nx, Ny, dx, dy, a, theta=0.37, shift=(1.3, 0.4))
    psi += 0.05 * rng.standard_normal(psi.shape)
    xy, (w, h), nrings, flags = co.coordinate(psi, dx, dy, a, 1.0)
    # rings: local minima of the same field in the same units
    ux, uy = dx / a, dy / a
    p = psi.copy()
    rings = []
    for j in range(Ny):
        for i in range(Nx):
            c = p[j, i]
            nb = [p[(j + v) % Ny, (i + u) % Nx] for u, v in
                  ((1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1))]
            if min(nb) < c:
                continue
            def vert(k, p0, p1, p2, u):
                x0, x1, x2 = k - 1.0, float(k), k + 1.0
                A = x2 * (p1 - p0) + x1 * (p0 - p2) + x0 * (p2 - p1)
                B = x2 * x2 * (p0 - p1) + x1 * x1 * (p2 - p0) + x0 * x0 * (p1 - p2)
                return k if abs(A) < 1e-14 else -B / (2 * A) * u
            rx = vert(i, p[j, (i - 1) % Nx], c, p[j, (i + 1) % Nx], ux)
            ry = vert(j, p[(j - 1) % Ny, i], c, p[(j + 1) % Ny, i], uy)
            rings.append(np.mod([rx, ry], [w, h]))
            p[j, i] -= 2000
            for u, v in ((1, 0), (1, 1), (0, 1), (-1, 1), (-1, 0), (-1, -1), (0, -1), (1, -1)):
                p[(j + v) % Ny, (i + u) % Nx] -= 1000
    assert len(rings) == nrings
    ref = brute_force_atoms(np.array(rings), w, h, 1.0)
    assert len(ref) == len(xy)
    cmp = co.compare_atoms(ref, xy, (w, h))
    assert cmp["max"] < 1e-9


def run(args, cwd):
    r = subprocess.run([co.BIN_PATH, *args], cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=60)
    return r.returncode, r.stdout


def test_command_contract(tmp_path):
    rc, out = run([], tmp_path)
    assert rc == 0 and out.startswith("This is a tool for converting PFC density fields")
    rc, out = run(["a", "b"], tmp_path)
    assert (rc, out) == (1, "Error: Invalid syntax.\n")
    rc, out = run(["none.n", "o.xyz", "7.26", "2.46"], tmp_path)
    assert (rc, out) == (1, "Error: Input file not found.\n")


def test_command_files(tmp_path):
    a = 7.2552
    dx = a / 16
    dy = a * math.sqrt(3) / 2 / 14
    Nx, Ny = 128, 112
    psi = triangular_minima_field(Nx, Ny, dx, dy, a, shift=(0.3 * dx, 0.4 * dy))
    with open(tmp_path / "d.n", "w") as f:
        f.write(f"{Nx} {Ny} {dx:.6f} {dy:.6f}\n")
        for v in psi.ravel():
            f.write(f"{v:.6e}\n")
    rc, out = run(["d.n", "c.xyz", "nh.nh", str(a), "2.46"], tmp_path)
    assert rc == 0 and out == ""
    lines = (tmp_path / "c.xyz").read_text().splitlines()
    w, h = (float(v) for v in lines[0].split())
    assert w == pytest.approx(Nx * round(dx, 6) * 2.46 / a)
    atoms = [l.split() for l in lines[1:]]
    assert len(atoms) == 2 * 64
    for t in atoms:
        assert t[0] == "A" and t[3] == "0.0" and len(t) == 7   # three neighbour indices
        assert all(0 <= int(k) < len(atoms) for k in t[4:])
    assert (tmp_path / "nh.nh").read_text() == ""              # all rings hexagons
