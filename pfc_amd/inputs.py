"""Seeded synthetic initial density fields in the shapes pfc-init builds.

The one-mode approximations follow OMA of the reference's
src/PFCInit.java:46-66 (sym 3 triangular, 4 square, 6 hexagonal); periodic
nearest-seed (Voronoi) grains follow PFCInit.java:389-462 and the
mean/amplitude transform T follows :649-667.  Randomness comes from numpy's
PCG64 with an explicit seed (not java.util.Random), so files are
reproducible but not byte-identical to pfc-init's.
"""
from __future__ import annotations

import math

import numpy as np


def one_mode(x, y, sym, a):
    """One-mode density pattern of lattice constant a (PFCInit.java:46-66)."""
    q = 2.0 * math.pi / a
    if sym == 3:
        qx, qy = q, q / math.sqrt(3.0)
        return np.cos(qx * x) * np.cos(qy * y) + 0.5 * np.cos(2.0 * qy * y)
    if sym == 4:
        return np.cos(q * x) * np.cos(q * y)
    if sym == 6:
        qx, qy = q, q / math.sqrt(3.0)
        return (-2.0 * np.cos(qx * x - qy * y) - 2.0 * np.cos(2.0 * qy * y)
                - 2.0 * np.cos(-qx * x - qy * y) + np.cos(-qx * x - 3.0 * qy * y)
                + np.cos(2.0 * qx * x) + np.cos(-qx * x + 3.0 * qy * y))
    raise ValueError("sym must be 3, 4 or 6")


def grid(Nx, Ny, dx, dy):
    y, x = np.meshgrid(np.arange(Ny) * dy, np.arange(Nx) * dx, indexing="ij")
    return x, y


def rotated(x, y, theta, tx=0.0, ty=0.0):
    """Coordinates rotated by -theta about the origin, then shifted
    (PFCInit.java:448-453 with xr = yr = tx2 = ty2 = 0)."""
    s, c = math.sin(-theta), math.cos(-theta)
    return x * c - y * s - tx, x * s + y * c - ty


def crystal(Nx, Ny, dx, dy, nbar, amp, a, sym=3, theta=0.0, noise=0.0, seed=0):
    """A single crystal n = nbar + amp*OMA + uniform noise in [-noise, noise)."""
    x, y = grid(Nx, Ny, dx, dy)
    rx, ry = rotated(x, y, theta)
    rng = np.random.default_rng(seed)
    return nbar + amp * one_mode(rx, ry, sym, a) + 2.0 * noise * (rng.random((Ny, Nx)) - 0.5)


def voronoi_labels(Nx, Ny, dx, dy, seeds_xy):
    """Periodic nearest-seed labels (PFCInit.java:68-93)."""
    W, H = Nx * dx, Ny * dy
    try:
        from scipy.spatial import cKDTree
    except ImportError:   # small grids only
        cKDTree = None
    if cKDTree is not None:
        pts = np.mod(np.asarray(seeds_xy, dtype=np.float64), [W, H])
        tree = cKDTree(pts, boxsize=[W, H])
        lab = np.empty((Ny, Nx), dtype=np.int64)
        xs = np.arange(Nx) * dx
        rows = max(1, (1 << 20) // Nx)
        for j0 in range(0, Ny, rows):
            j1 = min(Ny, j0 + rows)
            yy, xx = np.meshgrid(np.arange(j0, j1) * dy, xs, indexing="ij")
            q = np.column_stack([np.mod(xx.ravel(), W), np.mod(yy.ravel(), H)])
            lab[j0:j1] = tree.query(q)[1].reshape(j1 - j0, Nx)
        return lab
    x, y = grid(Nx, Ny, dx, dy)
    best = np.full((Ny, Nx), np.inf)
    lab = np.zeros((Ny, Nx), dtype=np.int64)
    for n, (sx, sy) in enumerate(seeds_xy):
        ddx = x - sx
        ddx = np.where(ddx > 0.5 * W, ddx - W, np.where(ddx < -0.5 * W, ddx + W, ddx))
        ddy = y - sy
        ddy = np.where(ddy > 0.5 * H, ddy - H, np.where(ddy < -0.5 * H, ddy + H, ddy))
        d2 = ddx * ddx + ddy * ddy
        m = d2 < best
        best[m] = d2[m]
        lab[m] = n
    return lab


def _rotate_labeled(x, y, theta_per_point):
    s, c = np.sin(-theta_per_point), np.cos(-theta_per_point)
    return x * c - y * s, x * s + y * c


def polycrystal(Nx, Ny, dx, dy, n_grains, nbar, amp, a, sym=3, noise=0.0, boundary=0,
                seed=0, max_angle=60.0):
    """Seeded Voronoi grains with uniform random orientations in
    [0, max_angle) degrees; sites within `boundary` grid points of a grain
    boundary are set to nbar (a disordered boundary at mean density)."""
    rng = np.random.default_rng(seed)
    seeds = np.column_stack([rng.random(n_grains) * Nx * dx, rng.random(n_grains) * Ny * dy])
    angles = rng.random(n_grains) * math.radians(max_angle)
    lab = voronoi_labels(Nx, Ny, dx, dy, seeds)
    x, y = grid(Nx, Ny, dx, dy)
    rx, ry = _rotate_labeled(x, y, angles[lab])
    out = nbar + amp * one_mode(rx, ry, sym, a)
    if boundary > 0:
        edge = np.zeros(lab.shape, dtype=bool)
        for s in range(1, boundary + 1):
            for ax in (0, 1):
                edge |= lab != np.roll(lab, s, axis=ax)
                edge |= lab != np.roll(lab, -s, axis=ax)
        out[edge] = nbar
    out += 2.0 * noise * (rng.random((Ny, Nx)) - 0.5)
    return out


def hetero_patchwork(Nx, Ny, dx, dy, domain=256, seed=0, noise=0.01, n_domains=None):
    """Three fields (C, B, N) of a graphene/hBN patchwork.

    Domains come from a periodic Voronoi tessellation with about one seed per
    domain^2 grid points (or n_domains seeds); half of them, chosen at random,
    are graphene (honeycomb in C, B and N at their mean) and the rest hBN
    (B and N triangular sublattices offset along y, C at its mean).  Lattice
    and amplitude choices follow the reference's demos/hetero/hetero.i0-2.
    """
    rng = np.random.default_rng(seed)
    if n_domains is None:
        n_domains = max(2, int(round(Nx * Ny / float(domain * domain))))
    seeds = np.column_stack([rng.random(n_domains) * Nx * dx, rng.random(n_domains) * Ny * dy])
    angles = rng.random(n_domains) * math.radians(60.0)
    is_g = np.zeros(n_domains, dtype=bool)
    is_g[rng.permutation(n_domains)[: n_domains // 2]] = True
    lab = voronoi_labels(Nx, Ny, dx, dy, seeds)
    x, y = grid(Nx, Ny, dx, dy)
    rx, ry = _rotate_labeled(x, y, angles[lab])
    del x, y
    g = is_g[lab]
    a_g, a_h = 7.2552, 7.3858
    f = np.empty((3, Ny, Nx))
    f[0] = np.where(g, 0.485 - 0.3 * one_mode(rx, ry, 3, a_g), 0.485)
    f[1] = np.where(g, -0.32, -0.32 + one_mode(rx, ry - 2.1, 3, a_h))
    f[2] = np.where(g, -0.32, -0.32 + one_mode(rx, ry + 2.1, 3, a_h))
    f += 2.0 * noise * (rng.random(f.shape) - 0.5)
    return f


def write_density(path, field, dx, dy, prec=6):
    """Writes a density file in the format pfc-init and pfc-relax share:
    header `Nx Ny dx dy`, then one value per line, row by row
    (PFCInit.java:670-686)."""
    Ny, Nx = field.shape
    fmt = "%." + str(prec) + "e"
    with open(path, "w") as fh:
        fh.write(f"{Nx} {Ny} " + (fmt % dx) + " " + (fmt % dy) + "\n")
        np.savetxt(fh, field.reshape(-1), fmt=fmt)
