"""pfc-init, the initial-state generator (src/PFCInit.java:96-698), pinned by
what the definitions fix rather than by a second copy of its formulas:
java.util.Random's published outputs, the lattice ansatz at points where it
has exact values, rotations and translations that turn one lattice into a
known closed form, Voronoi / disc / polygon regions recomputed by brute-force
geometry, the transform's mean and the loader's periodic shift."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INIT = os.path.join(ROOT, "pfc_amd", "bin", "pfc-init")


def run(text, cwd, name="x.i"):
    (cwd / name).write_text(text)
    r = subprocess.run([INIT, name], cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                       text=True, timeout=60)
    return r.returncode, r.stdout


def read(path):
    with open(path) as f:
        Nx, Ny, dx, dy = f.readline().split()
        v = np.array([float(l) for l in f])
    Nx, Ny = int(Nx), int(Ny)
    assert v.size == Nx * Ny
    return v.reshape(Ny, Nx), float(dx), float(dy)


def grain(sym=4, a=2.0, nave=0.0, amp=1.0, namp=0.0, tx1=0.0, ty1=0.0, theta=0.0,
          xr=0.0, yr=0.0, tx2=0.0, ty2=0.0):
    return f"{sym} {a!r} {nave!r} {amp!r} {namp!r}  {tx1!r} {ty1!r} {theta!r} {xr!r} {yr!r} {tx2!r} {ty2!r}"


def one_grain(tmp_path, Nx, Ny, dx, dy, prec=17, **kw):
    rc, out = run(f"D {Nx} {Ny} {dx!r} {dy!r}\nV 1\n 0.0 0.0 0 {grain(**kw)}\nO o.n {prec}\n", tmp_path)
    assert (rc, out) == (0, "")
    return read(tmp_path / "o.n")[0]


def grid(Nx, Ny, dx, dy):
    y, x = np.meshgrid(np.arange(Ny) * dy, np.arange(Nx) * dx, indexing="ij")
    return x, y


# -- java.util.Random -------------------------------------------------------

# The first nextDouble() values of new Random(0) and new Random(42), as the
# Java platform documents and every JVM prints them.
JAVA_RANDOM_0 = [0.730967787376657, 0.24053641567148587, 0.6374174253501083,
                 0.5504370051176339, 0.5975452777972018]
JAVA_RANDOM_42 = [0.7275636800328681, 0.6832234717598454, 0.30871945533265976]


@pytest.mark.parametrize("seed,ref", [(0, JAVA_RANDOM_0), (42, JAVA_RANDOM_42)])
def test_noise_is_java_random(tmp_path, seed, ref):
    """With nave = amp = 0 and namp = 0.5 a grain's value is nextDouble() - 0.5
    exactly, drawn in raster order (rows of x within y)."""
    text = f"D 5 2 1.0 1.0\nS {seed}\nV 1\n 0 0 0 {grain(amp=0.0, namp=0.5)}\nO o.n 17\n"
    rc, out = run(text, tmp_path)
    assert (rc, out) == (0, "")
    f = read(tmp_path / "o.n")[0].ravel()
    assert list(f[:len(ref)] + 0.5) == ref


def test_unseeded_runs_differ(tmp_path):
    text = f"D 8 8 1.0 1.0\nV 1\n 0 0 0 {grain(amp=0.0, namp=0.5)}\nO o.n 17\n"
    assert run(text, tmp_path)[0] == 0
    a = read(tmp_path / "o.n")[0]
    assert np.all((a >= -0.5) & (a < 0.5))


# -- the lattice ansatz -------------------------------------------------------

@pytest.mark.parametrize("sym,val", [(3, 1.5), (4, 1.0), (6, -3.0)])
def test_ansatz_at_origin(tmp_path, sym, val):
    f = one_grain(tmp_path, 4, 4, 0.5, 0.5, sym=sym, a=7.2552)
    assert f[0, 0] == val


def test_square_ansatz_on_its_grid(tmp_path):
    """sym 4 with a = 4 dx: cos(pi i / 2) cos(pi j / 2)."""
    f = one_grain(tmp_path, 8, 8, 0.5, 0.5, sym=4, a=2.0)
    i = np.arange(8)
    ref = np.outer(np.cos(np.pi * i / 2), np.cos(np.pi * i / 2))
    assert np.max(np.abs(f - ref)) < 1e-15


def test_triangular_ansatz_closed_form(tmp_path):
    """sym 3: cos(qx)cos(qy/sqrt3) + cos(2qy/sqrt3)/2, a triangular lattice of
    maxima: its value is 1.5 on the lattice sites a(i + j/2, j sqrt3/2)."""
    a = 7.2552
    Nx, Ny = 32, 28
    dx, dy = a / 8, a * math.sqrt(3) / 2 / 7
    f = one_grain(tmp_path, Nx, Ny, dx, dy, sym=3, a=a)
    for j in range(0, Ny, 7):
        for i in range(0, Nx, 8):
            ii = i + (4 if (j // 7) % 2 else 0)
            assert f[j, ii % Nx] == pytest.approx(1.5, abs=1e-12)
    assert f.max() <= 1.5 + 1e-12 and f.min() >= -0.75 - 1e-12


def test_honeycomb_ansatz_closed_form(tmp_path):
    """sym 6 is -2(cos A + cos B + cos C) + cos(A-B) + cos(B-C) + cos(C-A) with
    A + B + C = 0; it peaks at 6 where A = B = C = 2 pi / 3, i.e. at
    (a/2, a/(2 sqrt3)), has zero mean over a period and is invariant under a
    turn by 120 degrees about the origin (not by 30)."""
    a = 7.2552
    Nx, Ny = 48, 36
    dx, dy = a / 8, a * math.sqrt(3) / 18
    f = one_grain(tmp_path, Nx, Ny, dx, dy, sym=6, a=a)
    assert f[3, 4] == pytest.approx(6.0, abs=1e-12)
    assert f.max() <= 6.0 + 1e-12
    assert abs(f.mean()) < 1e-12
    f120 = one_grain(tmp_path, Nx, Ny, dx, dy, sym=6, a=a, theta=120.0)
    assert np.max(np.abs(f120 - f)) < 1e-12
    f30 = one_grain(tmp_path, Nx, Ny, dx, dy, sym=6, a=a, theta=30.0)
    assert np.max(np.abs(f30 - f)) > 1.0


@pytest.mark.parametrize("case", ["rot", "tx2", "xr"])
def test_rotation_and_translations(tmp_path, case):
    """Square lattice, q = 2 pi / a, turned by 90 degrees:
      tx1 = a/4 (before the turn)   -> sin(q y) cos(q x)
      tx2 = a/4 (after the turn)    -> cos(q y) sin(q x)
      axis at (a/4, 0)              -> -sin(q y) sin(q x)
    Each sign of the rotation or of a translation would flip one factor."""
    a, Nx, Ny, dx, dy = 2.0, 16, 16, 0.25, 0.25
    q = 2 * math.pi / a
    x, y = grid(Nx, Ny, dx, dy)
    kw = {"rot": dict(tx1=a / 4), "tx2": dict(tx2=a / 4), "xr": dict(xr=a / 4)}[case]
    ref = {"rot": np.sin(q * y) * np.cos(q * x), "tx2": np.cos(q * y) * np.sin(q * x),
           "xr": -np.sin(q * y) * np.sin(q * x)}[case]
    f = one_grain(tmp_path, Nx, Ny, dx, dy, sym=4, a=a, theta=90.0, **kw)
    assert np.max(np.abs(f - ref)) < 1e-14


def test_sixty_degrees_maps_triangular_onto_itself(tmp_path):
    a, Nx, Ny = 7.2552, 32, 28
    dx, dy = a / 8, a * math.sqrt(3) / 2 / 7
    f0 = one_grain(tmp_path, Nx, Ny, dx, dy, sym=3, a=a)
    f60 = one_grain(tmp_path, Nx, Ny, dx, dy, sym=3, a=a, theta=60.0)
    assert np.max(np.abs(f60 - f0)) < 1e-12


# -- regions ------------------------------------------------------------------

def test_voronoi_cells_brute_force(tmp_path):
    """amp = 0: every point takes the nave of its nearest seed under periodic
    minimum image (first seed on ties); seeds outside the box are wrapped."""
    Nx, Ny, dx, dy = 40, 30, 0.7, 0.6
    W, H = Nx * dx, Ny * dy
    seeds = [(3.0, 4.0, 0.1), (20.0, 15.0, 0.2), (-2.0, 9.0, 0.3), (27.0 + W, 1.0, 0.4)]
    text = f"D {Nx} {Ny} {dx} {dy}\nV {len(seeds)}\n"
    for sx, sy, nv in seeds:
        text += f" {sx} {sy} 0 {grain(amp=0.0, nave=nv)}\n"
    rc, out = run(text + "O o.n 17\n", tmp_path)
    assert (rc, out) == (0, "")
    f = read(tmp_path / "o.n")[0]
    x, y = grid(Nx, Ny, dx, dy)
    best = None
    for sx, sy, nv in seeds:
        sx, sy = sx % W, sy % H
        ddx = np.abs(x - sx)
        ddy = np.abs(y - sy)
        d2 = np.minimum(ddx, W - ddx) ** 2 + np.minimum(ddy, H - ddy) ** 2
        if best is None:
            best, ref = d2, np.full_like(d2, nv)
        else:
            ref = np.where(d2 < best - 1e-9, nv, ref)
            best = np.minimum(best, d2)
    assert np.array_equal(f, ref)


def test_voronoi_types(tmp_path):
    """type 1 adds to the field, type 2 leaves it unchanged, type 0 overwrites."""
    Nx, Ny = 16, 8
    base = f"D {Nx} {Ny} 1.0 1.0\nV 1\n 0 0 0 {grain(amp=0.0, nave=1.0)}\n"
    text = base + f"V 3\n 2 4 0 {grain(amp=0.0, nave=5.0)}\n 7 4 1 {grain(amp=0.0, nave=0.25)}\n" \
                  f" 12 4 2 {grain(amp=0.0, nave=9.0)}\nO o.n 17\n"
    rc, out = run(text, tmp_path)
    assert (rc, out) == (0, "")
    f = read(tmp_path / "o.n")[0]
    assert f[4, 2] == 5.0 and f[4, 7] == 1.25 and f[4, 12] == 1.0


def test_circle_wraps_periodically(tmp_path):
    Nx, Ny, dx, dy, R = 32, 24, 0.5, 0.5, 3.1
    xc, yc = 1.0, 11.0                      # crosses x = 0
    text = f"D {Nx} {Ny} {dx} {dy}\nC {xc} {yc} {R} 0 {grain(amp=0.0, nave=1.0)}\nO o.n 17\n"
    rc, out = run(text, tmp_path)
    assert (rc, out) == (0, "")
    f = read(tmp_path / "o.n")[0]
    ref = np.zeros((Ny, Nx))
    for j in range(int((yc - R) / dy), int((yc + R) / dy) + 1):
        for i in range(int((xc - R) / dx), int((xc + R) / dx) + 1):
            if (i * dx - xc) ** 2 + (j * dy - yc) ** 2 < R * R:
                ref[j % Ny, i % Nx] = 1.0
    assert np.array_equal(f, ref)
    assert ref[:, Nx - 4:].sum() > 0       # the part left of x = 0 landed at the far edge


def test_polygon_region(tmp_path):
    """A convex quadrilateral, either winding: the points on the inner side of
    every edge (boundary included) are set, by a half-plane test per edge."""
    Nx, Ny, dx, dy = 40, 40, 0.5, 0.5
    corners = [(2.0, 3.0), (15.0, 5.0), (17.0, 16.0), (4.0, 13.0)]
    for pts in (corners, corners[::-1]):
        text = f"D {Nx} {Ny} {dx} {dy}\nP 4\n" + "".join(f" {x} {y}\n" for x, y in pts)
        text += f"0 {grain(amp=0.0, nave=1.0)}\nO o.n 17\n"
        rc, out = run(text, tmp_path)
        assert (rc, out) == (0, "")
        f = read(tmp_path / "o.n")[0]
        x, y = grid(Nx, Ny, dx, dy)
        s = np.sign(sum(pts[n][0] * pts[(n + 1) % 4][1] - pts[(n + 1) % 4][0] * pts[n][1]
                        for n in range(4)))   # orientation from the shoelace area
        inside = np.ones_like(x, dtype=bool)
        for n in range(4):
            (x0, y0), (x1, y1) = pts[n], pts[(n + 1) % 4]
            inside &= s * ((x1 - x0) * (y - y0) - (y1 - y0) * (x - x0)) >= 0
        assert np.array_equal(f == 1.0, inside)
        assert f.sum() == inside.sum()


def test_concave_polygon_is_refused(tmp_path):
    text = "D 8 8 1 1\nP 4\n 0 0  4 1  1 1  1 4\n0 " + grain() + "\nO o.n 6\n"
    assert run(text, tmp_path) == (1, "Error: Concave grain detected. Convex grain expected.\n")


# -- transform, load, output --------------------------------------------------

def test_transform_sets_mean_and_scales(tmp_path):
    Nx, Ny = 16, 12
    text = f"D {Nx} {Ny} 0.5 0.5\nS 7\nV 1\n 0 0 0 {grain(sym=3, a=3.0, nave=0.2, namp=0.1)}\n" \
           "O a.n 17\nT -0.3 0.5\nO b.n 17\n"
    rc, out = run(text, tmp_path)
    assert (rc, out) == (0, "")
    a = read(tmp_path / "a.n")[0]
    b = read(tmp_path / "b.n")[0]
    assert b.mean() == pytest.approx(-0.3, abs=1e-15)
    assert np.max(np.abs((b + 0.3) - 0.5 * (a - a.mean()))) < 1e-15


def test_load_places_a_wrapped_block(tmp_path):
    """L copies a dNx x dNy block from (i0, j0) of a saved field to (i1, j1),
    both wrapped; type 1 adds it to what is there."""
    rng = np.random.default_rng(3)
    src = rng.standard_normal((10, 12))
    with open(tmp_path / "s.n", "w") as fh:
        fh.write("12 10 5.00000e-01 5.00000e-01\n")
        np.savetxt(fh, src.ravel(), fmt="%.17e")
    text = f"D 9 7 0.5 0.5\nV 1\n 0 0 0 {grain(amp=0.0, nave=1.0)}\n" \
           "L s.n 1  -3 8 5 4  7 -1\nO o.n 17\n"
    rc, out = run(text, tmp_path)
    assert (rc, out) == (0, "")
    f = read(tmp_path / "o.n")[0]
    ref = np.ones((7, 9))
    for j in range(4):
        for i in range(5):
            ref[(-1 + j) % 7, (7 + i) % 9] = 1.0 + src[(8 + j) % 10, (-3 + i) % 12]
    assert np.array_equal(f, ref)


def test_load_warnings(tmp_path):
    with open(tmp_path / "s.n", "w") as fh:
        fh.write("2 2 7.00000e-01 7.00000e-01\n1\n2\n3\n4\n")
    rc, out = run("D 2 2 0.5 0.5\nL s.n 0 0 0 2 2 0 0\nO o.n 3\n", tmp_path)
    assert rc == 0
    assert out == 'Warning: System from data file "s.n" has incompatible discretization.\n'
    assert read(tmp_path / "o.n")[0].ravel().tolist() == [1, 2, 3, 4]
    rc, out = run("D 2 2 0.5 0.5\nL t.n 0 0 0 2 2 0 0\nO o.n 3\n", tmp_path)
    assert (rc, out) == (1, 'Error: Data file "t.n" not found.\n')


def test_output_format(tmp_path):
    rc, out = run("D 3 2 0.7 0.65\nT 0.25 1.0\nO o.n 4\n", tmp_path)
    assert (rc, out) == (0, "")
    assert (tmp_path / "o.n").read_text() == "3 2 7.0000e-01 6.5000e-01\n" + "2.5000e-01\n" * 6


def test_comments_and_layout(tmp_path):
    """Words may be split over lines and tabs at will; '#' starts a comment
    only as the first character of a word."""
    text = "# header\nD\n\n#\tNx\tNy\n\t4\t\t2 # a comment 99\n 0.5\n0.5\nT 1.5 1 O o.n\n2\n"
    rc, out = run(text, tmp_path)
    assert (rc, out) == (0, "")
    assert read(tmp_path / "o.n")[0].ravel().tolist() == [1.5] * 8


def test_messages(tmp_path):
    r = subprocess.run([INIT], cwd=tmp_path, stdout=subprocess.PIPE, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.startswith("This is a tool for constructing 2D density fields")
    r = subprocess.run([INIT, "none.i"], cwd=tmp_path, stdout=subprocess.PIPE, text=True, timeout=60)
    assert (r.returncode, r.stdout) == (1, 'Error: Input file "none.i" not found.\n')
    assert run("D 4 4 0.5 0.5\nX\n", tmp_path) == (1, 'Error: Invalid input file. Label expected instead of "X".\n')
    assert run("D 4.0 4 0.5 0.5\n", tmp_path) == (1, 'Error: Invalid input file. Integer expected instead of "4.0".\n')
    assert run("D 4 4 0.5 half\n", tmp_path) == (1, 'Error: Invalid input file. Decimal number expected instead of "half".\n')
    assert run("D 4 4 0.5 0.5d\nO o.n 2\n", tmp_path) == (0, "")   # Java's suffix form
    assert run(f"D 4 4 0.5 0.5\nV 1\n 0 0 0 {grain(sym=5)}\n", tmp_path) == \
        (1, "Error: Invalid lattice symmetry specified. Only triangular, square and hexagonal "
            "symmetries are supported.\n")
    assert run(f"D 4 4 0.5 0.5\nC 1 1 1 2 {grain()}\nO o.n 2\n", tmp_path) == \
        (1, 'Warning: Phantom type used for a circular grain.\n'
            'Error: Invalid input file. Label expected instead of "4".\n')


HETERO_GRAINS = [  # seed x, y, symmetry, nave, amp, theta of the heterostructure demo's graphene layer
    (173.174, 252.155, 6, 0.66, 0.0, 150.487), (262.565, 189.181, 6, 0.66, 0.0, 280.869),
    (189.433, 195.483, 6, 0.66, 0.0, 109.178), (183.422, 109.088, 6, 0.66, 0.0, 214.594),
    (157.924, 104.086, 6, 0.66, 0.0, 4.48658), (283.674, 39.718, 3, 0.31, -1.0, 94.1102),
    (7.85441, 296.342, 3, 0.31, -1.0, 169.041), (237.943, 156.396, 3, 0.31, -1.0, 37.6864),
    (316.331, 247.728, 3, 0.31, -1.0, 216.745), (267.606, 2.47376, 3, 0.31, -1.0, 216.745)]


def test_heterostructure_demo_layer(tmp_path):
    """The demo's first layer (512^2, ten Voronoi grains, then T 0.485 1.0):
    flat 0.66 - mean + 0.485 in the five sym-6 cells with amp 0, a triangular
    pattern elsewhere, mean 0.485 to the printed precision."""
    text = "D 512 512 0.7 0.7\nV 10\n"
    for x, y, sym, nave, amp, th in HETERO_GRAINS:
        text += f"{x} {y} 0\n{sym} 7.2552 {nave} {amp} 0.0\n0.0 0.0 {th} 0.0 0.0 0.0 0.0\n"
    rc, out = run(text + "T 0.485 1.0\nO hetero.n0 6\n", tmp_path)
    assert (rc, out) == (0, "")
    f, dx, dy = read(tmp_path / "hetero.n0")
    assert f.shape == (512, 512) and (dx, dy) == (0.7, 0.7)
    assert f.mean() == pytest.approx(0.485, abs=2e-6)
    flat = f[int(252.155 / 0.7), int(173.174 / 0.7)]
    cells = [f[int(y / 0.7), int(x / 0.7)] for x, y, *_ in HETERO_GRAINS[:5]]
    assert all(c == flat for c in cells)
    assert np.count_nonzero(f == flat) > 512 * 512 // 4
