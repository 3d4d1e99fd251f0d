"""The spectral form of the free energy sums that pfc_evaluate_sums computes
(k_energy_local and k_energy_spec in pfc_amd/csrc/pfc_kernels.hpp), restated
on the half spectrum with the kernels' column weights, against the oracle's
real-space evaluate (src/pfc-relax.c:987-1130) on the same state.

The identity behind it is Parseval's: the grid mean of n_u times a filtered
field C2R(m N_v) is the half-spectrum sum of w m Re(N_u conj N_v) for
N = FFT2(n) / (Nx Ny), w = 1 on the kx = 0 and Nyquist columns and 2 between
them.  White-noise states excite every mode, including the ky = +-(Ny -+ 1)/2
pair of an odd Ny where the multiplier is uneven in ky.  A weight mistake
(the kx = 0 or Nyquist column counted twice, or an interior column once)
must break the agreement.
"""
import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import LATTICE_A, ENERGY_NU2, graphene_params, hetero_params, make_params
from pfc_amd.inputs import crystal, hetero_patchwork


def two_mode_params():
    """Two fields, two modes on field 0 with phi != 0, every pair term on."""
    return make_params(
        U=2, lam=[0.01, -0.02], alpha=[[-0.3, 0.05], [0, -0.2]], beta=[[1.0, 0.1], [0, 1.1]],
        gamma=[[0.8, 0.2], [0, 0.6]], delta=[1.0, 0.9], epsilon=[[0, -0.3], [0, 0]], sigma=0.25,
        nu=[[1.0, 0.58], [1.05]], omega=[[1.0, 0.7], [1.0]], phi=[[0.0, 0.02], [0.0]],
        mu=[[0, 1.02], [0, 0]], c=1.0, dt=0.5)


def spectral_energy(p, n, dx, dy, weights="kernel"):
    """fave from the densities n (U x Ny x Nx) the way the two kernels split
    it: local terms summed on the grid, the rest over the half spectrum."""
    U, Ny, Nx = n.shape
    f = np.zeros((Ny, Nx))
    for u in range(U):
        a = n[u]
        f += (p.lambda_[u] * a + 0.5 * p.alpha[u][u] * a * a + p.gamma[u][u] / 3.0 * a ** 3
              + 0.25 * p.delta[u] * a ** 4)
        for v in range(u + 1, U):
            b = n[v]
            f += p.alpha[u][v] * a * b + 0.5 * p.gamma[u][v] * (a * a * b + a * b * b)
    local = f.mean()
    N = np.fft.rfft2(n, axes=(1, 2)) / (Nx * Ny)          # U x Ny x (Nx/2 + 1)
    kx = np.arange(Nx // 2 + 1) * (2 * np.pi / (dx * Nx))
    j = np.arange(Ny)
    ky = np.where(j < Ny // 2, j, j - Ny) * (2 * np.pi / (dy * Ny))
    k2 = kx[None, :] ** 2 + ky[:, None] ** 2
    G = np.exp(-0.5 / p.sigma ** 2 * k2)
    e = np.zeros_like(k2)
    for u in range(U):
        P = np.ones_like(k2)
        for w in range(p.W[u]):
            P *= p.omega[u][w] * (p.nu2[u][w] - k2) ** 2 + p.phi[u][w]
        e += 0.5 * p.beta[u][u] * P * np.abs(N[u]) ** 2
        for v in range(u + 1, U):
            if p.energy_compat == ENERGY_NU2:
                c2 = p.nu2[u][v] if v < p.W[u] else 0.0
            else:
                c2 = p.mu2[u][v]
            m = p.beta[u][v] * (c2 - k2) ** 2 + p.epsilon[u][v] * G * G
            e += m * (N[u].real * N[v].real + N[u].imag * N[v].imag)
    col = np.arange(Nx // 2 + 1)
    once = (col == 0) | (2 * col == Nx)
    w = np.where(once, 1.0, 2.0)
    if weights == "kx0 twice":
        w[0] = 2.0
    elif weights == "nyquist twice":
        w[Nx // 2] = 2.0
    elif weights == "interior once":
        w[1] = 1.0
    return local + float((e * w[None, :]).sum())


CASES = {
    "graphene 64x48": (lambda: graphene_params(), 64, 48, 1),
    "graphene odd Nx 45x32": (lambda: graphene_params(), 45, 32, 1),
    "two fields two modes 48x64": (two_mode_params, 48, 64, 2),
    "hetero 64x64": (lambda: hetero_params(), 64, 64, 3),
    "hetero nu2 alias 64x40": (lambda: hetero_params(energy_compat=ENERGY_NU2), 64, 40, 3),
}


def relaxed_state(name):
    pf, Nx, Ny, U = CASES[name]
    p = pf()
    dx, dy = LATTICE_A / 8, LATTICE_A / 8 * 1.03
    if U == 1:
        f = crystal(Nx, Ny, dx, dy, -0.2, 0.2, LATTICE_A, sym=3, theta=0.2, noise=0.01, seed=3)[None]
    elif U == 2:
        rng = np.random.default_rng(2)
        f = np.stack([crystal(Nx, Ny, dx, dy, 0.2, -0.2, LATTICE_A, sym=3, noise=0.01, seed=2),
                      0.1 + 0.05 * rng.standard_normal((Ny, Nx))])
    else:
        f = hetero_patchwork(Nx, Ny, dx, dy, domain=24, seed=6)
    o = Oracle(p, Nx, Ny, dx, dy)
    o.set_fields(f)
    assert o.steps(0, 20) == 0
    return o, p, dx, dy


@pytest.mark.parametrize("name", sorted(CASES))
def test_spectral_sums_equal_real_space_evaluate(name):
    o, p, dx, dy = relaxed_state(name)
    fo, _ = o.evaluate()
    fs = spectral_energy(p, o.get_fields(), dx, dy)
    assert abs(fs - fo) <= 1e-12 * abs(fo), (fs, fo)


@pytest.mark.parametrize("Nx,Ny", [(64, 64), (64, 45), (45, 32), (45, 33)])
@pytest.mark.parametrize("U", [1, 3])
def test_spectral_sums_white_noise(Nx, Ny, U):
    p = graphene_params() if U == 1 else hetero_params()
    rng = np.random.default_rng(Nx * Ny + U)
    f = 0.3 * rng.standard_normal((U, Ny, Nx))
    o = Oracle(p, Nx, Ny, 0.7, 0.75)
    o.set_fields(f)
    fo, _ = o.evaluate()
    fs = spectral_energy(p, o.get_fields(), 0.7, 0.75)
    assert abs(fs - fo) <= 1e-12 * abs(fo), (fs, fo)


@pytest.mark.parametrize("mistake", ["kx0 twice", "nyquist twice", "interior once"])
def test_weight_mistakes_are_caught(mistake):
    o, p, dx, dy = relaxed_state("hetero 64x64")
    fo, _ = o.evaluate()
    fs = spectral_energy(p, o.get_fields(), dx, dy, weights=mistake)
    assert abs(fs - fo) > 1e-9 * abs(fo), (mistake, fs, fo)
