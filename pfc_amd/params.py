"""ctypes mirror of ``pfc_params`` (include/pfc_relax.h) and helpers that build
it from the quantities of a pfc-relax input file.

The reference reads these from the ``M`` and ``R`` blocks of its input file
(src/pfc-relax.c:541-718).  The input file stores 1/nu and 1/mu; the struct
holds 1/nu^2 and 1/mu^2 (src/pfc-relax.c:654, :699) and mu2 is 0 where
beta_uv is 0 (:698), so :func:`make_params` applies the same conversions.
"""
from __future__ import annotations

import ctypes
import math

MAX_FIELDS = 8
MAX_MODES = 8

ENERGY_MODEL = 0
ENERGY_NU2 = 1

FP64 = 0
MIXED = 1

_Mat = (ctypes.c_double * MAX_FIELDS) * MAX_FIELDS
_Modes = (ctypes.c_double * MAX_MODES) * MAX_FIELDS


class PfcParams(ctypes.Structure):
    _fields_ = [
        ("U", ctypes.c_int),
        ("W", ctypes.c_int * MAX_FIELDS),
        ("energy_compat", ctypes.c_int),
        ("reserved", ctypes.c_int),
        ("c", ctypes.c_double),
        ("sigma", ctypes.c_double),
        ("dt", ctypes.c_double),
        ("lambda_", ctypes.c_double * MAX_FIELDS),
        ("delta", ctypes.c_double * MAX_FIELDS),
        ("alpha", _Mat),
        ("beta", _Mat),
        ("gamma", _Mat),
        ("epsilon", _Mat),
        ("mu2", _Mat),
        ("nu2", _Modes),
        ("omega", _Modes),
        ("phi", _Modes),
    ]


def make_params(U, lam, alpha, beta, gamma, delta, epsilon, sigma, nu, omega, phi,
                mu, c, dt, energy_compat=ENERGY_MODEL):
    """Builds a :class:`PfcParams` from input-file quantities.

    ``alpha``, ``beta``, ``gamma`` are U x U (upper triangle incl. diagonal
    read), ``epsilon`` and ``mu`` U x U (strict upper triangle read), ``nu``,
    ``omega``, ``phi`` lists of per-field mode lists.  ``nu`` and ``mu`` are
    the reciprocal lengths exactly as written in the input file.
    """
    if not 1 <= U <= MAX_FIELDS:
        raise ValueError(f"U={U} out of range 1..{MAX_FIELDS}")
    p = PfcParams()
    p.U = U
    p.c = float(c)
    p.sigma = float(sigma)
    p.dt = float(dt)
    p.energy_compat = int(energy_compat)
    for u in range(U):
        p.lambda_[u] = float(lam[u])
        p.delta[u] = float(delta[u])
        W = len(nu[u])
        if not 1 <= W <= MAX_MODES:
            raise ValueError(f"W[{u}]={W} out of range 1..{MAX_MODES}")
        p.W[u] = W
        for w in range(W):
            p.nu2[u][w] = 1.0 / float(nu[u][w]) / float(nu[u][w])
            p.omega[u][w] = float(omega[u][w])
            p.phi[u][w] = float(phi[u][w])
        for v in range(u, U):
            p.alpha[u][v] = float(alpha[u][v])
            p.beta[u][v] = float(beta[u][v])
            p.gamma[u][v] = float(gamma[u][v])
        for v in range(u + 1, U):
            p.epsilon[u][v] = float(epsilon[u][v])
            if p.beta[u][v] == 0.0:
                p.mu2[u][v] = 0.0
            else:
                p.mu2[u][v] = 1.0 / float(mu[u][v]) / float(mu[u][v])
    return p


def graphene_params(alpha=-0.25, beta=1.0, gamma=1.0, delta=1.0, lam=0.0,
                    sigma=0.2, c=0.0, dt=1.0):
    """Single-component one-mode PFC with q0 = 1: f = n(alpha + beta(1+lap)^2)n/2
    + gamma n^3/3 + delta n^4/4 (+ lambda n)."""
    return make_params(
        U=1, lam=[lam], alpha=[[alpha]], beta=[[beta]], gamma=[[gamma]], delta=[delta],
        epsilon=[[0.0]], sigma=sigma, nu=[[1.0]], omega=[[1.0]], phi=[[0.0]],
        mu=[[0.0]], c=c, dt=dt)


def hetero_params(dt=0.25, c=1.0, sigma=0.2, energy_compat=ENERGY_MODEL):
    """The 3-field graphene/hBN (C, B, N) parameter set of the reference's
    demos/hetero/hetero.in:15-63."""
    U = 3
    alpha = [[-1.4, -0.04, -0.04], [0, -0.3, 0.5], [0, 0, -0.3]]
    beta = [[2.25, 0.0, 0.0], [0, 1.0, 0.02], [0, 0, 1.0]]
    gamma = [[0.0, 0.0, 0.0], [0, 0.0, 0.3], [0, 0, 0.0]]
    epsilon = [[0, -0.8, -0.8], [0, 0, 0.0], [0, 0, 0]]
    mu = [[0, 0.0, 0.0], [0, 0, 1.018], [0, 0, 0]]
    return make_params(
        U=U, lam=[0.0, 0.0, 0.0], alpha=alpha, beta=beta, gamma=gamma,
        delta=[2.25, 1.0, 1.0], epsilon=epsilon, sigma=sigma,
        nu=[[1.0], [1.018], [1.018]], omega=[[1.0], [1.0], [1.0]],
        phi=[[0.0], [0.0], [0.0]], mu=mu, c=c, dt=dt, energy_compat=energy_compat)


LATTICE_A = 4.0 * math.pi / math.sqrt(3.0)   # honeycomb lattice constant for q0 = 1
