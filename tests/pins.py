"""Closed-form checks of two parts of the reference's loop, shared by
tests/test_oracle.py (run against the oracle) and
tests/test_oracle_mutants.py (run against oracle builds with a planted
mistake, which they must reject).

- pair_gamma_error: the explicit pair terms of step()'s nonlinear part
  (src/pfc-relax.c:1191-1205).  For uniform fields only k = 0 survives, and
  the semi-implicit scheme reduces to a scalar recursion per field whose
  explicit part is the derivative of the free energy (the model of
  src/pfc-relax.c:9) minus the implicit linear part.  The derivative is taken
  by complex-step differentiation of the energy density, so the check does
  not restate the step's formula.
- cadence_errors: the event order of relax() (src/pfc-relax.c:1292-1312):
  optimise, then evaluate and print, then write, then step.  A linear
  single-mode model has closed-form fields and energies at every step.
"""
from __future__ import annotations

import math

import numpy as np

from oracle_lib import Oracle
from pfc_amd import make_params


# ------------------------------------------------------------- pair gamma

def _uniform_energy(n, lam, alpha, beta, gamma, delta, eps, mu2, P0):
    """Free energy density of uniform fields n (complex allowed).  At k = 0
    the operators reduce to P_u(0) and (mu2_uv)^2, and G(0) = 1."""
    U = len(n)
    f = 0.0
    for u in range(U):
        f = f + lam[u] * n[u] + 0.5 * (alpha[u][u] + beta[u][u] * P0[u]) * n[u] ** 2 \
            + gamma[u][u] * n[u] ** 3 / 3.0 + delta[u] * n[u] ** 4 / 4.0
        for v in range(u + 1, U):
            f = f + alpha[u][v] * n[u] * n[v] + beta[u][v] * mu2[u][v] ** 2 * n[u] * n[v] \
                + 0.5 * gamma[u][v] * (n[u] ** 2 * n[v] + n[u] * n[v] ** 2) \
                + eps[u][v] * n[u] * n[v]
    return f


def pair_gamma_case():
    """Three nonconserved fields with every pair coupling switched on and
    distinct gamma_uv, alpha_uv per pair."""
    U = 3
    lam = [0.01, -0.02, 0.015]
    alpha = [[-0.3, 0.07, -0.05], [0, -0.2, 0.04], [0, 0, -0.25]]
    beta = [[1.0, 0.03, 0.0], [0, 0.8, 0.02], [0, 0, 1.2]]
    gamma = [[0.4, 0.9, -0.6], [0, -0.3, 0.5], [0, 0, 0.2]]
    delta = [1.0, 0.7, 1.3]
    eps = [[0, -0.2, 0.1], [0, 0, 0.15], [0, 0, 0]]
    mu = [[0, 1.05, 0.0], [0, 0, 0.97], [0, 0, 0]]
    nu = [[1.0], [1.0 / 0.9], [1.1]]
    p = make_params(U=U, lam=lam, alpha=alpha, beta=beta, gamma=gamma, delta=delta,
                    epsilon=eps, sigma=0.3, nu=nu, omega=[[1.0], [0.8], [1.1]],
                    phi=[[0.0], [0.05], [0.0]], mu=mu, c=0.0, dt=0.3)
    n0 = np.array([0.21, -0.13, 0.34])
    return p, n0


def pair_gamma_expected(p, n0, T):
    U = p.U
    lam = [p.lambda_[u] for u in range(U)]
    A = [[p.alpha[u][v] for v in range(U)] for u in range(U)]
    B = [[p.beta[u][v] for v in range(U)] for u in range(U)]
    G = [[p.gamma[u][v] for v in range(U)] for u in range(U)]
    E = [[p.epsilon[u][v] for v in range(U)] for u in range(U)]
    M2 = [[p.mu2[u][v] for v in range(U)] for u in range(U)]
    dl = [p.delta[u] for u in range(U)]
    P0 = []
    for u in range(U):
        P = 1.0
        for w in range(p.W[u]):
            P *= p.omega[u][w] * p.nu2[u][w] ** 2 + p.phi[u][w]
        P0.append(P)
    a = [A[u][u] + B[u][u] * P0[u] for u in range(U)]
    Ad = [math.exp(-a[u] * p.dt) for u in range(U)]
    n = np.array(n0, dtype=float)
    h = 1e-30
    for _ in range(T):
        g = np.empty(U)
        for u in range(U):
            z = n.astype(complex)
            z[u] += 1j * h
            dfdn = _uniform_energy(z, lam, A, B, G, dl, E, M2, P0).imag / h
            g[u] = dfdn - a[u] * n[u]
        n = np.array([Ad[u] * n[u] + (Ad[u] - 1.0) / a[u] * g[u] for u in range(U)])
    return n


def pair_gamma_error(so=None, T=12):
    """Largest deviation of the oracle's uniform fields after T steps from
    the recursion built from the energy's derivative."""
    p, n0 = pair_gamma_case()
    Nx = Ny = 16
    o = Oracle(p, Nx, Ny, 0.8, 0.8, so=so)
    o.set_fields(np.broadcast_to(n0[:, None, None], (p.U, Ny, Nx)).copy())
    assert o.steps(1, T) == 0
    got = o.get_fields()
    o.close()
    ref = pair_gamma_expected(p, n0, T)
    return float(np.max(np.abs(got - ref[:, None, None])))


# ----------------------------------------------------------------- cadence

CAD = dict(Nx=32, Ny=24, dx=0.97, dy=0.95, alpha=-0.1, beta=1.0, dt=0.5, nbar=0.05,
           amp=0.2, ix=5, iy=0)


def _cad_params():
    return make_params(U=1, lam=[0.0], alpha=[[CAD["alpha"]]], beta=[[CAD["beta"]]],
                       gamma=[[0.0]], delta=[0.0], epsilon=[[0.0]], sigma=0.3, nu=[[1.0]],
                       omega=[[1.0]], phi=[[0.0]], mu=[[0.0]], c=0.0, dt=CAD["dt"])


def _a_of(k2):
    return CAD["alpha"] + CAD["beta"] * (1.0 - k2) ** 2


def _k2(dx, dy):
    kx = 2 * math.pi * CAD["ix"] / (CAD["Nx"] * dx)
    ky = 2 * math.pi * CAD["iy"] / (CAD["Ny"] * dy)
    return kx * kx + ky * ky


def cad_field(t, sched):
    """n(t) = nbar g0^t + amp prod g(k)^steps cos(k.x): each mode of the
    linear nonconserved model is multiplied by exp(-a(k) dt) per step, with k
    from the box in force at that step.  sched: [(t_from, dx, dy), ...] in
    ascending t_from, the first at 0."""
    Nx, Ny = CAD["Nx"], CAD["Ny"]
    g0 = math.exp(-_a_of(0.0) * CAD["dt"])
    amp = CAD["amp"]
    for i, (t_from, bx, by) in enumerate(sched):
        t_to = sched[i + 1][0] if i + 1 < len(sched) else t
        n = max(0, min(t, t_to) - t_from)
        amp *= math.exp(-_a_of(_k2(bx, by)) * CAD["dt"]) ** n
    y, x = np.meshgrid(np.arange(Ny), np.arange(Nx), indexing="ij")
    ph = 2 * math.pi * (CAD["ix"] * x / Nx + CAD["iy"] * y / Ny)
    return CAD["nbar"] * g0 ** t + amp * np.cos(ph), CAD["nbar"] * g0 ** t, amp


def cad_energy(nbar, amp, dx, dy):
    """Free energy per area of nbar + amp cos(k.x) with k from the box:
    a(0) nbar^2 / 2 + a(k) amp^2 / 4 (<cos^2> = 1/2)."""
    return 0.5 * _a_of(0.0) * nbar ** 2 + 0.25 * _a_of(_k2(dx, dy)) * amp ** 2


def _parabola_vertex(f0, fm, fp, x0, d):
    """Vertex of the parabola through (x0 - d, fm), (x0, f0), (x0 + d, fp):
    numpy's least-squares quadratic through the three points in the unit
    variable s = (x - x0)/d, then back to x."""
    c2, c1, _ = np.polyfit([-1.0, 0.0, 1.0], [fm, f0, fp], 2)
    return x0 + d * (-c1 / (2 * c2))


def cadence_errors(so=None):
    """Runs relax() with Tprint = 7, writes at [0 5 14 21 28] (tend = 28) and
    box searches (x only) every 14 steps; returns the largest deviations of
    the printed log, the printed box and the snapshots from the closed forms,
    and which steps printed."""
    Nx, Ny, dx, dy = CAD["Nx"], CAD["Ny"], CAD["dx"], CAD["dy"]
    p = _cad_params()
    sched = [(0, dx, dy)]
    f0, _, _ = cad_field(0, sched)
    tend, Tprint, Topt = 28, 7, 14
    snaps_at = (0, 5, 14, 21, 28)
    o = Oracle(p, Nx, Ny, dx, dy, so=so)
    o.set_fields(f0[None])
    log, snaps = o.relax(tend=tend, Tprint=Tprint, Topt=Topt, xopt=1, yopt=0,
                         snap_steps=snaps_at)
    o.close()
    # each box search samples the energy of the field at that step at dx and
    # dx -/+ d and moves dx to the vertex (limited to l0/4 of size change);
    # d starts at 1e-4 and halves when the move is below d, else doubles
    d = 1e-4
    l0 = 4 * math.pi / math.sqrt(3)
    for ts in range(Topt, tend + 1, Topt):
        _, nb, am = cad_field(ts, sched)
        x0 = sched[-1][1]
        fs = [cad_energy(nb, am, x, dy) for x in (x0, x0 - d, x0 + d)]
        x1 = _parabola_vertex(fs[0], fs[1], fs[2], x0, d)
        if abs(Nx * (x1 - x0)) > 0.25 * l0:
            x1 = x0 + math.copysign(0.25 * l0 / Nx, x1 - x0)
        d = d * 0.5 if abs(x1 - x0) < d else d * 2.0
        sched.append((ts, x1, dy))

    def box_at(t):
        return [s for s in sched if s[0] <= t][-1][1:]

    err_log = 0.0
    err_box = 0.0
    for row in log:
        t = int(round(row[0]))
        bx = box_at(t)
        _, nb, am = cad_field(t, sched)
        fe = cad_energy(nb, am, *bx)
        err_log = max(err_log, abs(row[3] - fe), abs(row[4] - nb))
        err_box = max(err_box, abs(row[1] - bx[0]), abs(row[2] - bx[1]))
    err_snap = 0.0
    for s, t in zip(snaps, snaps_at):
        ref, _, _ = cad_field(t, sched)
        err_snap = max(err_snap, float(np.max(np.abs(s[0] - ref))))
    return {"printed": [int(round(r[0])) for r in log], "log": err_log, "box": err_box,
            "snap": err_snap, "boxes": [s[1] for s in sched]}
