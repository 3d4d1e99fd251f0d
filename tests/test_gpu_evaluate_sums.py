"""pfc_evaluate_sums, the free energy's sums without its density field
(k_energy_local in real space, k_energy_spec over the half spectrum), against
the oracle's evaluate (src/pfc-relax.c:987-1130) and against pfc_evaluate on
the same engine.  FP64: fave to 1e-12 of the oracle's and 1e-13 of
pfc_evaluate's, nave to 1e-12; the re-synchronised state it leaves steps on
as the oracle's does; pfc_get_energy refuses until the next pfc_evaluate.
Mixed form: fave to 1e-6 of the oracle's."""
import numpy as np
import pytest

from oracle_lib import Oracle
from pfc_amd import FP64, MIXED, ENERGY_NU2, graphene_params, hetero_params, make_params
from pfc_amd.engine import Engine, PfcError
from pfc_amd.inputs import crystal, hetero_patchwork
from test_gpu_parity import DX, graphene_state, rel_l2, two_field_params

pytestmark = pytest.mark.gpu


def five_field_params():
    U = 5
    rng = np.random.default_rng(5)
    alpha = [[0.0] * U for _ in range(U)]
    beta = [[0.0] * U for _ in range(U)]
    gamma = [[0.0] * U for _ in range(U)]
    eps = [[0.0] * U for _ in range(U)]
    mu = [[0.0] * U for _ in range(U)]
    for u in range(U):
        alpha[u][u], beta[u][u], gamma[u][u] = -0.3, 1.0, 0.5
        for v in range(u + 1, U):
            alpha[u][v] = float(rng.uniform(-0.1, 0.1))
            beta[u][v] = float(rng.uniform(0.0, 0.03))
            gamma[u][v] = float(rng.uniform(-0.2, 0.2))
            eps[u][v] = float(rng.uniform(-0.3, 0.3))
            mu[u][v] = float(rng.uniform(0.95, 1.05))
    return make_params(U=U, lam=[0.01 * u for u in range(U)], alpha=alpha, beta=beta, gamma=gamma,
                       delta=[1.0] * U, epsilon=eps, sigma=0.3,
                       nu=[[1.0 + 0.01 * u] for u in range(U)], omega=[[1.0]] * U, phi=[[0.0]] * U,
                       mu=mu, c=1.0, dt=0.5)


def state(kind, Nx, Ny):
    if kind == "graphene":
        return graphene_params(), graphene_state(Nx, Ny)
    if kind == "two fields":
        rng = np.random.default_rng(8)
        return two_field_params(), np.stack([graphene_state(Nx, Ny, seed=8)[0],
                                             0.1 + 0.05 * rng.standard_normal((Ny, Nx))])
    if kind == "hetero":
        return hetero_params(), hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=40, seed=4)
    if kind == "hetero nu2 alias":
        return (hetero_params(energy_compat=ENERGY_NU2),
                hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=40, seed=4))
    rng = np.random.default_rng(5)
    return five_field_params(), 0.05 * rng.standard_normal((5, Ny, Nx)) + 0.1


# tuned kernels (2^k and 3 * 2^k sides, U <= 4), then the generic path
# (other sides, odd Nx and odd Ny, U = 5)
CASES = [("graphene", 128, 64), ("two fields", 128, 64), ("hetero", 256, 256),
         ("hetero nu2 alias", 256, 128), ("hetero", 384, 192), ("hetero", 4096, 64),
         ("hetero", 64, 4096), ("hetero", 8192, 64), ("graphene", 100, 90),
         ("hetero", 97, 64), ("hetero", 64, 45), ("five fields", 64, 64)]


@pytest.mark.parametrize("kind,Nx,Ny", CASES)
def test_sums_match_oracle_and_full_evaluate(gpu, kind, Nx, Ny):
    p, f = state(kind, Nx, Ny)
    print(f"case {kind} {Nx}x{Ny}", flush=True)
    o = Oracle(p, Nx, Ny, 0.7, 0.75)
    o.set_fields(f)
    assert o.steps(0, 7) == 0
    with Engine(p, Nx, Ny, 0.7, 0.75, FP64) as g:
        g.set_fields(f)
        g.step(0, 7)
        fo, no = o.evaluate()
        fs, ns = g.evaluate(density=False)
        assert abs(fs - fo) <= 1e-12 * abs(fo), (fs, fo)
        assert np.allclose(ns, no, rtol=1e-12, atol=1e-15)
        with pytest.raises(PfcError) as err:
            g.get_energy()
        assert err.value.code == -5      # PFC_ESTATE
        fd, nd = g.evaluate()
        assert abs(fs - fd) <= 1e-13 * abs(fd), (fs, fd)
        assert np.allclose(ns, nd, rtol=1e-13, atol=1e-15)
        assert rel_l2(g.get_energy(), o.get_energy()) < 1e-12
        # the refresh leaves the state the oracle's evaluate leaves
        g.evaluate(density=False)
        o.steps(7, 5)
        g.step(7, 5)
        assert rel_l2(g.get_fields(), o.get_fields()) < 1e-12


def test_sums_mixed_form(gpu):
    Nx, Ny = 256, 256
    p = hetero_params()
    f = hetero_patchwork(Nx, Ny, 0.7, 0.7, domain=64, seed=9)
    o = Oracle(p, Nx, Ny, 0.7, 0.7)
    o.set_fields(f)
    assert o.steps(0, 50) == 0
    with Engine(p, Nx, Ny, 0.7, 0.7, MIXED) as g:
        g.set_fields(f)
        g.step(0, 50)
        fo, no = o.evaluate()
        fs, ns = g.evaluate(density=False)
        fd, _ = g.evaluate()
        assert abs(fs - fo) <= 1e-6 * abs(fo), (fs, fo)
        assert abs(fs - fd) <= 1e-7 * abs(fd), (fs, fd)
        assert np.allclose(ns, no, rtol=1e-6, atol=1e-7)


def test_energy_density_before_any_evaluate_is_zero(gpu):
    p, f = state("graphene", 64, 64)
    with Engine(p, 64, 64, DX, DX, FP64) as g:
        g.set_fields(f)
        assert not np.any(g.get_energy())
