"""Pins the oracle's transposed sweep (orc_sweep_adjoint) and the gradient formulas built on it, with
checks that do not restate them:

* the bilinear identity <L^{-1}(x, h), g> = <x, zeta> + <h, dJ/dh> for random sources x, inflow halos h
  and seeds g (this fixes the transpose, inflow gradient included);
* the cross-section gradient of a single sweep, dJ/dsigma_t = -sum zeta psi, against central finite
  differences of the forward oracle sweep;
* the gradients of the scattering solve (sigma_t, sigma_s, Q, inflow) against central finite
  differences of the oracle's Peierls GMRES solve.

The reference has no adjoint sweep; only its forward system is restated by the oracle, and the
transpose here is of that system.
"""
import math

import numpy as np
import pytest

from _common import (Grid, adjoint_sigma_grads, halo_mask, interior, orc_gmres, orc_pmv, orc_solve_adjoint,
                     orc_sweep, orc_sweep_adjoint, random_ordinates)


def _setup(p, seed, nx=5, ny=4, npq=2):
    rng = np.random.default_rng(seed)
    g = Grid(nx, ny, p, 0.0, 1.25, -0.5, 0.5)
    o = random_ordinates(npq, rng)
    sigma_t = rng.uniform(0.4, 3.0, size=(nx + 2, ny + 2))
    return rng, g, o, sigma_t


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("dt", [math.inf, 0.1])
def test_adjoint_bilinear_identity(p, dt):
    rng, g, o, sigma_t = _setup(p, 40 + p)
    for _ in range(3):
        x = rng.normal(size=g.psi_shape(o.nq))      # interior: source, halos: inflow
        seed = rng.normal(size=g.psi_shape(o.nq))
        psi = orc_sweep(g, o, sigma_t, x, dt)
        adj = orc_sweep_adjoint(g, o, sigma_t, seed, dt)
        lhs = np.sum(interior(psi) * interior(seed))
        h = halo_mask(x.shape)
        rhs = np.sum(interior(x) * interior(adj)) + np.sum(x[h] * adj[h])
        assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + np.abs(interior(psi) * interior(seed)).sum())


def test_adjoint_halo_zero_on_outflow_sides():
    """Inflow gradients live only on each ordinate's inflow faces; corners and outflow faces are 0."""
    rng, g, o, sigma_t = _setup(1, 5)
    adj = orc_sweep_adjoint(g, o, sigma_t, rng.normal(size=g.psi_shape(o.nq)))
    for q in range(o.nq):
        out_i = g.nx + 1 if o.xi[q] > 0 else 0
        out_j = g.ny + 1 if o.eta[q] > 0 else 0
        assert not adj[out_i, :, q].any() and not adj[:, out_j, q].any()
        for ci in (0, g.nx + 1):
            for cj in (0, g.ny + 1):
                assert not adj[ci, cj, q].any()
        in_i = 0 if o.xi[q] > 0 else g.nx + 1
        assert np.abs(adj[in_i, 1:-1, q]).max() > 0


@pytest.mark.parametrize("p", [1, 2])
def test_sweep_sigma_t_gradient_matches_finite_differences(p):
    rng, g, o, sigma_t = _setup(p, 70 + p)
    src = rng.uniform(0.0, 1.0, size=g.psi_shape(o.nq))
    seed = rng.normal(size=g.psi_shape(o.nq))
    psi = orc_sweep(g, o, sigma_t, src)
    grad = adjoint_sigma_grads(psi, orc_sweep_adjoint(g, o, sigma_t, seed))

    def J(st):
        return np.sum(interior(orc_sweep(g, o, st, src)) * interior(seed))

    h = 1e-6
    for (i, j) in [(1, 1), (3, 2), (g.nx, g.ny), (2, g.ny)]:
        sp, sm = sigma_t.copy(), sigma_t.copy()
        sp[i, j] += h
        sm[i, j] -= h
        fd = (J(sp) - J(sm)) / (2 * h)
        assert abs(fd - grad[i, j]) <= 1e-7 * max(1.0, abs(fd))
    assert not grad[0].any() and not grad[:, 0].any()


def test_solve_gradients_match_finite_differences():
    rng, g, o, sigma_t = _setup(1, 99, nx=4, ny=4)
    sigma_s = sigma_t * rng.uniform(0.2, 0.8, size=sigma_t.shape)
    src = rng.uniform(0.0, 1.0, size=g.psi_shape(o.nq))   # halos: inflow
    seed = rng.normal(size=g.psi_shape(o.nq))

    def solve(st, ss, q):
        rc, psi, phi, _, _ = orc_gmres(g, o, st, ss, q, abs_tol=1e-14, restart=40)
        assert rc == 0
        return psi

    def J(st, ss, q):
        return np.sum(interior(solve(st, ss, q)) * interior(seed))

    psi = solve(sigma_t, sigma_s, src)
    phi = orc_pmv(g, o, psi)
    zeta = orc_solve_adjoint(g, o, sigma_t, sigma_s, seed)
    gt, gs = adjoint_sigma_grads(psi, zeta, phi)
    h = 1e-6
    for (i, j) in [(1, 1), (2, 3), (4, 4)]:
        for arr, grad in ((sigma_t, gt), (sigma_s, gs)):
            a_p, a_m = arr.copy(), arr.copy()
            a_p[i, j] += h
            a_m[i, j] -= h
            if arr is sigma_t:
                fd = (J(a_p, sigma_s, src) - J(a_m, sigma_s, src)) / (2 * h)
            else:
                fd = (J(sigma_t, a_p, src) - J(sigma_t, a_m, src)) / (2 * h)
            assert abs(fd - grad[i, j]) <= 1e-6 * max(1.0, abs(fd)), (i, j, fd, grad[i, j])
    # source (interior) and inflow (halo) entries: J is linear in them, one-sided differences are exact
    for idx in [(2, 2, 1, 0, 1), (1, 3, 0, 1, 1), (0, 2, 0, 1, 0), (g.nx + 1, 3, 3, 0, 1), (2, 0, 1, 1, 1)]:
        qp = src.copy()
        qp[idx] += 1.0
        fd = J(sigma_t, sigma_s, qp) - J(sigma_t, sigma_s, src)
        assert abs(fd - zeta[idx]) <= 1e-8 * max(1.0, abs(fd)), (idx, fd, zeta[idx])
