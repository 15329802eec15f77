"""Pins the CPU oracle (oracle/oracle.c) with checks that do not restate its formulas:

* element solver vs numpy.linalg.solve (LAPACK, partial pivoting);
* sweep vs a dense global solve of the upwind-DG weak form assembled by Gauss quadrature;
* polynomial exactness of the sweep (manufactured Q_p solutions) for every quadrant;
* Lmv on manufactured solutions equals Omega.grad psi + sigma psi, and Lmv o Sweep = I;
* Pmv / Smv closed forms (sum of weights, isotropic broadcast);
* source iteration: infinite-medium closed form phi = q/(sigma_t - sigma_s); particle balance;
* Peierls GMRES converges to the source-iteration fixed point.

The reference ships no golden vectors or tests for this path; it is compiled C++ that needs LAPACK
(src/utils/Quadrule/Quadrule.cpp:1240) and so is not built here.
"""
import math

import numpy as np
import pytest

from _common import (Grid, dg_weak_form_solve, oracle, orc_gmres, orc_lmv, orc_pmv, orc_si, orc_smv,
                     orc_sweep, orc_sweep_omp, project_field, ptr, random_ordinates)


@pytest.mark.parametrize("N", [1, 4, 9, 16])
def test_ge_matches_lapack(N):
    rng = np.random.default_rng(N)
    for _ in range(20):
        A = rng.normal(size=(N, N)) + (N + 2) * np.eye(N)
        b = rng.normal(size=N)
        ref = np.linalg.solve(A, b)
        Acm = np.asfortranarray(A).ravel(order="F").copy()
        x = b.copy()
        oracle().orc_ge_solve(N, ptr(Acm), ptr(x))
        np.testing.assert_allclose(x, ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("dt", [math.inf, 0.05])
def test_sweep_matches_weak_form(p, dt):
    rng = np.random.default_rng(100 + p)
    g = Grid(3, 4, p, 0.0, 1.5, -0.5, 1.0)
    o = random_ordinates(2, rng)
    sigma = rng.uniform(0.2, 3.0, size=(g.nx + 2, g.ny + 2))
    src = rng.normal(size=g.psi_shape(o.nq))
    # interior source, halo inflow values, all random
    psi = orc_sweep(g, o, sigma, src, dt)
    inv_dt = 0.0 if math.isinf(dt) else 1.0 / dt
    for q in range(o.nq):
        ref = dg_weak_form_solve(g, o.xi[q], o.eta[q], sigma, src[:, :, q], src[:, :, q], inv_dt)
        np.testing.assert_allclose(psi[1:-1, 1:-1, q], ref[1:-1, 1:-1], rtol=1e-10, atol=1e-11)
    # halos untouched
    np.testing.assert_array_equal(psi[0], src[0])
    np.testing.assert_array_equal(psi[-1], src[-1])


def manufactured(g, o, p_exact, sigma_val, rng):
    """psi_q(x,y) polynomial of degree p_exact in x and y (different per ordinate)."""
    coefs = rng.normal(size=(o.nq, p_exact + 1, p_exact + 1))

    def f(q):
        c = coefs[q]
        return lambda X, Y: sum(c[a, b] * X ** a * Y ** b for a in range(p_exact + 1) for b in range(p_exact + 1))

    def grad_term(q):
        c = coefs[q]
        xi, eta = o.xi[q], o.eta[q]

        def h(X, Y):
            s = 0.0 * X
            for a in range(p_exact + 1):
                for b in range(p_exact + 1):
                    if a:
                        s = s + xi * c[a, b] * a * X ** (a - 1) * Y ** b
                    if b:
                        s = s + eta * c[a, b] * b * X ** a * Y ** (b - 1)
            return s + sigma_val * f(q)(X, Y)
        return h

    psi = np.zeros(g.psi_shape(o.nq))
    src = np.zeros(g.psi_shape(o.nq))
    for q in range(o.nq):
        psi[:, :, q] = project_field(f(q), g)
        src[:, :, q] = project_field(grad_term(q), g)
    return psi, src


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_sweep_polynomial_exactness(p):
    rng = np.random.default_rng(7 + p)
    g = Grid(5, 4, p, -1.0, 1.0, 0.0, 0.8)
    o = random_ordinates(2, rng)
    sig = 1.3
    psi_exact, src = manufactured(g, o, p, sig, rng)
    # the sweep's input: source in the interior, exact solution (inflow) in the halo
    inp = src.copy()
    inp[0], inp[-1], inp[:, 0], inp[:, -1] = psi_exact[0], psi_exact[-1], psi_exact[:, 0], psi_exact[:, -1]
    sigma = np.full((g.nx + 2, g.ny + 2), sig)
    out = orc_sweep(g, o, sigma, inp)
    np.testing.assert_allclose(out[1:-1, 1:-1], psi_exact[1:-1, 1:-1], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_lmv_manufactured_and_inverse(p):
    rng = np.random.default_rng(11 + p)
    g = Grid(4, 3, p, 0.0, 2.0, 0.0, 1.0)
    o = random_ordinates(2, rng)
    sig = 0.7
    psi_exact, src = manufactured(g, o, p, sig, rng)
    sigma = np.full((g.nx + 2, g.ny + 2), sig)
    y = orc_lmv(g, o, sigma, psi_exact)
    np.testing.assert_allclose(y[1:-1, 1:-1], src[1:-1, 1:-1], rtol=1e-9, atol=1e-9)
    # L o L^{-1} = I on interior for arbitrary data (random sigma, random inflow)
    sigma = rng.uniform(0.1, 5.0, size=(g.nx + 2, g.ny + 2))
    q = rng.normal(size=g.psi_shape(o.nq))
    psi = orc_sweep(g, o, sigma, q)
    back = orc_lmv(g, o, sigma, psi)
    np.testing.assert_allclose(back[1:-1, 1:-1], q[1:-1, 1:-1], rtol=1e-10, atol=1e-10)


def test_pmv_smv_closed_forms():
    rng = np.random.default_rng(3)
    g = Grid(3, 3, 2)
    o = random_ordinates(3, rng)
    # isotropic psi = c(x) -> phi = 4 pi c
    c = rng.normal(size=g.phi_shape())
    psi = np.repeat(c[:, :, None], o.nq, axis=2)
    phi = orc_pmv(g, o, psi)
    np.testing.assert_allclose(phi[1:-1, 1:-1], 4 * math.pi * c[1:-1, 1:-1], rtol=1e-13)
    assert np.all(phi[0] == 0) and np.all(phi[:, -1] == 0)
    phi2 = orc_pmv(g, o, psi, phi=np.ones(g.phi_shape()), alpha=-2.0, beta=3.0, all_cells=1)
    np.testing.assert_allclose(phi2, 3.0 - 8 * math.pi * c, rtol=1e-12, atol=1e-12)
    # Smv then Pmv: P S phi = sigma_s phi (sum w = 4 pi)
    ss = rng.uniform(0, 2, size=(g.nx + 2, g.ny + 2))
    out = orc_smv(g, o, ss, c, np.zeros(g.psi_shape(o.nq)), beta=0.0)
    back = orc_pmv(g, o, out)
    np.testing.assert_allclose(back[1:-1, 1:-1], ss[1:-1, 1:-1, None, None] * c[1:-1, 1:-1], rtol=1e-12)
    assert np.all(out[0] == 0)


@pytest.mark.parametrize("p", [0, 1, 2])
def test_si_infinite_medium(p):
    rng = np.random.default_rng(5)
    g = Grid(6, 5, p)
    o = random_ordinates(3, rng)
    st, ss, q0 = 2.0, 1.2, 0.8
    phi_inf = q0 / (st - ss)
    src = np.zeros(g.psi_shape(o.nq))
    src[..., 0, 0] = q0 / (4 * math.pi)
    # inflow = the infinite-medium angular flux in every halo cell
    halo = np.ones((g.nx + 2, g.ny + 2), bool)
    halo[1:-1, 1:-1] = False
    src[halo] = 0.0
    src[halo, :, 0, 0] = phi_inf / (4 * math.pi)
    sigma_t = np.full((g.nx + 2, g.ny + 2), st)
    sigma_s = np.full((g.nx + 2, g.ny + 2), ss)
    psi, iters, err = orc_si(g, o, sigma_t, sigma_s, src, abs_tol=1e-13, max_its=400)
    assert err <= 1e-13 and iters < 400
    phi = orc_pmv(g, o, psi)
    np.testing.assert_allclose(phi[1:-1, 1:-1, 0, 0], phi_inf, rtol=1e-11)
    if p:
        assert np.abs(phi[1:-1, 1:-1, 1:, :]).max() < 1e-11
        assert np.abs(phi[1:-1, 1:-1, :, 1:]).max() < 1e-11


def test_si_balance_and_gmres_fixed_point():
    """Vacuum boundaries: absorption + leakage = emission; GMRES reaches the SI fixed point."""
    rng = np.random.default_rng(9)
    p = 1
    g = Grid(6, 6, p)
    o = random_ordinates(3, rng)
    sigma_t = rng.uniform(0.5, 2.0, size=(g.nx + 2, g.ny + 2))
    sigma_s = sigma_t * rng.uniform(0.1, 0.8, size=sigma_t.shape)
    src = np.zeros(g.psi_shape(o.nq))
    src[1:-1, 1:-1, :, 0, 0] = rng.uniform(0, 1, size=(g.nx, g.ny))[:, :, None] / (4 * math.pi)
    psi, iters, err = orc_si(g, o, sigma_t, sigma_s, src, abs_tol=1e-13, max_its=1000)
    assert err <= 1e-13
    phi = orc_pmv(g, o, psi)
    area = g.dx * g.dy
    emission = 4 * math.pi * np.sum(src[1:-1, 1:-1, 0, 0, 0]) * area
    absorption = np.sum((sigma_t - sigma_s)[1:-1, 1:-1] * phi[1:-1, 1:-1, 0, 0]) * area
    # outgoing traces on the four boundaries (Legendre trace at +-1: sum of (+-1)^a c_a)
    P = p + 1
    sp = np.ones(P)
    sm = np.array([(-1.0) ** a for a in range(P)])
    leak = 0.0
    for q in range(o.nq):
        xi, eta, w = o.xi[q], o.eta[q], o.w[q]
        c = psi[:, :, q]
        if xi > 0:   # right boundary, trace at x=+1, integrate in y: 2 * c[.,0] * dy/2
            leak += w * xi * np.sum(np.einsum("jab,a->jb", c[-2, 1:-1], sp)[:, 0]) * g.dy
        else:
            leak += w * -xi * np.sum(np.einsum("jab,a->jb", c[1, 1:-1], sm)[:, 0]) * g.dy
        if eta > 0:
            leak += w * eta * np.sum(np.einsum("iab,b->ia", c[1:-1, -2], sp)[:, 0]) * g.dx
        else:
            leak += w * -eta * np.sum(np.einsum("iab,b->ia", c[1:-1, 1], sm)[:, 0]) * g.dx
    assert abs(absorption + leak - emission) < 1e-10 * emission

    rc, psi_g, phi_g, its, res = orc_gmres(g, o, sigma_t, sigma_s, src, abs_tol=1e-12, restart=10)
    assert rc == 0 and res <= 1e-12
    np.testing.assert_allclose(phi_g[1:-1, 1:-1], phi[1:-1, 1:-1], rtol=1e-9, atol=1e-11)
    np.testing.assert_allclose(psi_g[1:-1, 1:-1], psi[1:-1, 1:-1], rtol=1e-9, atol=1e-11)


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_threaded_sweep_bitwise_equal(threads):
    """The multi-core baseline sweep distributes each diagonal's cells and ordinates over threads with a
    barrier per diagonal; every cell's arithmetic is unchanged, so it equals the serial sweep bitwise."""
    rng = np.random.default_rng(17)
    g = Grid(13, 9, 2, 0.0, 1.3, 0.0, 0.9)
    o = random_ordinates(3, rng)
    st = rng.uniform(0.4, 3.0, size=(g.nx + 2, g.ny + 2))
    src = rng.normal(size=g.psi_shape(o.nq))
    np.testing.assert_array_equal(orc_sweep_omp(g, o, st, src, 0.2, threads), orc_sweep(g, o, st, src, 0.2))
