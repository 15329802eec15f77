"""Pins the oracle's cross sections of DG degree p (orc_set_xs_degree; init.cpp:741 with piecewise_cross
false), with checks that do not restate its formulas:

* the triple-product tensor (ComputeTPI global.cpp:67-107, ComputeTensor CrossSection.cpp:165-189) against
  numpy Gauss quadrature of the Legendre triple products, and its closed form for a constant sigma;
* the sweep (general-sigma CalcA, SweepLinearSolveSpec.cpp:300-321) against a dense global solve of the
  upwind-DG weak form whose int sigma psi v term integrates sigma's expansion by quadrature;
* Smv (TransportOperator.cpp:375-381) against the quadrature L2 projection of sigma phi / (4 pi);
* Lmv (:1587-1596) on manufactured polynomials with a polynomial sigma, and Lmv o Sweep = I;
* a sigma with only its mean coefficient gives the cell-mean results (sweep, Smv, Lmv, GMRES);
* GMRES and source iteration reach the same fixed point with polynomial cross sections.
"""
import math

import numpy as np
import pytest

from _common import (Grid, dg_weak_form_solve, gauss, leg, orc_gmres, orc_lmv, orc_pmv, orc_si, orc_smv, orc_sweep,
                     orc_sweep_omp, orc_xs_tensor, project_field, random_ordinates, random_poly_sigma, rel_err,
                     xs_degree)


def _tpi_quadrature(p):
    xg, wg = gauss(3 * p + 4)
    P = p + 1
    L = np.array([leg(a, xg) for a in range(P)])
    return np.einsum("g,ag,bg,cg->abc", wg, L, L, L)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_tensor_against_quadrature(p):
    rng = np.random.default_rng(40 + p)
    g = Grid(2, 3, p)
    sig = random_poly_sigma(g, rng)
    t = orc_xs_tensor(g, sig)          # [i][j][d*P+e][a*P+b] = sigma(i,j,a,b,d,e)
    T = _tpi_quadrature(p)
    P = p + 1
    ref = np.einsum("ijuv,adu,bev->ijdeab", sig, T, T).reshape(g.nx + 2, g.ny + 2, P * P, P * P)
    np.testing.assert_allclose(t, ref, rtol=1e-13, atol=1e-14)
    # constant sigma: sigma(i,j,a,b,d,e) = s * 2/(2a+1) d_ad * 2/(2b+1) d_be
    const = np.zeros_like(sig)
    const[..., 0, 0] = 1.7
    t = orc_xs_tensor(g, const)
    m = np.array([2.0 / (2 * a + 1) for a in range(P)])
    diag = 1.7 * np.einsum("a,b->ab", m, m).ravel()
    np.testing.assert_allclose(t[1, 1], np.diag(diag), rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("dt", [math.inf, 0.05])
def test_sweep_general_sigma_matches_weak_form(p, dt):
    rng = np.random.default_rng(200 + p)
    g = Grid(3, 4, p, 0.0, 1.5, -0.5, 1.0)
    o = random_ordinates(2, rng)
    sig = random_poly_sigma(g, rng, rel=0.6)
    src = rng.normal(size=g.psi_shape(o.nq))
    with xs_degree(p):
        psi = orc_sweep(g, o, sig, src, dt)
        psi_omp = orc_sweep_omp(g, o, sig, src, dt, threads=3)
    np.testing.assert_array_equal(psi, psi_omp)
    inv_dt = 0.0 if math.isinf(dt) else 1.0 / dt
    for q in range(o.nq):
        ref = dg_weak_form_solve(g, o.xi[q], o.eta[q], None, src[:, :, q], src[:, :, q], inv_dt, sigma_coef=sig)
        np.testing.assert_allclose(psi[1:-1, 1:-1, q], ref[1:-1, 1:-1], rtol=1e-10, atol=1e-11)
    np.testing.assert_array_equal(psi[0], src[0])


def _poly2(c):
    """f(x, y) = sum c[a, b] x^a y^b."""
    return lambda X, Y: sum(c[a, b] * X ** a * Y ** b for a in range(c.shape[0]) for b in range(c.shape[1]))


@pytest.mark.parametrize("p", [1, 2])
def test_smv_general_sigma_is_projection(p):
    rng = np.random.default_rng(60 + p)
    g = Grid(3, 2, p, -1.0, 1.0, 0.0, 1.0)
    o = random_ordinates(2, rng)
    fs = _poly2(rng.uniform(0.1, 0.4, size=(p + 1, p + 1)) + np.eye(p + 1)[0:1].T @ np.eye(p + 1)[0:1])
    fp = _poly2(rng.normal(size=(p + 1, p + 1)))
    sig = project_field(fs, g)          # exact: sigma is a polynomial of degree p in each variable
    phi = project_field(fp, g)
    ref = project_field(lambda X, Y: fs(X, Y) * fp(X, Y) / (4 * math.pi), g)   # projection of the product
    base = rng.normal(size=g.psi_shape(o.nq))
    with xs_degree(p):
        out = orc_smv(g, o, sig, phi, base, alpha=-0.5, beta=2.0, all_cells=1)
    for q in range(o.nq):
        np.testing.assert_allclose(out[:, :, q], 2.0 * base[:, :, q] - 0.5 * ref, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("p", [1, 2, 3])
def test_lmv_general_sigma_manufactured_and_inverse(p):
    rng = np.random.default_rng(80 + p)
    g = Grid(4, 3, p, 0.0, 2.0, 0.0, 1.0)
    o = random_ordinates(2, rng)
    cs = np.zeros((p + 1, p + 1))
    cs[0, 0], cs[1, 0], cs[0, 1] = 1.5, 0.2, -0.3
    fs = _poly2(cs)
    sig = project_field(fs, g)
    coefs = rng.normal(size=(o.nq, p + 1, p + 1))
    x = np.zeros(g.psi_shape(o.nq))
    want = np.zeros(g.psi_shape(o.nq))
    for q in range(o.nq):
        c = coefs[q]
        f = _poly2(c)
        xi, eta = o.xi[q], o.eta[q]

        def h(X, Y, c=c, f=f, xi=xi, eta=eta):
            s = 0.0 * X
            for a in range(p + 1):
                for b in range(p + 1):
                    if a:
                        s = s + xi * c[a, b] * a * X ** (a - 1) * Y ** b
                    if b:
                        s = s + eta * c[a, b] * b * X ** a * Y ** (b - 1)
            return s + fs(X, Y) * f(X, Y)
        x[:, :, q] = project_field(f, g)
        want[:, :, q] = project_field(h, g)
    with xs_degree(p):
        y = orc_lmv(g, o, sig, x)
        np.testing.assert_allclose(y[1:-1, 1:-1], want[1:-1, 1:-1], rtol=1e-9, atol=1e-9)
        sigr = random_poly_sigma(g, rng, rel=0.5)
        src = rng.normal(size=g.psi_shape(o.nq))
        back = orc_lmv(g, o, sigr, orc_sweep(g, o, sigr, src))
    np.testing.assert_allclose(back[1:-1, 1:-1], src[1:-1, 1:-1], rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("p", [1, 2])
def test_mean_only_sigma_equals_cell_means(p):
    rng = np.random.default_rng(90 + p)
    g = Grid(4, 5, p)
    o = random_ordinates(2, rng)
    st = rng.uniform(0.5, 2.0, size=(g.nx + 2, g.ny + 2))
    ss = st * rng.uniform(0.1, 0.8, size=st.shape)
    src = rng.normal(size=g.psi_shape(o.nq))
    stc = np.zeros(g.phi_shape())
    ssc = np.zeros(g.phi_shape())
    stc[..., 0, 0] = st
    ssc[..., 0, 0] = ss
    phi = rng.normal(size=g.phi_shape())
    ref_sw, ref_smv, ref_lmv = orc_sweep(g, o, st, src, 0.3), orc_smv(g, o, ss, phi, src), orc_lmv(g, o, st, src)
    ref_gm = orc_gmres(g, o, st, ss, src, abs_tol=1e-13)
    with xs_degree(p):
        assert rel_err(orc_sweep(g, o, stc, src, 0.3), ref_sw) <= 1e-14
        assert rel_err(orc_smv(g, o, ssc, phi, src), ref_smv) <= 1e-14
        assert rel_err(orc_lmv(g, o, stc, src), ref_lmv) <= 1e-14
        gm = orc_gmres(g, o, stc, ssc, src, abs_tol=1e-13)
    assert gm[0] == 0 and rel_err(gm[1], ref_gm[1]) <= 1e-12


def test_gmres_and_si_fixed_point_with_polynomial_sigma():
    p = 2
    rng = np.random.default_rng(7)
    g = Grid(5, 4, p)
    o = random_ordinates(3, rng)
    st = random_poly_sigma(g, rng, lo=1.0, hi=2.0, rel=0.5)
    ss = st * 0.6
    src = rng.uniform(0.0, 1.0, size=g.psi_shape(o.nq))
    with xs_degree(p):
        rc, psi_g, phi_g, its, res = orc_gmres(g, o, st, ss, src, abs_tol=1e-12)
        psi_s, it_s, err_s = orc_si(g, o, st, ss, src, abs_tol=1e-13, max_its=2000)
        assert rc == 0 and res <= 1e-12
        phi_s = orc_pmv(g, o, psi_s)
        assert rel_err(phi_s[1:-1, 1:-1], phi_g[1:-1, 1:-1]) <= 1e-10
        # the fixed point: psi = L^{-1}(Q + S P psi)
        again = orc_sweep(g, o, st, orc_smv(g, o, ss, orc_pmv(g, o, psi_g), src))
    assert rel_err(again[1:-1, 1:-1], psi_g[1:-1, 1:-1]) <= 1e-10
