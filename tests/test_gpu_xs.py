"""Cross sections of DG degree p on the GPU (tsw_set_xs_degree; the reference's CrossSection of degree p,
init.cpp:741 with piecewise_cross false) against the oracle with the same coefficient arrays:

* sweeps (general-sigma CalcA, SweepLinearSolveSpec.cpp:300-321) for DG degrees 1..3 on every pattern,
  FP64 to 1e-12, FP32 to 2e-6 and TSW_MIXED to 1e-10 (relative l2), steady and with dt;
* Smv and Lmv with the triple-product tensor (TransportOperator.cpp:375-381, :1587-1596), and the in-place
  Lmv (x == y, Lmv_InPlace :602-1200) for both kinds of cross section, host and device arrays;
* GMRES (with and without DSA) and source iteration against the oracle's solves, and one IDC step;
* coefficient arrays holding only the cell mean reproduce the cell-mean path;
* the vector-Jacobian products refuse degree-p cross sections (TSW_E_UNSUPPORTED)."""
import math

import numpy as np
import pytest

import idc
import tsw
from _common import (OracleIdcOps, ls_ordinates, orc_gmres, orc_lmv, orc_pmv, orc_si, orc_smv, orc_sweep,
                     random_poly_sigma, random_problem, rel_err, xs_degree)

pytestmark = pytest.mark.gpu


def _problem(nx, ny, p, order, seed, halo_inflow=True):
    o = ls_ordinates(order)
    g, _, _, src = random_problem(nx, ny, p, o, seed=seed, halo_inflow=halo_inflow)
    rng = np.random.default_rng(seed + 7)
    st = random_poly_sigma(g, rng, lo=0.5, hi=3.0, rel=0.6)
    ss = st * rng.uniform(0.2, 0.8, size=st.shape[:2])[..., None, None]
    return g, o, st, ss, src


def _ctx(g, o, p, pattern="hip-staggered"):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern(pattern)
    c.set_xs_degree(p)
    return c


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("pattern", ["hip-staggered", "hip-wavefront", "hip-wavefront-eager"])
def test_sweep_degree_p_sigma_matches_oracle(p, pattern):
    g, o, st, _, src = _problem(11, 9, p, 4, seed=300 + p)
    print(f"xs sweep p={p} {pattern}", flush=True)
    with _ctx(g, o, p, pattern) as c:
        for dt in (math.inf, 0.2):
            with xs_degree(p):
                ref = orc_sweep(g, o, st, src, dt)
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6), (tsw.MIXED, 1e-10)):
                out = src.copy()
                c.sweep(out, st, dt, precision=prec)
                assert rel_err(out, ref) <= tol, (dt, prec)
                np.testing.assert_array_equal(out[0], src[0])


@pytest.mark.parametrize("p", [1, 2, 3])
def test_smv_lmv_degree_p_and_in_place_lmv(p):
    import torch
    g, o, st, ss, src = _problem(6, 7, p, 6, seed=320 + p)
    rng = np.random.default_rng(p)
    phi = rng.normal(size=g.phi_shape())
    y0 = rng.normal(size=g.psi_shape(o.nq))
    with xs_degree(p):
        ref_smv = orc_smv(g, o, ss, phi, y0, alpha=0.7, beta=-1.3, all_cells=1)
        ref_lmv = orc_lmv(g, o, st, src, y0, alpha=0.4, beta=2.0)
        ref_inplace = orc_lmv(g, o, st, y0, y0.copy(), alpha=-0.6, beta=1.5)
    with _ctx(g, o, p) as c:
        y = y0.copy()
        c.smv(ss, phi, y, alpha=0.7, beta=-1.3, all_cells=True)
        assert rel_err(y, ref_smv) <= 1e-13
        y = y0.copy()
        c.lmv(st, src, y, alpha=0.4, beta=2.0)
        assert rel_err(y[1:-1, 1:-1], ref_lmv[1:-1, 1:-1]) <= 1e-13
        y = y0.copy()
        c.lmv(st, y, y, alpha=-0.6, beta=1.5)           # in place, host arrays
        assert rel_err(y[1:-1, 1:-1], ref_inplace[1:-1, 1:-1]) <= 1e-13
        np.testing.assert_array_equal(y[0], y0[0])
        d = torch.from_numpy(y0).to("cuda:0")
        c.lmv(torch.from_numpy(st).to("cuda:0"), d, d, alpha=-0.6, beta=1.5)   # in place, device arrays
        assert rel_err(d.cpu().numpy()[1:-1, 1:-1], ref_inplace[1:-1, 1:-1]) <= 1e-13


@pytest.mark.parametrize("p", [0, 2])
def test_in_place_lmv_cell_means(p):
    import torch
    o = ls_ordinates(4)
    g, st, _, src = random_problem(9, 5, p, o, seed=340 + p)
    ref = orc_lmv(g, o, st, src, src.copy(), alpha=1.25, beta=-0.5)
    with tsw.Context(0) as c:
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        y = src.copy()
        c.lmv(st, y, y, alpha=1.25, beta=-0.5)
        assert rel_err(y[1:-1, 1:-1], ref[1:-1, 1:-1]) <= 1e-13
        d = torch.from_numpy(src).to("cuda:0")
        c.lmv(torch.from_numpy(st).to("cuda:0"), d, d, alpha=1.25, beta=-0.5)
        assert rel_err(d.cpu().numpy()[1:-1, 1:-1], ref[1:-1, 1:-1]) <= 1e-13


@pytest.mark.parametrize("p", [1, 2])
@pytest.mark.parametrize("dsa", [0, 1])
def test_gmres_si_degree_p_match_oracle(p, dsa):
    g, o, st, ss, src = _problem(10, 8, p, 4, seed=360 + p)
    with xs_degree(p):
        rc, ref, ref_phi, _, _ = orc_gmres(g, o, st, ss, src, abs_tol=1e-13, restart=40)
        assert rc == 0
        ref_si, _, _ = orc_si(g, o, st, ss, src, abs_tol=1e-13, max_its=3000)
    with _ctx(g, o, p) as c:
        c.set_tuning("use_dsa", dsa)
        for prec, tol in ((tsw.FP64, 1e-11), (tsw.FP32, 1e-8)):
            out = src.copy()
            phi = np.zeros(g.phi_shape())
            its, res = c.gmres_solve(out, st, ss, phi_out=phi, abs_tol=1e-13 if prec == tsw.FP64 else 1e-11,
                                     restart=30, precision=prec)
            assert rel_err(out, ref) <= tol, prec
            assert rel_err(phi[1:-1, 1:-1], ref_phi[1:-1, 1:-1]) <= tol, prec
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)  # zero SI state
        c.set_xs_degree(p)
        out = src.copy()
        it, err = c.si_solve(out, st, ss, abs_tol=1e-13, max_its=3000)
        assert err <= 1e-13 and rel_err(out, ref_si) <= 1e-11


def test_idc_step_degree_p_matches_oracle():
    p = 1
    g, o, st, ss, src = _problem(7, 6, p, 4, seed=380, halo_inflow=False)
    rng = np.random.default_rng(381)
    psi0 = np.zeros(g.psi_shape(o.nq))
    psi0[1:-1, 1:-1] = rng.uniform(0.0, 1.0, size=psi0[1:-1, 1:-1].shape)
    with xs_degree(p):
        ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
        ref = idc.idc_step(ops, psi0.copy(), src, 0.05, num_nodes=3, num_corr=2, kind="lobatto")
    with _ctx(g, o, p) as c:
        c.idc_set(3, 2, tsw.IDC_LOBATTO, False)
        psi = psi0.copy()
        c.idc_step(psi, src, st, ss, 0.05, abs_tol=1e-13)
        assert rel_err(psi, ref) <= 1e-10


def test_mean_only_coefficients_equal_cell_means():
    p = 2
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(12, 10, p, o, seed=390, materials=3)
    stc = np.zeros(g.phi_shape())
    ssc = np.zeros(g.phi_shape())
    stc[..., 0, 0], ssc[..., 0, 0] = st, ss
    with tsw.Context(0) as c:
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        a = src.copy()
        c.sweep(a, st, 0.3)
        ga = src.copy()
        c.gmres_solve(ga, st, ss, abs_tol=1e-13)
        c.set_xs_degree(p)
        b = src.copy()
        c.sweep(b, stc, 0.3)
        gb = src.copy()
        c.gmres_solve(gb, stc, ssc, abs_tol=1e-13)
    assert rel_err(b, a) <= 1e-13
    assert rel_err(gb, ga) <= 1e-11


def test_degree_p_refusals():
    g, o, st, ss, src = _problem(5, 4, 1, 2, seed=395)
    with _ctx(g, o, 1) as c:
        out = src.copy()
        c.sweep(out, st)
        with pytest.raises(tsw.TswError) as ei:
            c.sweep_vjp(src, np.zeros_like(src), st)
        assert ei.value.code == tsw.TSW_E_UNSUPPORTED
        with pytest.raises(tsw.TswError) as ei:
            c.set_xs_degree(2)
        assert ei.value.code == tsw.TSW_E_ARG
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        assert c.xs_degree == 0   # problem_set returns to cell means
