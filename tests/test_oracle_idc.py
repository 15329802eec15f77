"""Pins the oracle's IDC step (oracle/idc.py) without restating it:

* quadrature tables: Lobatto nodes (endpoints + roots of P'_{m-1}), rows of W integrate polynomials of
  degree < m exactly over each subinterval, and the rows 1.. sum to row 0; Radau IIA nodes in closed form
  for 2 and 3 nodes, and for every count their row 0 integrates degree 2m-2 exactly (the property that
  singles out the Radau points); Chebyshev and Chebyshev-Radau nodes in closed form;
* the library's host-only tables (tsw_idc_nodes) equal these to rounding for every family and count;
* order of accuracy: on a small spatial problem the semi-discrete system d psi/dt = Q_eff - B psi is
  solved exactly with a matrix exponential (B assembled column by column from the oracle's Lmv and
  S P); 4 Lobatto nodes with 3 corrections must converge at fourth order in dt, 1 correction at second;
* the update form gives the error form's step to the solve tolerance, and Radau nodes converge at the
  order their corrections give;
* a steady state of the transport problem is kept by the step.
"""
import math

import numpy as np
import pytest
from scipy.linalg import expm

import idc
import tsw
from _common import Grid, OracleIdcOps, orc_gmres, random_ordinates

KINDS = {"lobatto": tsw.IDC_LOBATTO, "legendre": tsw.IDC_LEGENDRE, "radau": tsw.IDC_RADAU,
         "chebyshev": tsw.IDC_CHEBYSHEV, "chebyshev-radau": tsw.IDC_CHEBYSHEV_RADAU}


def test_quadrature_tables():
    nodes, W, left, right = idc.idc_tables(4, "lobatto")
    assert left and right and nodes[0] == 0.0 and nodes[-1] == 1.0
    np.testing.assert_allclose(nodes, [0, (1 - 1 / math.sqrt(5)) / 2, (1 + 1 / math.sqrt(5)) / 2, 1], atol=1e-15)
    for deg in range(4):
        f = nodes ** deg
        for n in range(1, 4):
            exact = (nodes[n] ** (deg + 1) - nodes[n - 1] ** (deg + 1)) / (deg + 1)
            assert abs(W[n] @ f - exact) < 1e-14
        assert abs(W[0] @ f - 1.0 / (deg + 1)) < 1e-14
    np.testing.assert_allclose(W[1:].sum(axis=0), W[0], atol=1e-15)


def test_radau_and_chebyshev_nodes():
    r6 = math.sqrt(6.0)
    np.testing.assert_allclose(idc.radau_nodes(2), [1 / 3, 1.0], atol=1e-15)
    np.testing.assert_allclose(idc.radau_nodes(3), [(4 - r6) / 10, (4 + r6) / 10, 1.0], atol=1e-15)
    for m in range(1, 13):
        t = idc.radau_nodes(m)
        assert t[-1] == 1.0 and np.all(np.diff(t) > 0)
        w = idc.integration_weights(t, 0.0, 1.0)
        for deg in range(2 * m - 1):   # Radau with m points: exact to degree 2m - 2
            assert abs(w @ t ** deg - 1.0 / (deg + 1)) < 1e-13, (m, deg)
        if 2 <= m <= 6:
            assert abs(w @ t ** (2 * m - 1) - 1.0 / (2 * m)) > 1e-7   # and no further
    m = 5
    np.testing.assert_allclose(idc.chebyshev_nodes(m),
                               np.sort(0.5 * (np.cos((2 * np.arange(m) + 1) * np.pi / (2 * m)) + 1)), atol=1e-15)
    cr = idc.chebyshev_radau_nodes(m)
    assert cr[-1] == 1.0
    np.testing.assert_allclose(cr, np.sort(0.5 * (np.cos(2 * np.arange(m) * np.pi / (2 * m - 1)) + 1)), atol=1e-15)
    for kind, (_, left, right) in idc.NODE_KINDS.items():
        stages, W, l, r = idc.idc_tables(4, kind)
        assert (l, r) == (left, right)
        assert stages[0] == 0.0 and len(stages) == 4 + (0 if left else 1)
        assert (stages[-1] == 1.0) == right
        if right:   # the subintervals then tile [0, 1]
            np.testing.assert_allclose(W[1:].sum(axis=0), W[0], atol=1e-14)


@pytest.mark.parametrize("kind", list(KINDS))
def test_library_idc_tables_match(kind):
    """tsw_idc_nodes (host-only, quadruple-precision roots) against the oracle's tables."""
    for m in range(2 if kind == "lobatto" else 1, 13):
        stages, W, _, _ = idc.idc_tables(m, kind)
        ls, lW = tsw.idc_nodes(KINDS[kind], m)
        np.testing.assert_allclose(ls, stages, rtol=0, atol=4e-16 * m)
        np.testing.assert_allclose(lW, W, rtol=0, atol=2e-14 * m)
    with pytest.raises(tsw.TswError):
        tsw.idc_nodes(KINDS[kind], 13)
    with pytest.raises(tsw.TswError):
        tsw.idc_nodes(7, 3)


def _problem():
    rng = np.random.default_rng(3)
    g = Grid(3, 3, 1)
    o = random_ordinates(1, rng)
    st = rng.uniform(0.5, 2.0, size=(5, 5))
    ss = st * rng.uniform(0.2, 0.7, size=st.shape)
    src = np.zeros(g.psi_shape(o.nq))
    src[1:-1, 1:-1] = rng.uniform(0.0, 1.0, size=src[1:-1, 1:-1].shape)
    # vacuum inflow: the reference's step keeps the initial condition's ghost cells at zero
    # (idc.cpp:833-835), so its node-0 residual has no inflow term and the step is consistent only
    # with zero inflow, the case its time-dependent problems use
    return g, o, st, ss, src


def _exact(g, o, ops, psi0, src, T):
    shape = g.psi_shape(o.nq)
    interior = np.zeros(shape, dtype=bool)
    interior[1:-1, 1:-1] = True
    idx = np.flatnonzero(interior)
    n = idx.size
    B = np.zeros((n, n))
    for k, i in enumerate(idx):
        e = np.zeros(shape)
        e.flat[i] = 1.0
        B[:, k] = (ops.lmv(e) - ops.k(e)).flat[idx]
    halo_only = src.copy()
    halo_only[1:-1, 1:-1] = 0.0
    q_eff = src.flat[idx] - ops.lmv(halo_only).flat[idx]
    x0 = psi0.flat[idx]
    xs = np.linalg.solve(B, q_eff)
    xT = xs + expm(-B * T) @ (x0 - xs)
    out = np.zeros(shape)
    out.flat[idx] = xT
    return out


@pytest.mark.parametrize("num_corr,order,steps", [(3, 4, (8, 16)), (1, 2, (16, 32)), (0, 1, (16, 32))])
def test_idc_order_of_accuracy(num_corr, order, steps):
    """Each correction raises the order by one (measured rates 1.0, 2.0, 4.1 at these step counts)."""
    g, o, st, ss, src = _problem()
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    psi0 = np.zeros(g.psi_shape(o.nq))
    psi0[1:-1, 1:-1] = 0.5
    T = 0.8
    ref = _exact(g, o, ops, psi0, src, T)
    errs = []
    for nsteps in steps:
        psi = psi0.copy()
        for _ in range(nsteps):
            psi = idc.idc_step(ops, psi, src, T / nsteps, num_nodes=4, num_corr=num_corr)
        errs.append(np.abs(psi - ref).max())
    rate = math.log2(errs[0] / errs[1])
    assert rate >= order - 0.25, (errs, rate)


def test_idc_collocation_legendre_nodes_converge():
    g, o, st, ss, src = _problem()
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    psi0 = np.zeros(g.psi_shape(o.nq))
    psi0[1:-1, 1:-1] = 0.5
    ref = _exact(g, o, ops, psi0, src, 0.5)
    errs = []
    for nsteps in (10, 20):
        psi = psi0.copy()
        for _ in range(nsteps):
            psi = idc.idc_step(ops, psi, src, 0.5 / nsteps, num_nodes=3, num_corr=3, kind="legendre")
        errs.append(np.abs(psi - ref).max())
    assert errs[1] < 2e-7 and math.log2(errs[0] / errs[1]) >= 4.0, errs   # measured 7.6e-8, rate 4.56


@pytest.mark.parametrize("kind,num_nodes,num_corr,colloc", [("lobatto", 4, 3, False), ("radau", 3, 2, False),
                                                             ("legendre", 3, 2, True), ("chebyshev-radau", 3, 1, True)])
def test_update_form_equals_error_form(kind, num_nodes, num_corr, colloc):
    """In exact arithmetic the two forms are one step; with solves to 1e-14 they agree to rounding."""
    g, o, st, ss, src = _problem()
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    rng = np.random.default_rng(9)
    psi0 = np.zeros(g.psi_shape(o.nq))
    psi0[1:-1, 1:-1] = rng.uniform(0.0, 1.0, size=psi0[1:-1, 1:-1].shape)
    a = idc.idc_step(ops, psi0, src, 0.2, num_nodes=num_nodes, num_corr=num_corr, kind=kind, collocation=colloc)
    b = idc.idc_step(ops, psi0, src, 0.2, num_nodes=num_nodes, num_corr=num_corr, kind=kind, collocation=colloc,
                     form="update")
    assert np.abs(a - b).max() <= 1e-12 * np.abs(a).max()
    # the corrections do act: without them the step differs at first order
    c = idc.idc_step(ops, psi0, src, 0.2, num_nodes=num_nodes, num_corr=0, kind=kind, collocation=colloc)
    assert np.abs(a - c).max() > 1e-6 * np.abs(a).max()


@pytest.mark.parametrize("form", ["error", "update"])
def test_radau_order_of_accuracy(form):
    """3 Radau IIA nodes: the rate rises by one per correction, up to the nodes' order.  Stage 0 is no node
    here, so the first subinterval's integrals extrapolate the nodes' interpolant, and the rates approach
    their limits from oscillating values (measured from 8 to 32 steps: 0.93, 1.89, 2.68); the check is
    over that span."""
    g, o, st, ss, src = _problem()
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    psi0 = np.zeros(g.psi_shape(o.nq))
    psi0[1:-1, 1:-1] = 0.5
    T = 0.8
    ref = _exact(g, o, ops, psi0, src, T)
    for num_corr, low in ((0, 0.85), (1, 1.8), (2, 2.5)):
        errs = []
        for nsteps in (8, 32):
            psi = psi0.copy()
            for _ in range(nsteps):
                psi = idc.idc_step(ops, psi, src, T / nsteps, num_nodes=3, num_corr=num_corr, kind="radau", form=form)
            errs.append(np.abs(psi - ref).max())
        rate = math.log2(errs[0] / errs[1]) / 2
        assert rate >= low, (num_corr, errs, rate)


def test_idc_keeps_steady_state():
    g, o, st, ss, src = _problem()
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    rc, psi_ss, _, _, _ = orc_gmres(g, o, st, ss, src, abs_tol=1e-14, restart=40)
    assert rc == 0
    psi0 = psi_ss.copy()
    ops.zero_halo(psi0)
    out = idc.idc_step(ops, psi0, src, 0.3)
    assert np.abs(out - psi0).max() <= 1e-11 * np.abs(psi0).max()
