"""IDC time steps on the GPU (tsw_idc_step) against the oracle's restatement of the reference's
non-hybrid IDC (oracle/idc.py, pinned by order-of-accuracy tests): several steps with Lobatto nodes
(FP64 to 1e-10 relative, FP32 sweeps with FP64 Krylov vectors and correction sums to 1e-8), Legendre
nodes with collocation, Radau, Chebyshev and Chebyshev-Radau nodes in the error and update forms,
sources with and without inflow halos, and the largest node counts (12 Lobatto or Legendre nodes, whose
correction sources combine 15 arrays).  The adjoint of
several steps is pinned by the dot-product identity and by central differences of the cross-section
gradients, with and without inflow halos.  The line-source configuration itself is in
test_gpu_line_source.py."""
import math

import numpy as np
import pytest

import idc
import tsw
from _common import Grid, OracleIdcOps, ls_ordinates, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _vacuum_problem(nx, ny, p, order, seed, halo_inflow=False):
    o = ls_ordinates(order)
    g, st, ss, src = random_problem(nx, ny, p, o, seed=seed, halo_inflow=halo_inflow, materials=4)
    rng = np.random.default_rng(seed + 1)
    psi0 = np.zeros(g.psi_shape(o.nq))
    psi0[1:-1, 1:-1] = rng.uniform(0.0, 1.0, size=psi0[1:-1, 1:-1].shape)
    return g, o, st, ss, src, psi0


@pytest.mark.parametrize("kind,num_nodes,num_corr,colloc", [(tsw.IDC_LOBATTO, 4, 3, False),
                                                             (tsw.IDC_LOBATTO, 3, 1, True),
                                                             (tsw.IDC_LEGENDRE, 3, 2, True)])
@pytest.mark.parametrize("pattern", ["hip-staggered", "hip-wavefront"])
@pytest.mark.parametrize("halo_inflow", [False, True])
def test_idc_steps_match_oracle(kind, num_nodes, num_corr, colloc, pattern, halo_inflow):
    g, o, st, ss, src, psi0 = _vacuum_problem(9, 12, 1, 4, seed=71, halo_inflow=halo_inflow)
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    ref = psi0.copy()
    name = "lobatto" if kind == tsw.IDC_LOBATTO else "legendre"
    dt = 0.05
    for _ in range(3):
        ref = idc.idc_step(ops, ref, src, dt, num_nodes=num_nodes, num_corr=num_corr, kind=name,
                           collocation=colloc)
    with tsw.Context(0) as c:
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        c.set_pattern(pattern)
        c.idc_set(num_nodes, num_corr, kind, colloc)
        for prec, tol in ((tsw.FP64, 1e-10), (tsw.FP32, 1e-8)):
            psi = psi0.copy()
            for _ in range(3):
                c.idc_step(psi, src, st, ss, dt, abs_tol=1e-13 if prec == tsw.FP64 else 1e-11, precision=prec)
            assert rel_err(psi, ref) <= tol, prec
            assert not psi[0].any() and not psi[:, -1].any()


NODE_NAMES = {tsw.IDC_LOBATTO: "lobatto", tsw.IDC_LEGENDRE: "legendre", tsw.IDC_RADAU: "radau",
              tsw.IDC_CHEBYSHEV: "chebyshev", tsw.IDC_CHEBYSHEV_RADAU: "chebyshev-radau"}


@pytest.mark.parametrize("kind,num_nodes,num_corr,colloc", [(tsw.IDC_RADAU, 3, 2, False),
                                                             (tsw.IDC_RADAU, 2, 1, True),
                                                             (tsw.IDC_CHEBYSHEV, 3, 2, False),
                                                             (tsw.IDC_CHEBYSHEV_RADAU, 4, 3, False),
                                                             (tsw.IDC_LOBATTO, 4, 3, False)])
@pytest.mark.parametrize("form", [tsw.IDC_ERROR, tsw.IDC_UPDATE])
def test_idc_node_families_and_forms_match_oracle(kind, num_nodes, num_corr, colloc, form):
    """The reference deck's configuration (idc_type update, idc_nodes_type gauss-radau) and the other node
    families, in both forms, with inflow halos, against the oracle's step of the same form."""
    g, o, st, ss, src, psi0 = _vacuum_problem(7, 9, 2, 4, seed=29, halo_inflow=True)
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    name = NODE_NAMES[kind]
    fname = "update" if form == tsw.IDC_UPDATE else "error"
    dt = 0.05
    ref = psi0.copy()
    for _ in range(2):
        ref = idc.idc_step(ops, ref, src, dt, num_nodes=num_nodes, num_corr=num_corr, kind=name,
                           collocation=colloc, form=fname)
    with tsw.Context(0) as c:
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        c.idc_set(num_nodes, num_corr, kind, colloc, form=form)
        for prec, tol in ((tsw.FP64, 1e-10), (tsw.FP32, 1e-8)):
            psi = psi0.copy()
            for _ in range(2):
                c.idc_step(psi, src, st, ss, dt, abs_tol=1e-13 if prec == tsw.FP64 else 1e-11, precision=prec)
            assert rel_err(psi, ref) <= tol, prec
            assert not psi[0].any() and not psi[:, -1].any()


@pytest.mark.parametrize("kind,colloc", [(tsw.IDC_LOBATTO, False), (tsw.IDC_LOBATTO, True),
                                         (tsw.IDC_LEGENDRE, True)])
def test_idc_twelve_nodes_match_oracle(kind, colloc):
    """The largest node count tsw_idc_set accepts: the correction source combines 3 + 12 arrays."""
    g, o, st, ss, src, psi0 = _vacuum_problem(4, 5, 1, 2, seed=17, halo_inflow=True)
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    name = "lobatto" if kind == tsw.IDC_LOBATTO else "legendre"
    dt = 0.05
    ref = idc.idc_step(ops, psi0.copy(), src, dt, num_nodes=12, num_corr=2, kind=name, collocation=colloc)
    with tsw.Context(0) as c:
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        c.idc_set(12, 2, kind, colloc)
        psi = psi0.copy()
        c.idc_step(psi, src, st, ss, dt, abs_tol=1e-13)
        assert rel_err(psi, ref) <= 1e-10


def test_idc_device_arrays_and_errors():
    import torch
    g, o, st, ss, src, psi0 = _vacuum_problem(6, 5, 2, 2, seed=3)
    dev = torch.device("cuda:0")
    with tsw.Context(0) as c:
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off)
        host = psi0.copy()
        c.idc_step(host, src, st, ss, 0.1)
        d = torch.from_numpy(psi0).to(dev)
        c.idc_step(d, torch.from_numpy(src).to(dev), torch.from_numpy(st).to(dev), torch.from_numpy(ss).to(dev), 0.1)
        assert rel_err(d.cpu().numpy(), host) <= 1e-12
        with pytest.raises(tsw.TswError):
            c.idc_set(1, 3, tsw.IDC_LOBATTO, False)
        with pytest.raises(tsw.TswError):
            c.idc_step(host, src, st, ss, math.inf)


def _ctx_for(g, o, num_nodes=3, num_corr=2):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.idc_set(num_nodes, num_corr, tsw.IDC_LOBATTO, False)
    return c


@pytest.mark.parametrize("halo_inflow", [False, True])
@pytest.mark.parametrize("kind,colloc,form", [(tsw.IDC_LOBATTO, False, tsw.IDC_ERROR),
                                              (tsw.IDC_LEGENDRE, True, tsw.IDC_ERROR),
                                              (tsw.IDC_RADAU, False, tsw.IDC_UPDATE)])
def test_idc_adjoint_dot_product_over_steps(halo_inflow, kind, colloc, form):
    """The IDC map is affine in psi^0: <g, run(x) - run(0)> = <dJ/dpsi^0, x> pins the transposed step
    (adjoint solves, (L - S P)^T residual terms, integration-matrix weights) over several steps; with
    inflow halos the affine part changes and the linear part must not.  The update form's forward runs
    check that the one transposed step serves both forms."""
    g, o, st, ss, src, psi0 = _vacuum_problem(8, 7, 1, 4, seed=90, halo_inflow=halo_inflow)
    rng = np.random.default_rng(91)
    seed = np.zeros(g.psi_shape(o.nq))
    seed[1:-1, 1:-1] = rng.normal(size=seed[1:-1, 1:-1].shape)
    with _ctx_for(g, o) as c:
        c.idc_set(3, 2, kind, colloc, form=form)
        dt, nsteps = 0.07, 4
        a = psi0.copy()
        c.idc_run(a, src, st, ss, dt, nsteps, abs_tol=1e-14)
        b = np.zeros_like(psi0)
        c.idc_run(b, src, st, ss, dt, nsteps, abs_tol=1e-14)
        grad = np.zeros_like(psi0)
        c.idc_adjoint(seed, grad, src, st, ss, dt, nsteps, abs_tol=1e-14)
    lhs = float(np.sum(seed * (a - b)))
    rhs = float(np.sum(grad * psi0))
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs), (lhs, rhs)
    assert not grad[0].any() and not grad[:, -1].any()


@pytest.mark.parametrize("halo_inflow", [False, True])
def test_idc_adjoint_cross_section_gradients_fd_and_checkpoints(halo_inflow):
    g, o, st, ss, src, psi0 = _vacuum_problem(7, 6, 1, 2, seed=33, halo_inflow=halo_inflow)
    rng = np.random.default_rng(34)
    seed = np.zeros(g.psi_shape(o.nq))
    seed[1:-1, 1:-1] = rng.normal(size=seed[1:-1, 1:-1].shape)
    dt, nsteps = 0.1, 3
    with _ctx_for(g, o) as c:
        def J(stv, ssv):
            x = psi0.copy()
            c.idc_run(x, src, stv, ssv, dt, nsteps, abs_tol=1e-14)
            return float(np.sum(seed * x))

        res = {}
        for ck in (1, 2, 5):
            gp = np.zeros_like(psi0)
            gst, gss = np.zeros(st.shape), np.zeros(st.shape)
            c.idc_adjoint(seed, gp, src, st, ss, dt, nsteps, psi0=psi0, checkpoint_every=ck, grad_sigma_t=gst,
                          grad_sigma_s=gss, abs_tol=1e-14)
            res[ck] = (gp, gst, gss)
        for ck in (2, 5):
            for a, b in zip(res[1], res[ck]):
                np.testing.assert_allclose(a, b, rtol=1e-12, atol=1e-15 * np.abs(b).max())
        gp0 = np.zeros_like(psi0)
        c.idc_adjoint(seed, gp0, src, st, ss, dt, nsteps, abs_tol=1e-14)   # no cross sections: no states
        np.testing.assert_allclose(gp0, res[1][0], rtol=1e-10, atol=1e-14 * np.abs(gp0).max())
        gp, gst, gss = res[1]
        scale = max(np.abs(gst).max(), np.abs(gss).max())
        for (i, j) in [(1, 1), (3, 4), (7, 6), (5, 2)]:
            h = 1e-5 * st[i, j]
            sp, sm = st.copy(), st.copy()
            sp[i, j] += h
            sm[i, j] -= h
            fd = (J(sp, ss) - J(sm, ss)) / (2 * h)
            assert abs(fd - gst[i, j]) <= 1e-6 * scale, ("sigma_t", i, j, fd, gst[i, j])
            sp, sm = ss.copy(), ss.copy()
            sp[i, j] += h
            sm[i, j] -= h
            fd = (J(st, sp) - J(st, sm)) / (2 * h)
            assert abs(fd - gss[i, j]) <= 1e-6 * scale, ("sigma_s", i, j, fd, gss[i, j])
