"""Adjoint (vector-Jacobian) path on the GPU: tsw_sweep_vjp / tsw_solve_vjp through the C-ABI and the
torch.autograd functions over them.

* sweep vjp vs the oracle's directly transposed sweep (interior adjoint flux, inflow gradients, sigma_t
  gradient), every pattern, p 0..3, steady and dt, FP64 to 1e-12 and FP32 to 2e-5 (relative l2);
* solve vjp vs the oracle's adjoint source iteration (1e-9);
* detector response of a 128x128 Q1 S8 scattering problem: gradients in sigma_t, sigma_s and Q against
  central finite differences of the GPU forward solve on 64 seeded cells (to 1e-4 of the largest
  gradient: the difference quotient of solves converged to 1e-13 carries noise near 1e-11);
* torch.autograd.gradcheck of both functions on small problems;
* the bilinear identity <L^{-1}x, g> = <x, L^{-T}g> (+ inflow terms) on the full 280x280 Q2 S12 lattice.
"""
import math

import numpy as np
import pytest

from _common import (Grid, adjoint_sigma_grads, halo_mask, interior, ls_ordinates, orc_gmres, orc_pmv,
                     orc_solve_adjoint, orc_sweep, orc_sweep_adjoint, random_ordinates, random_problem, rel_err)

pytestmark = pytest.mark.gpu

PATTERNS = ["hip-staggered", "hip-staggered-ge", "hip-wavefront"]


def _ctx(g, o, pattern="hip-staggered"):
    import tsw
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern(pattern)
    return c


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("dt", [math.inf, 0.05])
def test_sweep_vjp_matches_oracle(pattern, p, dt):
    import tsw
    o = ls_ordinates(6)
    g, sigma_t, _, src = random_problem(9, 7, p, o, seed=300 + p, ax=-0.5, bx=1.0, materials=3)
    rng = np.random.default_rng(17 + p)
    seed = rng.normal(size=g.psi_shape(o.nq))
    psi_ref = orc_sweep(g, o, sigma_t, src, dt)
    adj_ref = orc_sweep_adjoint(g, o, sigma_t, seed, dt)
    gst_ref = adjoint_sigma_grads(psi_ref, adj_ref)
    h = halo_mask(seed.shape)
    gsrc_ref = adj_ref.copy()
    gsrc_ref[h] += seed[h]
    with _ctx(g, o, pattern) as c:
        for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-5)):
            gsrc = np.zeros_like(seed)
            gst = np.full(sigma_t.shape, np.nan)
            c.sweep_vjp(seed, gsrc, sigma_t, psi=psi_ref, grad_sigma_t=gst, dt=dt, precision=prec)
            assert rel_err(interior(gsrc), interior(gsrc_ref)) <= tol
            assert rel_err(gsrc[h], gsrc_ref[h]) <= tol
            assert rel_err(gst, gst_ref) <= tol
            assert not gst[0].any() and not gst[-1].any() and not gst[:, 0].any() and not gst[:, -1].any()


def test_sweep_vjp_random_ordinates_no_materials_and_aliasing():
    """Unstructured ordinate set, per-cell cross sections (elimination path), grad_src aliasing seed."""
    import tsw
    rng = np.random.default_rng(5)
    o = random_ordinates(3, rng)
    g, sigma_t, _, src = random_problem(11, 6, 2, o, seed=8)
    seed = rng.normal(size=g.psi_shape(o.nq))
    adj_ref = orc_sweep_adjoint(g, o, sigma_t, seed)
    with _ctx(g, o) as c:
        buf = seed.copy()
        c.sweep_vjp(buf, buf, sigma_t)
    h = halo_mask(seed.shape)
    assert rel_err(interior(buf), interior(adj_ref)) <= 1e-12
    assert rel_err(buf[h], seed[h] + adj_ref[h]) <= 1e-12


@pytest.mark.parametrize("pattern", ["hip-staggered", "hip-wavefront"])
def test_solve_vjp_matches_oracle_adjoint(pattern):
    import tsw
    rng = np.random.default_rng(21)
    o = ls_ordinates(4)
    g, sigma_t, sigma_s, src = random_problem(8, 10, 1, o, seed=22, materials=4)
    seed = rng.normal(size=g.psi_shape(o.nq))
    rc, psi, phi, _, _ = orc_gmres(g, o, sigma_t, sigma_s, src, abs_tol=1e-13, restart=40)
    assert rc == 0
    zeta = orc_solve_adjoint(g, o, sigma_t, sigma_s, seed)
    gt_ref, gs_ref = adjoint_sigma_grads(psi, zeta, orc_pmv(g, o, psi))
    h = halo_mask(seed.shape)
    gsrc_ref = zeta.copy()
    gsrc_ref[h] += seed[h]
    with _ctx(g, o, pattern) as c:
        for prec, tol in ((tsw.FP64, 1e-9), (tsw.FP32, 1e-7)):
            gsrc = np.zeros_like(seed)
            gst, gss = np.zeros(sigma_t.shape), np.zeros(sigma_t.shape)
            its, res = c.solve_vjp(seed, gsrc, sigma_t, sigma_s, psi=psi, grad_sigma_t=gst, grad_sigma_s=gss,
                                   abs_tol=1e-12, restart=30, precision=prec)
            assert res <= 1e-12 and its > 0
            assert rel_err(gsrc, gsrc_ref) <= tol
            assert rel_err(gst, gt_ref) <= tol
            assert rel_err(gss, gs_ref) <= tol


# ------------------------------------------------------------------------------------------------------
# Detector response: 128x128, Q1, S8, per-cell sigma_t log-uniform [0.5, 5], scattering ratio c ~ U[0.2, 0.9],
# isotropic q = 1 on x < 0.1 (exact projection of the partly covered column), vacuum inflow,
# J = integral of phi over the detector [0.85, 0.95] x [0.45, 0.55] (exact integral over partial cells).
# ------------------------------------------------------------------------------------------------------
def _leg_antider(d, t):
    """Antiderivative of P_d at t (P_{d+1} - P_{d-1})/(2d+1), t for d = 0."""
    from numpy.polynomial import legendre as L
    if d == 0:
        return t
    c1 = np.zeros(d + 2)
    c1[d + 1] = 1.0
    c0 = np.zeros(d + 2)
    c0[d - 1] = 1.0
    return (L.legval(t, c1) - L.legval(t, c0)) / (2 * d + 1)


def _interval_moments(n, a, b, lo, hi, p):
    """m[k, d] = integral over [lo, hi] of P_d(local coordinate of cell k) dx, k = 1..n (row 0 and n+1 zero)."""
    h = (b - a) / n
    m = np.zeros((n + 2, p + 1))
    for k in range(1, n + 1):
        x0, x1 = a + (k - 1) * h, a + k * h
        u0, u1 = max(x0, lo), min(x1, hi)
        if u1 <= u0:
            continue
        t0, t1 = 2 * (u0 - x0) / h - 1, 2 * (u1 - x0) / h - 1
        for d in range(p + 1):
            m[k, d] = h / 2 * (_leg_antider(d, t1) - _leg_antider(d, t0))
    return m


def detector_problem(n=128, p=1, order=8, seed=4):
    rng = np.random.default_rng(seed)
    o = ls_ordinates(order)
    g = Grid(n, n, p)
    sigma_t = np.exp(rng.uniform(math.log(0.5), math.log(5.0), size=(n + 2, n + 2)))
    c = rng.uniform(0.2, 0.9, size=(n + 2, n + 2))
    sigma_s = c * sigma_t
    # source coefficients: (2d+1)(2e+1)/(dx dy) * integral of q P_d P_e over the covered part of the cell
    mx = _interval_moments(n, 0.0, 1.0, 0.0, 0.1, p)
    my = _interval_moments(n, 0.0, 1.0, 0.0, 1.0, p)
    w2 = np.array([2 * d + 1 for d in range(p + 1)], float)
    qcell = np.einsum("id,je,d,e->ijde", mx, my, w2, w2) / (g.dx * g.dy)
    src = np.zeros(g.psi_shape(o.nq))
    src[:] = (qcell / (4 * math.pi))[:, :, None]
    src[halo_mask(src.shape)] = 0.0
    # detector seed: J = sum_q w_q sum_{c,d,e} psi_{q,c,de} Ix_d Iy_e
    ix = _interval_moments(n, 0.0, 1.0, 0.85, 0.95, p)
    iy = _interval_moments(n, 0.0, 1.0, 0.45, 0.55, p)
    det = np.einsum("id,je->ijde", ix, iy)
    seed = o.w[None, None, :, None, None] * det[:, :, None]
    return g, o, sigma_t, sigma_s, src, seed


def test_detector_gradients_match_finite_differences():
    import tsw
    g, o, sigma_t, sigma_s, src, seed = detector_problem()
    tol_fwd = 1e-13
    with _ctx(g, o) as c:
        def solve(st, ss, q):
            psi = q.copy()
            c.gmres_solve(psi, st, ss, abs_tol=tol_fwd, restart=30)
            return psi

        def J(st, ss, q):
            return float(np.sum(solve(st, ss, q) * seed))

        psi = solve(sigma_t, sigma_s, src)
        gsrc = np.zeros_like(src)
        gst, gss = np.zeros(sigma_t.shape), np.zeros(sigma_t.shape)
        c.solve_vjp(seed, gsrc, sigma_t, sigma_s, psi=psi, grad_sigma_t=gst, grad_sigma_s=gss, abs_tol=1e-13)
        rng = np.random.default_rng(64)
        cells = [(int(i), int(j)) for i, j in zip(rng.integers(1, g.nx + 1, 64), rng.integers(1, g.ny + 1, 64))]
        # half of the cells near the detector, where the response is largest
        cells[32:] = [(int(i), int(j)) for i, j in zip(rng.integers(100, 125, 32), rng.integers(50, 80, 32))]
        scale = max(np.abs(gst).max(), np.abs(gss).max())
        for k, (i, j) in enumerate(cells):
            h = 1e-4 * sigma_t[i, j]
            sp, sm = sigma_t.copy(), sigma_t.copy()
            sp[i, j] += h
            sm[i, j] -= h
            fd = (J(sp, sigma_s, src) - J(sm, sigma_s, src)) / (2 * h)
            assert abs(fd - gst[i, j]) <= 1e-4 * (scale + abs(fd)), ("sigma_t", i, j, fd, gst[i, j])
            if k % 4 == 0:
                sp, sm = sigma_s.copy(), sigma_s.copy()
                sp[i, j] += h
                sm[i, j] -= h
                fd = (J(sigma_t, sp, src) - J(sigma_t, sm, src)) / (2 * h)
                assert abs(fd - gss[i, j]) <= 1e-4 * (scale + abs(fd)), ("sigma_s", i, j, fd, gss[i, j])
        # J is linear in the source: a unit perturbation of one coefficient gives the gradient entry
        base = J(sigma_t, sigma_s, src)
        for (i, j) in cells[::8]:
            for q in (0, o.nq // 2, o.nq - 1):
                qp = src.copy()
                qp[i, j, q, 1, 0] += 1.0
                fd = J(sigma_t, sigma_s, qp) - base
                assert abs(fd - gsrc[i, j, q, 1, 0]) <= 1e-6 * (np.abs(gsrc).max() + abs(fd))


def test_autograd_gradcheck():
    import torch
    import tsw
    from tsw.autograd import solve, sweep
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(3)
    o = ls_ordinates(2)
    g, sigma_t, sigma_s, src = random_problem(3, 4, 1, o, seed=12)
    with _ctx(g, o) as c:
        s = torch.tensor(src, device=dev, requires_grad=True)
        st = torch.tensor(sigma_t, device=dev, requires_grad=True)
        ss = torch.tensor(sigma_s, device=dev, requires_grad=True)
        assert torch.autograd.gradcheck(lambda a, b: sweep(a, b, c, 0.2), (s, st), eps=1e-6, atol=1e-7, rtol=1e-6)
        assert torch.autograd.gradcheck(lambda a, b, d: solve(a, b, d, c, abs_tol=1e-13, adj_tol=1e-13),
                                        (s, st, ss), eps=1e-6, atol=1e-6, rtol=1e-5)
        # a scalar loss through the sweep: gradient equals the oracle's transposed sweep of dloss/dpsi
        psi = sweep(s, st, c)
        loss = (psi[1:-1, 1:-1] ** 2).sum()
        loss.backward()
        psi_np = psi.detach().cpu().numpy()
        seed = np.zeros_like(psi_np)
        seed[1:-1, 1:-1] = 2 * psi_np[1:-1, 1:-1]
        adj = orc_sweep_adjoint(g, o, sigma_t, seed)
        assert rel_err(interior(s.grad.cpu().numpy()), interior(adj)) <= 1e-12


def test_bilinear_identity_full_size_lattice():
    """<L^{-1}x, g> = <x_int, zeta> + <x_halo, dJ/dinflow> on the 280x280 Q2 S12 lattice (59.3 M unknowns)."""
    import torch
    import tsw
    from tsw import problems
    P = problems.lattice()
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(11)
    shape = (P.nx + 2, P.ny + 2, P.nq, P.p + 1, P.p + 1)
    x = torch.rand(shape, generator=gen, device=dev, dtype=torch.float64)
    seed = torch.randn(shape, generator=gen, device=dev, dtype=torch.float64)
    st = torch.from_numpy(P.sigma_t).to(dev)
    with tsw.Context(0) as c:
        c.problem_set(P.nx, P.ny, P.p, P.xi, P.eta, P.w, P.quad_off, P.ax, P.bx, P.ay, P.by)
        y = x.clone()
        c.sweep(y, st)
        gsrc = torch.empty_like(seed)
        zero_halo_seed = seed.clone()
        zero_halo_seed[0] = 0
        zero_halo_seed[-1] = 0
        zero_halo_seed[:, 0] = 0
        zero_halo_seed[:, -1] = 0
        c.sweep_vjp(zero_halo_seed, gsrc, st)
    lhs = float((y[1:-1, 1:-1] * seed[1:-1, 1:-1]).sum())
    rhs = float((x * gsrc).sum())
    mag = float((y[1:-1, 1:-1] * seed[1:-1, 1:-1]).abs().sum())
    assert abs(lhs - rhs) <= 1e-12 * mag
