"""The line-source IDC configuration on the GPU: 200x200 cells on [-1.5, 1.5]^2, Q2, level-symmetric
S16 (144 ordinates), sigma_s = 1, sigma_a = 0, vacuum, a Gaussian pulse (width 0.03) of unit mass, IDC with
4 Gauss-Lobatto nodes and 3 corrections, dt = 0.01, FP32 sweeps with FP64 correction sums.

- Full size, two steps: particle balance (no absorption, and the pulse is far from the boundary, so the
  mass of phi is kept to the solve tolerance), D4 symmetry of phi (the mesh and the S16 set are both
  symmetric under x -> -x, y -> -y and x <-> y), FP32 sweeps against FP64 sweeps.
- Reduced mesh (20x20, same degree, ordinates, nodes and dt): every step against the oracle's IDC
  restatement (oracle/idc.py over the oracle's GMRES), FP64 and FP32 sweeps.
- Full size, 20 steps: the adjoint of the ring response J = int_{0.8 <= r <= 1} phi(t) with respect to the
  initial condition and sigma_s, with checkpoints every 10 steps: the dot-product identity (J is affine in
  psi0) and independence of the checkpoint interval.
Reference: IDCIntegrator::Step_Nonhybrid (src/time_integrators/RKDG/idc.cpp:846-940)."""
import math

import numpy as np
import pytest

import idc
import tsw
from _common import Grid, OracleIdcOps, Ordinates, rel_err
from tsw import problems

pytestmark = pytest.mark.gpu

DT = 0.01


def _ctx(P):
    c = tsw.Context(0)
    P.set_on(c)
    c.idc_set(4, 3, tsw.IDC_LOBATTO, False)
    return c


def _phi(psi, w):
    """Scalar flux sum_q w_q psi_q (torch, psi layout -> phi layout)."""
    import torch
    return torch.einsum("ijqde,q->ijde", psi, w)


def _mass(phi, h):
    return float(phi[1:-1, 1:-1, 0, 0].sum()) * h * h


def _d4_defect(phi):
    """Largest relative defect of phi under the mirror images x -> -x, y -> -y and the transpose x <-> y
    (coefficient (d, e) changes sign by (-1)^d under x -> -x and becomes (e, d) under the transpose)."""
    import torch
    P = phi.shape[2]
    s = torch.tensor([(-1.0) ** d for d in range(P)], dtype=phi.dtype, device=phi.device)
    x = phi[1:-1, 1:-1]
    images = [torch.flip(x, [0]) * s[:, None], torch.flip(x, [1]) * s[None, :], x.permute(1, 0, 3, 2)]
    scale = float(x.norm())
    return max(float((x - im).norm()) / scale for im in images)


def test_line_source_full_size_balance_symmetry_precision():
    import torch
    dev = torch.device("cuda:0")
    P, psi0_h = problems.line_source()
    h = (P.bx - P.ax) / P.nx
    psi0 = torch.from_numpy(psi0_h).to(dev)
    del psi0_h
    src = torch.zeros_like(psi0)
    st = torch.from_numpy(P.sigma_t).to(dev)
    ss = torch.from_numpy(P.sigma_s).to(dev)
    w = torch.from_numpy(P.w).to(dev)
    m0 = _mass(_phi(psi0, w), h)
    assert abs(m0 - 1.0) <= 1e-6   # the projected Gaussian has unit mass
    out = {}
    with _ctx(P) as c:
        for prec in (tsw.FP64, tsw.FP32):
            psi = psi0.clone()
            print(f"line source 200x200 Q2 S16, 2 IDC steps, precision {prec}", flush=True)
            c.idc_run(psi, src, st, ss, DT, 2, abs_tol=1e-10, precision=prec)
            out[prec] = _phi(psi, w)
            del psi
    for prec, sym_tol in ((tsw.FP64, 1e-12), (tsw.FP32, 1e-6)):
        phi = out[prec]
        assert abs(_mass(phi, h) - m0) <= 1e-8 * m0, (prec, _mass(phi, h), m0)
        assert _d4_defect(phi) <= sym_tol, prec
    inner = (slice(1, -1), slice(1, -1))
    err = float((out[tsw.FP32][inner] - out[tsw.FP64][inner]).norm() / out[tsw.FP64][inner].norm())
    assert err <= 1e-6, err


def test_line_source_reduced_mesh_matches_oracle():
    P, psi0 = problems.line_source(n=20)
    g = Grid(P.nx, P.ny, P.p, P.ax, P.bx, P.ay, P.by)
    o = Ordinates(P.xi, P.eta, P.w, P.quad_off)
    src = np.zeros_like(psi0)
    ops = OracleIdcOps(g, o, P.sigma_t, P.sigma_s, tol=1e-14)
    refs = [psi0]
    for _ in range(2):
        refs.append(idc.idc_step(ops, refs[-1], src, DT, num_nodes=4, num_corr=3, kind="lobatto"))
    with _ctx(P) as c:
        for prec, tol in ((tsw.FP64, 1e-10), (tsw.FP32, 1e-8)):
            psi = psi0.copy()
            for k in (1, 2):
                c.idc_step(psi, src, P.sigma_t, P.sigma_s, DT, abs_tol=1e-13 if prec == tsw.FP64 else 1e-12,
                           precision=prec)
                assert rel_err(psi, refs[k]) <= tol, (prec, k)


def test_line_source_ring_adjoint_checkpoints():
    import torch
    dev = torch.device("cuda:0")
    P, psi0_h = problems.line_source()
    seed = torch.from_numpy(problems.ring_seed(P)).to(dev)
    psi0 = torch.from_numpy(psi0_h).to(dev)
    del psi0_h
    src = torch.zeros_like(psi0)
    st = torch.from_numpy(P.sigma_t).to(dev)
    ss = torch.from_numpy(P.sigma_s).to(dev)
    nsteps = 20
    with _ctx(P) as c:
        x = psi0.clone()
        c.idc_run(x, src, st, ss, DT, nsteps, abs_tol=1e-10, precision=tsw.FP32)
        J0 = float((seed * x).sum())
        assert J0 > 0.0
        grads = {}
        for ck in (10, 20):
            print(f"line source ring adjoint, {nsteps} steps, checkpoint every {ck}", flush=True)
            gp = torch.zeros_like(psi0)
            gss = torch.zeros_like(ss)
            c.idc_adjoint(seed, gp, src, st, ss, DT, nsteps, psi0=psi0, checkpoint_every=ck, grad_sigma_s=gss,
                          abs_tol=1e-10, precision=tsw.FP32)
            grads[ck] = (gp, gss)
        for a, b in zip(grads[10], grads[20]):
            assert float((a - b).norm()) <= 1e-12 * float(b.norm())
        gp, gss = grads[10]
        assert float(gss[1:-1, 1:-1].abs().sum()) > 0.0
        assert float(gp[0].abs().sum()) == 0.0 and float(gp[:, -1].abs().sum()) == 0.0
        # J is affine in psi0: J(psi0 + d) - J(psi0) = <dJ/dpsi0, d>
        gen = torch.Generator(device=dev).manual_seed(5)
        d = torch.zeros_like(psi0)
        d[1:-1, 1:-1] = 1e-3 * torch.rand(d[1:-1, 1:-1].shape, device=dev, dtype=torch.float64, generator=gen)
        x = psi0 + d
        c.idc_run(x, src, st, ss, DT, nsteps, abs_tol=1e-10, precision=tsw.FP32)
        dJ = float((seed * x).sum()) - J0
        pred = float((gp * d).sum())
        assert abs(dJ - pred) <= 1e-6 * abs(dJ), (dJ, pred)
