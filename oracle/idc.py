"""idc.py -- CPU restatement of the reference's non-hybrid IDC time step (error and update forms) over the
oracle's transport operators and Peierls GMRES (crockatt/idc-transport-solver,
src/time_integrators/RKDG/idc.cpp).

TEST INFRASTRUCTURE ONLY: the checker for the library's tsw_idc_step; only tests/ load it.

Step (IDCIntegrator::Step_Nonhybrid, idc.cpp:846-940), nodes t_0 < ... < t_{S-1} on [0, 1]:
  predictor  (1/(h dt) + L - S P) psi_n = psi_{n-1}/(h dt) + Q            (:863-876), h = t_n - t_{n-1}
  corrections, K times:
    R_l = (L - S P) psi_l for every node                                  (ComputeResiduals, :1658-1667)
    for n = 1..S-1, error form (IDCType::Error):
      (1/(h dt) + L - S P) e = (psi_n_new_{n-1} - psi_n)/(h dt) - sum_l W(n,l)/h R_l + Q, zero halos
                                                                          (ComputeSource, :1686-1709)
      psi_n += e                                                          (:905-907)
    or update form (IDCType::Update):
      (1/(h dt) + L - S P) psi_n = psi_n_new_{n-1}/(h dt) + sum_l (delta_nl - W(n,l)/h) R_l + Q,
      halos = Q's                                                         (:1711-1730, :909-911)
  result psi_{S-1} when the nodes hold the right end of the step and no collocation is asked,
  else the collocation formula                                            (:933-939, :1541-1566)
W(0,l) integrates the Lagrange basis of the nodes over [0,1], W(n,l) over [t_{n-1}, t_n]
(SetNodes, idc.cpp:2305-2376).  Stage 0 is the initial condition with zero halos (t = 0 is a stage
even where it is no quadrature node); the other stages carry the source's halos (the boundary
condition).  Node families (Quadrule::ComputeQuadratureNodes, Quadrule.cpp:1077-1145): Gauss-Lobatto,
Gauss-Legendre, Gauss-Radau IIA (roots of P_n - P_{n-1}, right end included: Quadrule::Radau :412-446),
Chebyshev cos((2i+1) pi / 2n) and Chebyshev-Radau cos(2 i pi / (2n - 1)).
"""
from __future__ import annotations

import numpy as np
from numpy.polynomial import legendre as npleg


def lobatto_nodes(m):
    """Gauss-Lobatto nodes on [0, 1]: the endpoints and the roots of P'_{m-1}."""
    c = np.zeros(m)
    c[m - 1] = 1.0
    inner = npleg.legroots(npleg.legder(c)) if m > 2 else np.array([])
    x = np.concatenate([[-1.0], np.sort(inner), [1.0]])
    return 0.5 * (x + 1.0)


def legendre_nodes(m):
    x, _ = npleg.leggauss(m)
    return 0.5 * (np.sort(x) + 1.0)


def radau_nodes(m):
    """Gauss-Radau IIA nodes on [0, 1]: the roots of P_m - P_{m-1} (their largest is 1), ascending, each
    polished by Newton steps on that polynomial."""
    c = np.zeros(m + 1)
    c[m] = 1.0
    c[m - 1] = -1.0
    x = np.sort(np.real(npleg.legroots(c)))
    d = npleg.legder(c)
    for _ in range(3):
        x = x - npleg.legval(x, c) / npleg.legval(x, d)
    x[-1] = 1.0
    return 0.5 * (x + 1.0)


def chebyshev_nodes(m):
    """cos((2i+1) pi / (2m)), i = 0..m-1, ascending, on [0, 1]."""
    i = np.arange(m)
    x = np.cos((2 * i + 1) * np.pi / (2 * m))[::-1]
    return 0.5 * (x + 1.0)


def chebyshev_radau_nodes(m):
    """cos(2 i pi / (2m - 1)), i = 0..m-1, ascending (the last is 1), on [0, 1]."""
    i = np.arange(m)
    x = np.cos(2 * i * np.pi / (2 * m - 1))[::-1]
    return 0.5 * (x + 1.0)


# node families: generator, holds the left end, holds the right end
NODE_KINDS = {
    "lobatto": (lobatto_nodes, True, True),
    "legendre": (legendre_nodes, False, False),
    "radau": (radau_nodes, False, True),
    "chebyshev": (chebyshev_nodes, False, False),
    "chebyshev-radau": (chebyshev_radau_nodes, False, True),
}


def integration_weights(nodes, a, b):
    """w_l = integral over [a, b] of the Lagrange basis polynomial l of `nodes`."""
    m = len(nodes)
    xg, wg = npleg.leggauss(m + 2)
    t = 0.5 * (b - a) * (xg + 1.0) + a
    w = np.zeros(m)
    for l in range(m):
        lag = np.ones_like(t)
        for k in range(m):
            if k != l:
                lag *= (t - nodes[k]) / (nodes[l] - nodes[k])
        w[l] = 0.5 * (b - a) * np.sum(wg * lag)
    return w


def idc_tables(num_nodes, kind="lobatto"):
    """(stage nodes, W, left, right) with W[0] over [0,1] and W[n] over [t_{n-1}, t_n]; left endpoint added
    as stage 0 for nodes that exclude it (the reference's num_stages = num_nodes + !left_endpoint)."""
    if kind not in NODE_KINDS:
        raise ValueError(kind)
    make, left, right = NODE_KINDS[kind]
    nodes = make(num_nodes)
    stages = nodes if left else np.concatenate([[0.0], nodes])
    S = len(stages)
    W = np.zeros((S, S))
    off = 0 if left else 1
    W[0, off:] = integration_weights(nodes, 0.0, 1.0)
    for n in range(1, S):
        W[n, off:] = integration_weights(nodes, stages[n - 1], stages[n])
    return stages, W, left, right


def idc_step(ops, psi0, source, dt, num_nodes=4, num_corr=3, kind="lobatto", collocation=False, form="error"):
    """One IDC step.  ops: object with
         solve(rhs, dt_sub) -> psi   implicit transport solve (1/dt_sub + L - S P) psi = rhs (halos = inflow)
         lmv(x) -> y                 L x on the interior (inflow from x's halos), zero halos
         k(x) -> y                   S P x on the interior (sigma_s phi / 4 pi broadcast), zero halos
         zero_halo(x)                in place
    psi0: initial condition (zero halos); source: Q with inflow halos, constant over the step.
    form: "error" or "update" (IDCType)."""
    if form not in ("error", "update"):
        raise ValueError(form)
    nodes, W, left, right = idc_tables(num_nodes, kind)
    S = len(nodes)
    stages = [psi0.copy()]
    for n in range(1, S):
        h = nodes[n] - nodes[n - 1]
        rhs = stages[n - 1] / (h * dt)
        ops.zero_halo(rhs)
        rhs = rhs + source
        stages.append(ops.solve(rhs, h * dt))
    l0 = 0 if left else 1
    for _ in range(num_corr):
        R = [None] * S
        for l in range(l0, S):
            R[l] = ops.lmv(stages[l]) - ops.k(stages[l])
        for n in range(1, S):
            h = nodes[n] - nodes[n - 1]
            if form == "error":
                e = -stages[n] / (h * dt)
                for l in range(l0, S):
                    e = e - (W[n, l] / h) * R[l]
                e = e + stages[n - 1] / (h * dt)
                e = e + source
                ops.zero_halo(e)
                stages[n] = stages[n] + ops.solve(e, h * dt)
            else:
                u = np.zeros_like(stages[n])
                for l in range(l0, S):
                    u = u + ((1.0 if n == l else 0.0) - W[n, l] / h) * R[l]
                u = u + stages[n - 1] / (h * dt)
                ops.zero_halo(u)
                u = u + source
                stages[n] = ops.solve(u, h * dt)
    if (not right) or collocation:
        # collocation formula psi0 + dt Q + dt sum_l W(0,l) (S P - L) psi_l (ComputeCollocation_Nonhybrid)
        out = psi0 + dt * source
        for l in range(l0, S):
            out = out + dt * W[0, l] * (ops.k(stages[l]) - ops.lmv(stages[l]))
    else:
        out = stages[S - 1].copy()
    ops.zero_halo(out)
    return out
