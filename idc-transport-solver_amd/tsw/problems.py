"""Problem generators producing the reference-layout inputs of a sweep or solve: cross sections,
angular source with inflow halos, grid and ordinate set.

The reference defines problems as classes evaluated at cell centres and projected into the DG basis
(src/utils/init.cpp:422-946, src/utils/Problems/*).  All problems here have cross sections and
sources constant per cell, so the projection is exact: sigma is the cell value (DG degree 0,
init.cpp:741) and the angular source is q/(4 pi) in the (0,0) Legendre coefficient
(init.cpp:853,924)."""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import level_symmetric


@dataclass
class Problem:
    name: str
    nx: int
    ny: int
    p: int
    ax: float
    bx: float
    ay: float
    by: float
    xi: np.ndarray
    eta: np.ndarray
    w: np.ndarray
    quad_off: np.ndarray
    sigma_t: np.ndarray   # [nx+2][ny+2]
    sigma_s: np.ndarray   # [nx+2][ny+2]
    q_cell: np.ndarray    # [nx+2][ny+2] isotropic source density per cell (halos: isotropic inflow density)
    q_dg: np.ndarray = None  # optional [nx+2][ny+2][p+1][p+1] DG source density; replaces q_cell
    # optional degree-p cross sections [nx+2][ny+2][p+1][p+1] (sigma_t / sigma_s hold their cell means)
    sigma_t_dg: np.ndarray = None
    sigma_s_dg: np.ndarray = None
    reflect_mask: int = 0  # reflecting edges (REFLECT_X_MIN | ...), from the deck keys reflect_*

    @property
    def xs_degree(self):
        """DG degree of the cross sections the solver is given: p with sigma_t_dg, else 0."""
        return self.p if self.sigma_t_dg is not None else 0

    def cross_sections(self):
        """(sigma_t, sigma_s) in the form tsw takes at xs_degree: coefficient arrays or cell means."""
        if self.sigma_t_dg is not None:
            return self.sigma_t_dg, self.sigma_s_dg
        return self.sigma_t, self.sigma_s

    @property
    def nq(self):
        return int(self.xi.size)

    @property
    def source(self):
        """Angular source [nx+2][ny+2][nq][p+1][p+1]: the source density / (4 pi) in every ordinate,
        all DG coefficients of q_dg when given, else q_cell in the (0,0) coefficient."""
        if self.q_dg is not None:
            return np.ascontiguousarray(np.broadcast_to(self.q_dg[:, :, None] / (4.0 * math.pi),
                                                        (self.nx + 2, self.ny + 2, self.nq, self.p + 1, self.p + 1)))
        return _isotropic_source(self.q_cell, self.nq, self.p)

    def source_device(self, device):
        """The same angular source built directly in device memory (torch)."""
        import torch
        src = torch.zeros((self.nx + 2, self.ny + 2, self.nq, self.p + 1, self.p + 1), dtype=torch.float64,
                          device=device)
        if self.q_dg is not None:
            src[:] = torch.from_numpy(self.q_dg / (4.0 * math.pi)).to(device)[:, :, None]
            return src
        q = torch.from_numpy(self.q_cell / (4.0 * math.pi)).to(device)
        src[:, :, :, 0, 0] = q[:, :, None]
        return src

    @property
    def unknowns(self):
        """Angular unknowns per sweep, nq * nx * ny * (p+1)^2."""
        return self.nq * self.nx * self.ny * (self.p + 1) ** 2

    def set_on(self, ctx):
        ctx.problem_set(self.nx, self.ny, self.p, self.xi, self.eta, self.w, self.quad_off, self.ax, self.bx,
                        self.ay, self.by)
        if self.xs_degree:
            ctx.set_xs_degree(self.xs_degree)
        if self.reflect_mask:
            ctx.set_boundaries(self.reflect_mask)


def _centres(n, a, b):
    h = (b - a) / n
    return a + (np.arange(n + 2) - 0.5) * h  # includes halo cells 0 and n+1


def _isotropic_source(q_cell, nq, p):
    nx2, ny2 = q_cell.shape
    src = np.zeros((nx2, ny2, nq, p + 1, p + 1))
    src[:, :, :, 0, 0] = q_cell[:, :, None] / (4.0 * math.pi)
    return src


def lattice(cells_per_unit=40, p=2, order=12) -> Problem:
    """Checkerboard lattice on [0,7]^2 (LatticeProblem.cpp:173-255): absorber unit blocks with
    sigma_t = 10, sigma_s = 0; elsewhere sigma_t = sigma_s = 1; q = 1 on (3,4)^2; vacuum inflow.
    The absorber set is the one TotalCross codes (:231-255): 11 blocks, the centre column holding
    only the (3,5) block."""
    n = 7 * cells_per_unit
    x = _centres(n, 0.0, 7.0)
    X, Y = np.meshgrid(x, x, indexing="ij")
    absorbers = [(1, 1), (1, 3), (1, 5), (2, 2), (2, 4), (3, 5), (4, 2), (4, 4), (5, 1), (5, 3), (5, 5)]
    mask = np.zeros_like(X, dtype=bool)
    for bx_, by_ in absorbers:
        mask |= (X > bx_) & (X <= bx_ + 1) & (Y > by_) & (Y <= by_ + 1)
    sigma_t = np.where(mask, 10.0, 1.0)
    sigma_s = np.where(mask, 0.0, 1.0)
    q = np.where((X > 3) & (X < 4) & (Y > 3) & (Y < 4), 1.0, 0.0)
    halo = np.ones_like(mask)
    halo[1:-1, 1:-1] = False
    q[halo] = 0.0  # vacuum inflow
    xi, eta, w, off = level_symmetric(order)
    return Problem("lattice", n, n, p, 0.0, 7.0, 0.0, 7.0, xi, eta, w, off, sigma_t, sigma_s, q)


def central_source(n=64, p=1, order=8, sigma_t=1.0, sigma_s=0.5) -> Problem:
    """Unit square, uniform sigma, isotropic q = 1 on [0.375, 0.625]^2, vacuum boundaries."""
    x = _centres(n, 0.0, 1.0)
    X, Y = np.meshgrid(x, x, indexing="ij")
    q = np.where((X > 0.375) & (X < 0.625) & (Y > 0.375) & (Y < 0.625), 1.0, 0.0)
    q[0, :] = q[-1, :] = q[:, 0] = q[:, -1] = 0.0
    xi, eta, w, off = level_symmetric(order)
    st = np.full((n + 2, n + 2), sigma_t)
    ss = np.full((n + 2, n + 2), sigma_s)
    return Problem("central-source", n, n, p, 0.0, 1.0, 0.0, 1.0, xi, eta, w, off, st, ss, q)


def batched(seed=2026, count=256, n_range=(32, 256), degrees=(1, 2, 3), orders=(4, 6, 8, 10, 12)):
    """Uneven batch of independent problems on the unit square, generated from `seed` with a
    counter-based generator (numpy Philox): N ~ U{32..256} cells per side, DG degree ~ U{1,2,3},
    S_N ~ U{4..12}; an 8x8 block pattern (block edges at floor(k N / 8)) with sigma_t log-uniform in
    [0.1, 100] and source q ~ U[0,1] per block; one scattering ratio c ~ U[0, 0.999] per problem,
    sigma_s = c sigma_t; vacuum boundaries."""
    rng = np.random.Generator(np.random.Philox(seed))
    out = []
    sets = {}
    for b in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        p = int(rng.choice(degrees))
        order = int(rng.choice(orders))
        if order not in sets:
            sets[order] = level_symmetric(order)
        xi, eta, w, off = sets[order]
        sig_blk = np.exp(rng.uniform(np.log(0.1), np.log(100.0), size=(8, 8)))
        q_blk = rng.uniform(0.0, 1.0, size=(8, 8))
        c = float(rng.uniform(0.0, 0.999))
        edges = (np.arange(9) * n) // 8
        blk = np.searchsorted(edges, np.arange(n), side="right") - 1  # block of each cell index
        st = np.ones((n + 2, n + 2))
        q = np.zeros((n + 2, n + 2))
        st[1:-1, 1:-1] = sig_blk[blk[:, None], blk[None, :]]
        q[1:-1, 1:-1] = q_blk[blk[:, None], blk[None, :]]
        ss = c * st
        out.append(Problem(f"batch{b}", n, n, p, 0.0, 1.0, 0.0, 1.0, xi, eta, w, off, st, ss, q))
    return out


def as_batch_member(P: Problem):
    """The dict tsw.Context.batch_set expects for one problem."""
    return {"nx": P.nx, "ny": P.ny, "p": P.p, "xi": P.xi, "eta": P.eta, "w": P.w, "quad_off": P.quad_off,
            "ax": P.ax, "bx": P.bx, "ay": P.ay, "by": P.by}


def line_source(n=200, p=2, order=16, width=0.03):
    """Line-source pulse on [-1.5, 1.5]^2 with n x n cells, Q_p, level-symmetric S_order, sigma_s = 1,
    sigma_a = 0, no source, vacuum boundaries (the reference has no line-source problem; its nearest is
    the bump, BumpProblem.cpp:107-128).  Returns (problem, psi0): the initial condition is an isotropic
    Gaussian exp(-r^2 / (2 width^2)) / (2 pi width^2) of unit mass, divided by 4 pi per ordinate and
    projected exactly per cell (separable 12-point Gauss-Legendre quadrature of each 1D factor)."""
    from numpy.polynomial import legendre as L
    xi, eta, w, off = level_symmetric(order)
    a, b = -1.5, 1.5
    h = (b - a) / n
    xg, wg = L.leggauss(12)
    P = p + 1
    leg = np.array([L.legval(xg, np.eye(P)[d]) for d in range(P)])  # [d][node]
    centres = a + (np.arange(n) + 0.5) * h
    X = centres[:, None] + 0.5 * h * xg[None, :]                    # [cell][node]
    gx = np.exp(-X ** 2 / (2 * width ** 2)) / math.sqrt(2 * math.pi * width ** 2)
    cx = np.einsum("cn,dn,n->cd", gx, leg, wg) * (2 * np.arange(P) + 1)[None, :] / 2.0
    coef = np.einsum("id,je->ijde", cx, cx) / (4 * math.pi)
    nq = xi.size
    psi0 = np.zeros((n + 2, n + 2, nq, P, P))
    psi0[1:-1, 1:-1] = coef[:, :, None]
    st = np.ones((n + 2, n + 2))
    ss = np.ones((n + 2, n + 2))
    q = np.zeros((n + 2, n + 2))
    prob = Problem("line-source", n, n, p, a, b, a, b, xi, eta, w, off, st, ss, q)
    return prob, psi0


def ring_seed(P: Problem, r0=0.8, r1=1.0, sub=8):
    """Seed of J = integral of phi over the ring r0 <= r <= r1 about the origin, in the psi layout:
    w_q * int_{ring and cell} P_d P_e dx dy by sub x sub midpoint sub-cells per cell (cut cells resolved
    to that quadrature), so that J = <seed, psi> over the interior coefficients."""
    from numpy.polynomial import legendre as Lg
    n, p, a = P.nx, P.p, P.ax
    h = (P.bx - P.ax) / n
    Pd = p + 1
    t = (np.arange(sub) + 0.5) / sub * 2 - 1
    leg = np.array([Lg.legval(t, np.eye(Pd)[d]) for d in range(Pd)])
    centres = a + (np.arange(n) + 0.5) * h
    X = centres[:, None, None, None] + 0.5 * h * t[None, None, :, None]
    Y = centres[None, :, None, None] + 0.5 * h * t[None, None, None, :]
    R = np.sqrt(X ** 2 + Y ** 2)
    ind = ((R >= r0) & (R <= r1)).astype(float) * (h / sub) ** 2
    M = np.einsum("ijuv,du,ev->ijde", ind, leg, leg)
    seed = np.zeros((n + 2, n + 2, P.nq, Pd, Pd))
    seed[1:-1, 1:-1] = P.w[None, None, :, None, None] * M[:, :, None]
    return seed
