"""Reference-style input decks: "key value" files read by ParameterList::ReadFromFile's rules, and the
problem inputs Init builds from them (src/utils/init.cpp:422-946): cross sections, source and initial
condition projected into the DG space, inflow halos, and the named ordinate set.

Problems: "lattice" (LatticeProblem.cpp:100-324), "hohlraum" (HohlraumProblem.cpp:60-272),
"hohlraum2" (Hohlraum2Problem.cpp:14-148), "bump" (BumpProblem.cpp) and "inhomogeneous-sphere"
(InhomogeneousSphereProblem.cpp).  Cross sections are projected at DG degree 0 when piecewise_cross is
true and at the flux's degree otherwise (init.cpp:741); the reflect_* keys and periodic give the problem's
reflecting edges or periodic boundaries (init.cpp:448-510); mollified problem functions are not available
(ValueError)."""
from __future__ import annotations

import math

import numpy as np

from . import gauss_legendre, ordinates
from .fields import Mesh, legendre_table
from .problems import Problem


def read_deck(path, params=None):
    """Key/value pairs of a deck (ParameterList.cpp:202-245): text from '#' on is dropped, the line is
    lower-cased, its first two words are key and value.  As in the reference, key and value carry over
    from the previous line when a line lacks them (so a key alone takes the previous value and a blank
    line repeats the previous pair), and keys already present are not replaced."""
    out = {} if params is None else params
    key, value = "", ""
    with open(path) as f:
        lines = f.read().split("\n")
    if lines[-1] == "":  # std::getline yields no empty line after the final newline
        lines.pop()
    for line in lines:
        words = line.split("#", 1)[0].lower().split()
        if words:
            key = words[0]
            if len(words) > 1:
                value = words[1]
        out.setdefault(key, value)
    return out


def write_deck(path, params):
    """"key value" lines in key order (ParameterList.cpp:254-285)."""
    with open(path, "w") as f:
        for k in sorted(params):
            f.write(f"{k} {params[k]}\n")


def _bool(params, key, default):
    """std::boolalpha reading: "true" / "false"; other text raises ValueError."""
    if key not in params:
        return default
    v = params[key]
    if v not in ("true", "false"):
        raise ValueError(f"Error converting value '{v}' for key '{key}' to type 'bool'")
    return v == "true"


# ---------------------------------------------------------------------------------------------------
def _in(v, a):
    return (v > a) & (v <= a + 1.0)


def _zeros(x, y):
    return np.zeros(np.broadcast(x, y).shape)


class ProblemFunctions:
    """Problem functions of position (vectorised); defaults: no source, no initial condition, vacuum."""
    domain = None  # (ax, bx, ay, by) the problem imposes, or None

    def source(self, x, y):
        return _zeros(x, y)

    def initial(self, x, y):
        return _zeros(x, y)

    def inflow(self, q):
        pass


class Lattice(ProblemFunctions):
    domain = (0.0, 7.0, 0.0, 7.0)

    def source(self, x, y):
        return np.where((x > 3.0) & (x < 4.0) & (y > 3.0) & (y < 4.0), 1.0, 0.0)

    @staticmethod
    def absorber(x, y):
        a = (_in(x, 1.0) | _in(x, 5.0)) & (_in(y, 1.0) | _in(y, 3.0) | _in(y, 5.0))
        a |= (_in(x, 2.0) | _in(x, 4.0)) & (_in(y, 2.0) | _in(y, 4.0))
        return a | (_in(x, 3.0) & _in(y, 5.0))

    def total(self, x, y):
        return np.where(self.absorber(x, y), 10.0, 1.0)

    def scatter(self, x, y):
        return np.where(self.absorber(x, y), 0.0, 1.0)


class Hohlraum(ProblemFunctions):
    domain = (0.0, 1.3, 0.0, 1.3)

    @staticmethod
    def wall(x, y):
        w = (x > 1.25) | (y <= 0.05) | (y > 1.25)
        return w | ((y > 0.25) & (y <= 1.05) & ((x <= 0.05) | ((x > 0.45) & (x <= 0.85))))

    def total(self, x, y):
        return np.where(self.wall(x, y), 100.0, 0.1)

    def scatter(self, x, y):
        return np.where(self.wall(x, y), 1.0, 0.1)

    def inflow(self, q):
        q[0, :, 0, 0] = 1.0


class Hohlraum2(Hohlraum):
    def scatter(self, x, y):
        x, y = np.broadcast_arrays(x, y)
        s = np.full(x.shape, 0.1)
        conds = [  # first matching rule wins, so apply in reverse order
            ((x > 1.25) | (y > 1.25) | (y < 0.05), 50.0),
            ((x > 0.50) & (x <= 0.85) & (y > 0.30) & (y <= 1.00), 0.0),
            ((x > 0.45) & (x <= 0.50) & (y > 0.25) & (y <= 1.05), 90.0),
            ((x > 0.50) & (x <= 0.85) & (y > 1.00) & (y <= 1.05), 90.0),
            ((x > 0.50) & (x <= 0.85) & (y > 0.25) & (y <= 0.30), 90.0),
            ((x <= 0.05) & (y > 0.25) & (y <= 1.05), 95.0),
        ]
        for c, v in reversed(conds):
            s = np.where(c, v, s)
        return s


class Bump(ProblemFunctions):
    """Normalised bump exp(1/(r^2/R^2 - 1)), r < R, in uniform sigma_t = sigma_s (BumpProblem.cpp); the
    disc integral is pi R^2 int_0^1 exp(-1/t) dt."""

    def __init__(self, params):
        self.sigma = float(params["sigma_val"])
        self.R = float(params["bump_radius"])
        if not self.R > 0:
            raise ValueError("bump_radius must be positive")
        g, w = gauss_legendre(20)
        panels = 64
        t = (np.arange(panels)[:, None] + 0.5 * (g[None, :] + 1.0)) / panels
        integral = float(np.sum(0.5 / panels * w[None, :] * np.exp(-1.0 / t)))
        self.norm = 1.0 / (math.pi * self.R * self.R * integral)

    def initial(self, x, y):
        u = (x * x + y * y) / (self.R * self.R)
        with np.errstate(divide="ignore", over="ignore"):
            return np.where(u < 1.0, np.exp(1.0 / np.where(u < 1.0, u - 1.0, -1.0)) * self.norm, 0.0)

    def total(self, x, y):
        return np.full(np.broadcast(x, y).shape, self.sigma)

    scatter = total


class InhomogeneousSphere(ProblemFunctions):
    """Disc of pure scatterer with smooth transitions (InhomogeneousSphereProblem.cpp)."""

    def __init__(self, params):
        g = lambda k: float(params["sphere_" + k])
        self.ri, self.ro = g("inner_radius"), g("outer_radius")
        self.wi, self.wo = g("inner_trans_radius"), g("outer_trans_radius")
        self.si, self.sm, self.so = g("inner_sigma_val"), g("interm_sigma_val"), g("outer_sigma_val")

    @staticmethod
    def step(r, R, w):
        rho = 0.5 * w
        inside = (r > R - rho) & (r < R + rho)
        rr = np.where(inside, r, R)  # keep the exponent finite outside the transition
        with np.errstate(over="ignore"):
            return 1.0 / (1.0 + np.exp(2 * rho / ((R + rho) - rr) - 2 * rho / (rr - (R - rho))))

    def source(self, x, y):
        r = np.sqrt(x * x + y * y)
        return np.where(r <= self.ri - 0.5 * self.wi, 1.0,
                        np.where(r < self.ri + 0.5 * self.wi, self.step(r, self.ri, self.wi), 0.0))

    def total(self, x, y):
        r = np.sqrt(x * x + y * y)
        out = np.full(r.shape, self.so)
        out = np.where(r < self.ro + 0.5 * self.wo, self.so + (self.sm - self.so) * self.step(r, self.ro, self.wo), out)
        out = np.where(r <= self.ro - 0.5 * self.wo, self.sm, out)
        out = np.where(r < self.ri + 0.5 * self.wi, self.sm + (self.si - self.sm) * self.step(r, self.ri, self.wi), out)
        return np.where(r <= self.ri - 0.5 * self.wi, self.si, out)

    scatter = total


PROBLEMS = {"lattice": lambda p: Lattice(), "hohlraum": lambda p: Hohlraum(), "hohlraum2": lambda p: Hohlraum2(),
            "bump": Bump, "inhomogeneous-sphere": InhomogeneousSphere}


def project(fn, mesh: Mesh, p, apply_limiter=True, proj_order=14):
    """ComputeProjection (init.cpp:50-240): [nx+2][ny+2][p+1][p+1] coefficients from a proj_order^2
    Gauss rule per interior cell; with apply_limiter, a cell with non-negative mean whose polynomial is
    negative somewhere on the 21 x 21 sample grid has its moments of degree > 0 scaled by the largest
    theta (bisection to 1e-14) that removes the negative samples."""
    g, w = gauss_legendre(proj_order)
    X = mesh.ax + np.arange(mesh.nx)[:, None] * mesh.dx + (g[None, :] + 1.0) * mesh.dx / 2.0
    Y = mesh.ay + np.arange(mesh.ny)[:, None] * mesh.dy + (g[None, :] + 1.0) * mesh.dy / 2.0
    F = fn(X[:, None, :, None], Y[None, :, None, :])  # [nx][ny][l][m]
    L = legendre_table(p, g) * ((2 * np.arange(p + 1) + 1) / 2.0)[None, :]
    c = np.einsum("ijlm,l,m,ld,me->ijde", F, w, w, L, L)
    if apply_limiter and p > 0:
        S = legendre_table(p, -1.0 + 0.1 * np.arange(21))
        hi = np.ones((p + 1, p + 1))
        hi[0, 0] = 0.0

        def negative(cc, theta):
            scaled = cc * (theta * hi + (1.0 - hi))
            return np.any(np.einsum("...de,ld,me->...lm", scaled, S, S) < 0.0, axis=(-1, -2))

        flag = (c[..., 0, 0] >= 0.0) & negative(c, 1.0)
        for i, j in zip(*np.nonzero(flag)):
            lo, up = 0.0, 1.0
            while up - lo > 1e-14:
                mid = 0.5 * (lo + up)
                if negative(c[i, j], mid):
                    up = mid
                else:
                    lo = mid
            c[i, j] = c[i, j] * (lo * hi + (1.0 - hi))
    out = np.zeros((mesh.nx + 2, mesh.ny + 2, p + 1, p + 1))
    out[1:-1, 1:-1] = c
    return out


def init(params):
    """(Problem, extras) for a deck given as a dict (see read_deck).  Problem carries the mesh, DG
    degree, ordinate set, per-cell cross sections and the source density with inflow halos
    (q_dg); extras holds the projections by the reference's file names, the initial angular flux psi0
    (zero halos), and final_time / max_step_size (tfinal, cfl * dx)."""
    params = dict(params)
    name = params["problem"]
    if name not in PROBLEMS:
        raise KeyError(f"problem '{name}' is not available")
    if float(params["moll_radius"]) >= 1e-14:
        raise ValueError("mollified problem functions (moll_radius > 0) are not available")
    prob = PROBLEMS[name](params)
    for k, v in zip(("ax", "bx", "ay", "by"), prob.domain or ()):  # OverrideOptions: only keys the deck names
        if k in params and float(params[k]) != v:
            params[k] = repr(v)
    if name.startswith("hohlraum"):
        for k in ("periodic", "reflect_x_min"):
            if _bool(params, k, False):
                params[k] = "false"
    reflect_mask = 0  # init.cpp:448-510
    for b, k in enumerate(("reflect_x_min", "reflect_x_max", "reflect_y_min", "reflect_y_max")):
        if _bool(params, k, False):
            reflect_mask |= 1 << b
    if _bool(params, "periodic", False):
        if reflect_mask:
            raise RuntimeError("Combined reflecting and periodic boundary conditions not supported.")
        reflect_mask = 0x100  # TSW_PERIODIC
    sdeg_pwc = _bool(params, "piecewise_cross", False)
    mesh = Mesh(int(params["nx"]), int(params["ny"]), float(params["ax"]), float(params["bx"]),
                float(params["ay"]), float(params["by"]))
    p = int(params.get("dg_degree_x", params.get("dg_degree", -1)))
    if p < 0:
        raise ValueError("Failed to determine degree of DG spatial discretization")
    otype = params.get("ordinate_type", "chebyshev-legendre")
    order = int(params["ang_order"])
    reduce = _bool(params, "ordinate_sym_reduce", False)
    xi, eta, w, off = ordinates(otype, order, reduce)

    sdeg = 0 if sdeg_pwc else p   # init.cpp:741
    st = project(prob.total, mesh, sdeg)
    ss = project(prob.scatter, mesh, sdeg)
    q0 = project(prob.source, mesh, p)
    ic = project(prob.initial, mesh, p)
    q = q0.copy()
    prob.inflow(q)
    P = Problem(name, mesh.nx, mesh.ny, p, mesh.ax, mesh.bx, mesh.ay, mesh.by, xi, eta, w, off,
                st[:, :, 0, 0].copy(), ss[:, :, 0, 0].copy(), q[:, :, 0, 0].copy(), q_dg=q,
                sigma_t_dg=st if sdeg else None, sigma_s_dg=ss if sdeg else None, reflect_mask=reflect_mask)
    psi0 = np.zeros((mesh.nx + 2, mesh.ny + 2, xi.size, p + 1, p + 1))
    psi0[1:-1, 1:-1] = ic[1:-1, 1:-1, None] / (4.0 * math.pi)
    nxs = f"{mesh.nx:07d}"
    extras = {
        "psi0": psi0,
        f"sigma_t_dg{sdeg}_nx{nxs}.cs2": st, f"sigma_s_dg{sdeg}_nx{nxs}.cs2": ss,
        f"initial_dg{p}_nx{nxs}.sd2": ic, f"source_dg{p}_nx{nxs}.sd2": q0,
        "ordinate_info": (otype, order, reduce), "mesh": mesh,
        "final_time": float(params["tfinal"]) if "tfinal" in params else math.nan,
        "max_step_size": float(params["cfl"]) * mesh.dx if "cfl" in params else math.nan,
    }
    return P, extras
