"""Diffusion synthetic acceleration of the Peierls GMRES (tsw_set_tuning "use_dsa").  DSA is a left
preconditioner, so it changes the Krylov iterates and the stopping norm but not the solution: solves with
and without it must agree with each other and with the oracle's GMRES to the solve tolerance, on both
engines and in both precisions; on an optically thick, scattering-dominated problem it must cut the
iteration count.  The reference's 2D DSA does not exist (its DSA is 1D), so iteration histories are not
compared with it."""
import math

import numpy as np
import pytest

import tsw
from _common import Grid, ls_ordinates, orc_gmres, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, o, pattern, dsa, cycles=2):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern(pattern)
    c.set_tuning("use_dsa", dsa)
    c.set_tuning("dsa_cycles", cycles)
    return c


@pytest.mark.parametrize("pattern", ["hip-staggered", "hip-wavefront"])
@pytest.mark.parametrize("dt", [math.inf, 0.2])
def test_dsa_solution_unchanged(pattern, dt):
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(21, 34, 1, o, seed=61, materials=5)
    rc, ref, _, _, _ = orc_gmres(g, o, st, ss, src, dt=dt, abs_tol=1e-13, restart=30)
    assert rc == 0
    for prec, tol in ((tsw.FP64, 1e-9), (tsw.FP32, 1e-8)):
        with _ctx(g, o, pattern, 1) as c:
            psi = src.copy()
            its, res = c.gmres_solve(psi, st, ss, dt=dt, abs_tol=1e-12 if prec == tsw.FP64 else 1e-11,
                                     precision=prec)
            assert rel_err(psi, ref) <= tol, (prec, its, res)


def _thick_problem(n=48, sigma=40.0, c=0.999, p=1, order=4):
    o = ls_ordinates(order)
    g = Grid(n, n, p, 0.0, 2.0, 0.0, 2.0)
    st = np.full((n + 2, n + 2), sigma)
    ss = c * st
    src = np.zeros(g.psi_shape(o.nq))
    src[1:-1, 1:-1, :, 0, 0] = 1.0 / (4 * math.pi)
    return g, o, st, ss, src


def test_dsa_cuts_iterations_on_thick_scattering_problem():
    g, o, st, ss, src = _thick_problem()
    out = {}
    for dsa in (0, 1):
        with _ctx(g, o, "hip-staggered", dsa) as c:
            psi = src.copy()
            its, res = c.gmres_solve(psi, st, ss, abs_tol=1e-9, restart=30, max_its=5000)
            out[dsa] = (psi, its)
    print("iterations without / with DSA:", out[0][1], out[1][1])
    assert out[1][1] <= out[0][1] // 3
    assert rel_err(out[1][0], out[0][0]) <= 1e-6


def test_dsa_adjoint_solve_matches_plain():
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(16, 20, 1, o, seed=5, materials=3)
    rng = np.random.default_rng(9)
    seed = rng.normal(size=g.psi_shape(o.nq))
    res = []
    for dsa in (0, 1):
        with _ctx(g, o, "hip-staggered", dsa) as c:
            psi = src.copy()
            c.gmres_solve(psi, st, ss, abs_tol=1e-13)
            gsrc = np.zeros_like(seed)
            gst = np.zeros(st.shape)
            c.solve_vjp(seed, gsrc, st, ss, psi=psi, grad_sigma_t=gst, abs_tol=1e-12)
            res.append((gsrc, gst))
    assert rel_err(res[1][0], res[0][0]) <= 1e-9
    assert rel_err(res[1][1], res[0][1]) <= 1e-9


def test_dsa_tuning_keys():
    o = ls_ordinates(2)
    g, st, ss, src = random_problem(4, 4, 0, o, seed=1)
    with _ctx(g, o, "hip-staggered", 1, 3) as c:
        assert c.get_tuning("use_dsa") == 1 and c.get_tuning("dsa_cycles") == 3
        with pytest.raises(tsw.TswError):
            c.set_tuning("dsa_cycles", 0)
        with pytest.raises(tsw.TswError):
            c.set_tuning("use_dsa", 3)
