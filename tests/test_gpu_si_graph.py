"""Source iteration replayed from captured HIP graphs (tsw_set_tuning "si_graph"): the loop test runs on
the device and iterations after it stops do nothing, so the graph path must give bitwise the same
angular flux, iteration count and last change as the eager loop, whether the loop stops inside a
chunk, at max_its, or not at all (max_its = 0), and across calls that continue from the solver state.
Both are compared with the oracle's source iteration (SISolver.cpp:206-249)."""
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, orc_si, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, o, si_graph):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern("hip-staggered")
    c.set_tuning("si_graph", si_graph)
    return c


@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32])
@pytest.mark.parametrize("max_its,tol", [(400, 1e-11), (5, 1e-14), (13, 1e-14), (0, 1e-10)])
def test_graph_si_equals_eager(precision, max_its, tol):
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(30, 70, 1, o, seed=31, materials=3)
    res = []
    for mode in (1, 0):
        with _ctx(g, o, mode) as c:
            out = []
            for _ in range(2):   # the second call continues from the solver state
                psi = src.copy()
                it, err = c.si_solve(psi, st, ss, dt=0.5, abs_tol=tol, max_its=max_its, precision=precision)
                out.append((psi, it, err))
            res.append(out)
    for (pa, ia, ea), (pb, ib, eb) in zip(res[0], res[1]):
        assert ia == ib and ea == eb
        np.testing.assert_array_equal(pa, pb)
    if max_its == 0:
        assert res[0][0][1] == 0
    if precision == tsw.FP64 and max_its == 400:
        ref, it_ref, err_ref = orc_si(g, o, st, ss, src, dt=0.5, abs_tol=tol, max_its=max_its)
        assert res[0][0][1] == it_ref
        assert rel_err(res[0][0][0], ref) <= 1e-11


def test_graph_si_configuration_one_problem():
    """The CPU-runnable case: 64x64, Q1, S8, sigma_t = 1, sigma_s = 0.5, q = 1 on [0.375, 0.625]^2,
    vacuum, SI to 1e-10 (absolute change, the reference's test)."""
    from tsw import problems
    P = problems.central_source()
    g_src = P.source
    with tsw.Context(0) as c:
        P.set_on(c)
        psi = g_src.copy()
        it, err = c.si_solve(psi, P.sigma_t, P.sigma_s, abs_tol=1e-10, max_its=500)
        assert err <= 1e-10 and it > 5
    from _common import Grid, Ordinates
    g = Grid(P.nx, P.ny, P.p, P.ax, P.bx, P.ay, P.by)
    o = Ordinates(P.xi, P.eta, P.w, P.quad_off)
    ref, it_ref, _ = orc_si(g, o, P.sigma_t, P.sigma_s, g_src, abs_tol=1e-10, max_its=500)
    assert it == it_ref
    assert rel_err(psi, ref) <= 1e-11
