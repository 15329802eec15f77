"""Reflecting boundaries (tsw_set_boundaries, tsw_reflect_boundaries): the sweep with reflecting edges
against the reference's own SweepOperator::Apply on reflecting decks (tests/golden/ref_fixtures.npz) and
against the oracle's restatement on random problems (every edge combination, DG degrees 0-3, FP64 to 1e-12,
FP32 to 2e-6, TSW_MIXED = the FP64 sweep), ghost cells included; source iteration with reflecting edges
against the reference's SISolver (its iteration count, its angular flux) and the oracle; periodic
boundaries (deck key periodic) the same way; the halo updates alone bit for bit; the configurations
refused."""
import math
import os

import numpy as np
import pytest

import tsw
from _common import (REPO, ls_ordinates, orc_periodic_halos, orc_reflect_boundaries, orc_si_reflect,
                     orc_sweep_reflect, random_problem, rel_err)

pytestmark = pytest.mark.gpu
FIX = os.path.join(REPO, "tests", "golden", "ref_fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _ctx(g, o):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern("hip-staggered")
    return c


@pytest.mark.parametrize("name", ["reflect_xmin", "reflect_xmax_ymin", "reflect_all", "reflect_ymax_q3"])
def test_reflecting_sweep_matches_reference(fx, name):
    nx, ny, p, nq, sd = (int(v) for v in fx[name + "/dims"])
    ax, bx, ay, by = (float(v) for v in fx[name + "/box"])
    mask = int(fx[name + "/mask"])
    st = fx[name + "/sigma_t"][..., 0, 0].copy()
    print(f"reference reflecting fixture {name} mask {mask}", flush=True)
    with tsw.Context(0) as c:
        c.problem_set(nx, ny, p, fx[name + "/xi"], fx[name + "/eta"], fx[name + "/w"], fx[name + "/quad"], ax, bx, ay, by)
        c.set_pattern("hip-staggered")
        c.set_boundaries(mask)
        for key, dt in (("steady", math.inf), ("dt03", 0.3)):
            ref = fx[name + "/" + key]
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6), (tsw.MIXED, 1e-12)):
                out = fx[name + "/src"].copy()
                c.sweep(out, st, dt, precision=prec)
                assert rel_err(out, ref) <= tol, (key, prec)
                # ghost cells: the reflected values (or the untouched inflow on other edges)
                halo = np.ones(out.shape[:2], dtype=bool)
                halo[1:-1, 1:-1] = False
                assert rel_err(out[halo], ref[halo]) <= tol, (key, prec)


def test_reflecting_sweep_refuses_degree_p_cross_sections(fx):
    name = "reflect_xmin_ymax_ho"
    nx, ny, p, nq, sd = (int(v) for v in fx[name + "/dims"])
    ax, bx, ay, by = (float(v) for v in fx[name + "/box"])
    with tsw.Context(0) as c:
        c.problem_set(nx, ny, p, fx[name + "/xi"], fx[name + "/eta"], fx[name + "/w"], fx[name + "/quad"], ax, bx, ay, by)
        c.set_xs_degree(sd)
        c.set_boundaries(int(fx[name + "/mask"]))
        with pytest.raises(tsw.TswError):
            c.sweep(fx[name + "/src"].copy(), fx[name + "/sigma_t"], math.inf)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("mask", [1, 2, 4, 8, 5, 10, 15])
def test_reflecting_sweep_matches_oracle(p, mask):
    o = ls_ordinates(8 if p < 3 else 4)
    g, st, _, src = random_problem(17, 13, p, o, seed=4000 + 16 * p + mask, materials=4)
    print(f"reflecting sweep p={p} mask={mask}", flush=True)
    with _ctx(g, o) as c:
        c.set_boundaries(mask)
        for dt in (math.inf, 0.25):
            ref = orc_sweep_reflect(g, o, st, src, mask, dt)
            outs = {}
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6), (tsw.MIXED, 1e-12)):
                out = src.copy()
                c.sweep(out, st, dt, precision=prec)
                assert rel_err(out, ref) <= tol, (dt, prec)
                outs[prec] = out
            np.testing.assert_array_equal(outs[tsw.MIXED], outs[tsw.FP64])
        # many right-hand sides with reflecting edges: one after the other, the same bits
        srcs = [src, src * 0.5 + 0.1]
        one = []
        for s in srcs:
            a = s.copy()
            c.sweep(a, st, math.inf)
            one.append(a)
        many = [s.copy() for s in srcs]
        c.sweep_multi(many, st, math.inf)
        for a, b in zip(one, many):
            np.testing.assert_array_equal(a, b)
        # back to inflow edges
        c.set_boundaries(0)
        a = src.copy()
        c.sweep(a, st, math.inf)
        np.testing.assert_array_equal(a[0], src[0])


@pytest.mark.parametrize("mask", [1, 6, 9, 15])
def test_reflect_boundaries_bitwise(mask):
    o = ls_ordinates(6)
    g, _, _, src = random_problem(9, 11, 2, o, seed=4100 + mask)
    rng = np.random.default_rng(mask)
    psi = rng.uniform(-1.0, 1.0, size=src.shape)
    with _ctx(g, o) as c:
        c.set_boundaries(mask)
        out = psi.copy()
        c.reflect_boundaries(out)
        np.testing.assert_array_equal(out, orc_reflect_boundaries(g, o, psi, mask))
        np.testing.assert_array_equal(out[1:-1, 1:-1], psi[1:-1, 1:-1])


def test_reflecting_configurations_refused():
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(8, 8, 1, o, seed=4200)
    with _ctx(g, o) as c:
        with pytest.raises(tsw.TswError):
            c.set_boundaries(16)
        c.set_boundaries(15)
        with pytest.raises(tsw.TswError):  # GMRES is not built for reflecting edges
            c.gmres_solve(src.copy(), st, ss)
        # tsw_problem_set restores inflow edges
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        c.gmres_solve(src.copy(), st, ss)
    # an unreduced set has ordinates sharing (xi, eta): the map must come from the caller
    xi, eta, w, off = tsw.ordinates("chebyshev-legendre", 4, False)
    with tsw.Context(0) as c:
        c.problem_set(6, 6, 1, xi, eta, w, off)
        with pytest.raises(tsw.TswError):
            c.set_boundaries(1)
        nq = len(xi)
        xmirror = np.array([int(np.argmin(np.abs(xi + xi[q]) + np.abs(eta - eta[q]))) for q in range(nq)])
        ymirror = np.array([int(np.argmin(np.abs(xi - xi[q]) + np.abs(eta + eta[q]))) for q in range(nq)])
        c.set_boundaries(1, np.stack([xmirror, ymirror]))
    # taller than one band
    g, st, _, src = random_problem(6, 460, 1, o, seed=4201)
    with _ctx(g, o) as c:
        c.set_boundaries(4)
        with pytest.raises(tsw.TswError):
            c.sweep(src.copy(), st, math.inf)


def test_reflecting_source_iteration_matches_reference(fx):
    name = "si_reflect"
    nx, ny, p, nq, sd = (int(v) for v in fx[name + "/dims"])
    ax, bx, ay, by = (float(v) for v in fx[name + "/box"])
    st = fx[name + "/sigma_t"][..., 0, 0].copy()
    ss = fx[name + "/sigma_s"][..., 0, 0].copy()
    print("reference reflecting SI fixture", flush=True)
    with tsw.Context(0) as c:
        c.problem_set(nx, ny, p, fx[name + "/xi"], fx[name + "/eta"], fx[name + "/w"], fx[name + "/quad"], ax, bx, ay, by)
        c.set_pattern("hip-staggered")
        c.set_boundaries(int(fx[name + "/mask"]))
        out = fx[name + "/src"].copy()
        its, err = c.si_solve(out, st, ss, abs_tol=float(fx[name + "/tol"]), max_its=5000)
        assert its == int(fx[name + "/its"])
        assert rel_err(out, fx[name + "/psi"]) <= 1e-11


@pytest.mark.parametrize("mask", [2, 12, 15])
def test_reflecting_source_iteration_matches_oracle(mask):
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(14, 12, 2, o, seed=4300 + mask, materials=3)
    ref, rits, _ = orc_si_reflect(g, o, st, ss, src, mask, abs_tol=1e-11, max_its=2000)
    with _ctx(g, o) as c:
        c.set_boundaries(mask)
        out = src.copy()
        its, _ = c.si_solve(out, st, ss, abs_tol=1e-11, max_its=2000)
        assert its == rits
        assert rel_err(out, ref) <= 1e-10
        # the solver state persists: a second solve starts from it, as SISolver::psi_ptr does
        ref2, rits2, _ = orc_si_reflect(g, o, st, ss, src, mask, abs_tol=1e-11, max_its=2000, psi0=ref)
        out2 = src.copy()
        its2, _ = c.si_solve(out2, st, ss, abs_tol=1e-11, max_its=2000)
        assert its2 == rits2
        assert rel_err(out2, ref2) <= 1e-10
        out32 = src.copy()
        c.si_solve(out32, st, ss, abs_tol=1e-11, max_its=2000, precision=tsw.FP32)
        assert rel_err(out32, ref) <= 1e-5


def test_periodic_source_iteration_matches_reference(fx):
    name = "si_periodic"
    nx, ny, p, nq, sd = (int(v) for v in fx[name + "/dims"])
    ax, bx, ay, by = (float(v) for v in fx[name + "/box"])
    st = fx[name + "/sigma_t"][..., 0, 0].copy()
    ss = fx[name + "/sigma_s"][..., 0, 0].copy()
    print("reference periodic SI fixture", flush=True)
    with tsw.Context(0) as c:
        c.problem_set(nx, ny, p, fx[name + "/xi"], fx[name + "/eta"], fx[name + "/w"], fx[name + "/quad"], ax, bx, ay, by)
        c.set_pattern("hip-staggered")
        c.set_boundaries(tsw.PERIODIC)
        out = fx[name + "/src"].copy()
        its, _ = c.si_solve(out, st, ss, abs_tol=float(fx[name + "/tol"]), max_its=5000)
        assert its == int(fx[name + "/its"])
        assert rel_err(out, fx[name + "/psi"]) <= 1e-11
        with pytest.raises(tsw.TswError):  # GMRES is not built for periodic boundaries
            c.gmres_solve(fx[name + "/src"].copy(), st, ss)
        with pytest.raises(tsw.TswError):  # the reference refuses the combination
            c.set_boundaries(tsw.PERIODIC | tsw.REFLECT_X_MIN)


@pytest.mark.parametrize("bc", [tsw.PERIODIC, 5, 15])
def test_synchronize_halos_bitwise(bc):
    o = ls_ordinates(4)
    g, _, _, src = random_problem(10, 7, 1, o, seed=4500 + bc)
    psi = np.random.default_rng(bc).uniform(-1.0, 1.0, size=src.shape)
    with _ctx(g, o) as c:
        c.set_boundaries(bc)
        out = psi.copy()
        c.synchronize_halos(out)
        ref = orc_periodic_halos(g, o, psi) if bc == tsw.PERIODIC else orc_reflect_boundaries(g, o, psi, bc)
        np.testing.assert_array_equal(out, ref)


def test_periodic_deck_source_iteration_matches_oracle(tmp_path):
    from tsw import deck
    from _common import Grid, Ordinates
    text = ("ax 0.0\nbx 7.0\nay 0.0\nby 7.0\nnx 14\nny 12\ndg_degree 2\npiecewise_cross true\nproblem lattice\n"
            "moll_radius 0.0\nordinate_type spherical-triangle\nang_order 2\nordinate_sym_reduce true\n"
            "solve_type source-iteration\nsweep_pattern wavefront\nabs_tol 1e-11\nmax_its 3000\nperiodic true\n")
    (tmp_path / "per.deck").write_text(text)
    P, _ = deck.init(deck.read_deck(str(tmp_path / "per.deck")))
    assert P.reflect_mask == tsw.PERIODIC
    g = Grid(P.nx, P.ny, P.p, P.ax, P.bx, P.ay, P.by)
    o = Ordinates(P.xi, P.eta, P.w, P.quad_off)
    ref, its, _ = orc_si_reflect(g, o, P.sigma_t, P.sigma_s, P.source, 0, abs_tol=1e-11, max_its=3000, periodic=True)
    with tsw.Context(0) as c:
        P.set_on(c)
        c.set_pattern("hip-staggered")
        out = P.source.copy()
        gits, _ = c.si_solve(out, P.sigma_t, P.sigma_s, abs_tol=1e-11, max_its=3000)
    assert gits == its
    assert rel_err(out, ref) <= 1e-10
