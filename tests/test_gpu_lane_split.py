"""Lane-split staggered sweep (tsw_set_tuning "lane_split"): 2 (Q1) or 4 (Q2, Q3) lanes per mesh row
on the material-table path, coefficients exchanged between a row's lanes by DPP quad permutations.
It is compared with the oracle over band sizes (one band, bands of a few rows, uneven last bands),
degrees, precisions, steady and time-dependent sweeps, repeated sweeps (the exchange records are reused)
and the scattering sources of source iteration and GMRES; and with the one-lane-per-row kernel, which
evaluates the same expressions in the same order, bit for bit."""
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, orc_gmres, orc_si, orc_sweep, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, o, band_rows, split, prefetch=0):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern("hip-staggered")
    c.set_tuning("band_rows", band_rows)
    c.set_tuning("lane_split", split)
    if prefetch:
        c.set_tuning("prefetch", prefetch)
    return c


def _built(p, prec):
    """Where the lane-split kernel exists: Q1-Q3 in both precisions (a lane's rows of the inverse in
    registers: FP64 Q3 holds 64 doubles per lane)."""
    return p >= 1


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("shape,band_rows", [((37, 150), 64), ((37, 150), 33), ((20, 100), -1), ((12, 64), 16),
                                             ((1, 9), -1), ((70, 3), -1), ((5, 130), 128), ((9, 257), 96)])
def test_lane_split_sweep_matches_oracle(shape, band_rows, p):
    o = ls_ordinates(6)
    g, st, _, src = random_problem(shape[0], shape[1], p, o, seed=1300 + p + shape[1], ax=0.0, bx=2.0, by=3.0,
                                   materials=5)
    print(f"lane split {shape} bands {band_rows} p={p}", flush=True)
    with _ctx(g, o, band_rows, 1) as c:
        for dt in (math.inf, 0.3):
            ref = orc_sweep(g, o, st, src, dt)
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
                for _ in range(2):  # the second sweep reuses the emptied exchange records
                    out = src.copy()
                    c.sweep(out, st, dt, precision=prec)
                    assert c.get_tuning("lane_split") == (1 if _built(p, prec) else 0), prec
                    assert rel_err(out, ref) <= tol, (dt, prec)
                    np.testing.assert_array_equal(out[0], src[0])
                    np.testing.assert_array_equal(out[:, 0], src[:, 0])


@pytest.mark.parametrize("p", [1, 2, 3])
@pytest.mark.parametrize("band_rows", [-1, 32, 64])
@pytest.mark.parametrize("prefetch", [2, 4])
def test_lane_split_bitwise_equals_one_lane_kernel(p, band_rows, prefetch):
    o = ls_ordinates(8)
    g, st, ss, src = random_problem(50, 120, p, o, seed=1400 + p, materials=4)
    outs = {}
    for split in (0, 1):
        with _ctx(g, o, band_rows if split else 0, split, prefetch) as c:
            for prec in (tsw.FP64, tsw.FP32):
                out = src.copy()
                c.sweep(out, st, 0.4, precision=prec)
                outs[split, prec] = out
                psi = src.copy()  # GMRES: sweeps with the scattering source only, and with both sources
                c.gmres_solve(psi, st, ss, abs_tol=1e-12 if prec == tsw.FP64 else 1e-9, precision=prec)
                outs[split, prec, "gmres"] = psi
    for prec in (tsw.FP64, tsw.FP32):
        if _built(p, prec):
            np.testing.assert_array_equal(outs[1, prec], outs[0, prec])
            np.testing.assert_array_equal(outs[1, prec, "gmres"], outs[0, prec, "gmres"])


def test_lane_split_without_material_table_uses_one_lane_kernel():
    o = ls_ordinates(4)
    g, st, _, src = random_problem(30, 40, 2, o, seed=1500)  # per-cell sigma_t: no material table
    ref = orc_sweep(g, o, st, src, math.inf)
    with _ctx(g, o, 0, 1) as c:
        out = src.copy()
        c.sweep(out, st, math.inf)
        assert c.get_tuning("lane_split") == 0
        assert rel_err(out, ref) <= 1e-12


@pytest.mark.parametrize("band_rows", [-1, 32])
def test_lane_split_solves_match_oracle(band_rows):
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(40, 100, 2, o, seed=1600, materials=6)
    ref_si, _, _ = orc_si(g, o, st, ss, src, abs_tol=1e-12, max_its=400)
    rc, ref_g, _, _, _ = orc_gmres(g, o, st, ss, src, abs_tol=1e-13, restart=30)
    assert rc == 0
    with _ctx(g, o, band_rows, 1) as c:
        for prec, tol in ((tsw.FP64, 1e-9), (tsw.FP32, 1e-5)):
            psi = src.copy()
            c.si_solve(psi, st, ss, abs_tol=1e-12, max_its=400, precision=prec)
            assert rel_err(psi, ref_si) <= tol, ("si", prec)
        for prec, tol in ((tsw.FP64, 1e-10), (tsw.FP32, 1e-8)):
            psi = src.copy()
            c.gmres_solve(psi, st, ss, abs_tol=1e-13 if prec == tsw.FP64 else 1e-11, precision=prec)
            assert rel_err(psi, ref_g) <= tol, ("gmres", prec)
        assert c.get_tuning("lane_split") == 1


def test_lane_split_tuning_key():
    o = ls_ordinates(2)
    g, st, _, src = random_problem(5, 5, 1, o, seed=1)
    with _ctx(g, o, 0, 0) as c:
        with pytest.raises(tsw.TswError):
            c.set_tuning("lane_split", 2)
