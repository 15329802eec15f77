"""Row-band decomposition of the staggered sweep (tsw_set_tuning "band_rows"): each ordinate's rows are
split over several pipelined workgroups that hand the y-traces of their top row on through device
memory.  Results must not depend on the split: every band size is compared with the oracle, FP64 to
1e-12 and FP32 to 2e-6 (relative l2), including uneven last bands, bands of a few rows, one-column
meshes, the material-table and elimination paths, scattering sources (source iteration and GMRES) and
repeated sweeps (the exchange records are reused from sweep to sweep)."""
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, orc_gmres, orc_si, orc_sweep, random_ordinates, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, o, band_rows, pattern="hip-staggered"):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern(pattern)
    c.set_tuning("band_rows", band_rows)
    return c


@pytest.mark.parametrize("band_rows", [-1, 0, 7, 50, 64, 100, 128])
@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("pattern", ["hip-staggered", "hip-staggered-ge"])
def test_banded_sweep_matches_oracle(band_rows, p, pattern):
    o = ls_ordinates(6)
    g, st, _, src = random_problem(37, 150, p, o, seed=500 + p, ax=0.0, bx=2.0, by=3.0, materials=5)
    print(f"bands {band_rows} p={p} {pattern}", flush=True)
    with _ctx(g, o, band_rows, pattern) as c:
        for dt in (math.inf, 0.3):
            ref = orc_sweep(g, o, st, src, dt)
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
                for _ in range(2):   # second sweep reuses the emptied exchange records
                    out = src.copy()
                    c.sweep(out, st, dt, precision=prec)
                    assert rel_err(out, ref) <= tol, (dt, prec)
                    np.testing.assert_array_equal(out[:, 0], src[:, 0])


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("shape", [(37, 150), (1, 9), (12, 64), (5, 65), (70, 3), (1, 200)])
@pytest.mark.parametrize("materials", [5, None])
def test_one_wave_sweep_matches_oracle(shape, p, materials):
    """One wavefront per band of 64 rows ("one_wave"): rows pass y-traces lane to lane and bands through
    device memory; FP64, FP32 and repeated sweeps (the emptied records are reused), steady and dt."""
    o = ls_ordinates(6)
    g, st, _, src = random_problem(shape[0], shape[1], p, o, seed=700 + p, ax=0.0, bx=2.0, by=3.0,
                                   materials=materials)
    print(f"one wave {shape} p={p} materials={materials}", flush=True)
    with _ctx(g, o, 0) as c:
        c.set_tuning("one_wave", 1)
        for dt in (math.inf, 0.3):
            ref = orc_sweep(g, o, st, src, dt)
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
                for _ in range(2):
                    out = src.copy()
                    c.sweep(out, st, dt, precision=prec)
                    assert c.get_tuning("one_wave") == 1
                    assert c.get_tuning("band_rows") == (64 if g.ny > 64 else 0)
                    assert rel_err(out, ref) <= tol, (dt, prec)
                    np.testing.assert_array_equal(out[0], src[0])


@pytest.mark.parametrize("shape", [(1, 200), (2, 129), (300, 65), (90, 448), (60, 500)])
def test_banded_sweep_shapes(shape):
    nx, ny = shape
    rng = np.random.default_rng(nx * 1000 + ny)
    o = random_ordinates(2, rng)
    g, st, _, src = random_problem(nx, ny, 1, o, seed=nx + ny)
    ref = orc_sweep(g, o, st, src, 0.5)
    for br in (0, 33, 64):
        with _ctx(g, o, br) as c:
            out = src.copy()
            c.sweep(out, st, 0.5)
            assert rel_err(out, ref) <= 1e-12, br


def test_banded_solves_match_oracle():
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(40, 130, 1, o, seed=77, materials=6)
    ref_si, _, _ = orc_si(g, o, st, ss, src, abs_tol=1e-12, max_its=400)
    rc, ref_g, _, _, _ = orc_gmres(g, o, st, ss, src, abs_tol=1e-13, restart=30)
    assert rc == 0
    for br in (-1, 16, 64):
        with _ctx(g, o, br) as c:
            for prec, tol in ((tsw.FP64, 1e-9), (tsw.FP32, 1e-5)):
                psi = src.copy()
                c.si_solve(psi, st, ss, abs_tol=1e-12, max_its=400, precision=prec)
                assert rel_err(psi, ref_si) <= tol, ("si", br, prec)
            for prec, tol in ((tsw.FP64, 1e-10), (tsw.FP32, 1e-8)):
                psi = src.copy()
                c.gmres_solve(psi, st, ss, abs_tol=1e-13 if prec == tsw.FP64 else 1e-11, precision=prec)
                assert rel_err(psi, ref_g) <= tol, ("gmres", br, prec)


def test_band_tuning_errors():
    o = ls_ordinates(2)
    g, st, _, src = random_problem(5, 5, 0, o, seed=1)
    with _ctx(g, o, 0) as c:
        with pytest.raises(tsw.TswError):
            c.set_tuning("band_rows", -2)
        with pytest.raises(tsw.TswError):
            c.set_tuning("no_such_key", 1)
        with pytest.raises(tsw.TswError):
            c.set_tuning("prefetch", 5)


@pytest.mark.parametrize("prefetch", [2, 3, 4, 6])
@pytest.mark.parametrize("band_rows", [-1, 64])
def test_prefetch_depths_match_oracle(prefetch, band_rows):
    """Every prefetch depth (steps of inputs in flight per row) gives the same sweep, both precisions,
    with and without the scattering source of the GMRES matvec."""
    o = ls_ordinates(4)
    for p in (1, 2, 3):
        g, st, ss, src = random_problem(23, 90, p, o, seed=900 + p + prefetch, materials=3)
        with _ctx(g, o, band_rows) as c:
            c.set_tuning("prefetch", prefetch)
            ref = orc_sweep(g, o, st, src, 0.7)
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
                out = src.copy()
                c.sweep(out, st, 0.7, precision=prec)
                assert rel_err(out, ref) <= tol, (p, prec)
            rc, ref_g, _, _, _ = orc_gmres(g, o, st, ss, src, abs_tol=1e-13, restart=30)
            psi = src.copy()
            c.gmres_solve(psi, st, ss, abs_tol=1e-13)
            assert rel_err(psi, ref_g) <= 1e-10


def test_autotuner_choice_is_valid_and_result_identical():
    """With band_rows and prefetch automatic the first sweep times the candidates (one-lane and lane-split
    kernels); the sweep result is bitwise the same as with fixed settings, and the reported choice is one of
    the candidates."""
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(60, 200, 2, o, seed=4242, materials=4)
    outs = []
    for autotune in (1, 0):
        with _ctx(g, o, 0) as c:
            c.set_tuning("ref_split", -1)  # the diagonal-layout kernels (not the lane split on the caller's layout)
            c.set_tuning("autotune", autotune)
            for prec in (tsw.FP64, tsw.FP32):
                for _ in range(2):
                    out = src.copy()
                    c.sweep(out, st, precision=prec)
                    outs.append(out)
                br, pf = c.get_tuning("band_rows"), c.get_tuning("prefetch")
                split = c.get_tuning("lane_split")
                # one-lane shapes: one band or 64..192-row bands; lane-split shapes: 32..128-row bands
                shapes = (32, 64, 96, 128) if split else (0, 64, 128, 192)
                assert br in shapes and pf in (2, 3, 4, 6), (br, pf, split)
                if not autotune:
                    assert br == 64 and pf == (2 if prec == tsw.FP64 else 3) and split == 0
            psi = src.copy()
            c.gmres_solve(psi, st, ss, abs_tol=1e-12)
            outs.append(psi)
    n = len(outs) // 2
    for a, b in zip(outs[:n], outs[n:]):
        np.testing.assert_array_equal(a, b)
    ref = orc_sweep(g, o, st, src)
    assert rel_err(outs[0], ref) <= 1e-12
    with _ctx(g, o, 0) as c:
        with pytest.raises(tsw.TswError):
            c.set_tuning("autotune", 2)


@pytest.mark.parametrize("lane_split", [0, 1])
def test_band_stall_is_reported_and_recovered(lane_split):
    """The stall path of the banded sweep, forced from the host (tuning key stall_test raises the stall
    flag before the sweep, so every bounded wait gives up within one poll interval and every workgroup
    still exits): the call fails with TSW_E_HIP, and the next sweep, after the exchange records are
    refilled, matches the oracle again."""
    o = ls_ordinates(4)
    g, st, _, src = random_problem(30, 140, 2, o, seed=811, materials=3)
    ref = orc_sweep(g, o, st, src)
    with _ctx(g, o, 32) as c:
        c.set_tuning("lane_split", lane_split)
        for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
            out = src.copy()
            c.sweep(out, st, precision=prec)
            assert rel_err(out, ref) <= tol
            c.set_tuning("stall_test", 1)
            out = src.copy()
            with pytest.raises(tsw.TswError) as ei:
                c.sweep(out, st, precision=prec)
            assert ei.value.code == tsw.TSW_E_HIP and "stalled" in str(ei.value)
            for _ in range(2):
                out = src.copy()
                c.sweep(out, st, precision=prec)
                assert rel_err(out, ref) <= tol, prec
