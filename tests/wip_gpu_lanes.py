"""Ordinate-lane sweep engine (tsw_set_tuning "lanes" = 1): the ordinates of a quadrant packed across the
lanes of a wavefront, rows across row groups and waves, the sweep reading and writing the caller's
reference-layout arrays with no layout conversion.  Compared with the oracle over DG degrees 0-2, ordinate
sets S2-S16 (1 to 36 ordinates per quadrant), row-group widths from one ordinate to a whole quadrant, one
to four waves per workgroup, bands (uneven last bands, one row per band), degenerate meshes, inflow halos,
steady and time-dependent sweeps, FP64 and FP32; and with the staggered kernels (diagonal layout), which
evaluate the same expressions in the same order, bit for bit."""
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, orc_sweep, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, o, lanes=1, block=0, waves=0):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern("hip-staggered")
    c.set_tuning("lanes", lanes)
    c.set_tuning("lanes_block", block)
    c.set_tuning("lanes_waves", waves)
    return c


@pytest.mark.parametrize("p", [0, 1, 2])
@pytest.mark.parametrize("order,block,waves", [(2, 0, 0), (4, 0, 0), (6, 0, 1), (8, 5, 2), (12, 0, 0),
                                               (12, 21, 4), (12, 3, 4), (16, 0, 0), (16, 36, 2), (12, 1, 1)])
def test_lanes_sweep_matches_oracle(p, order, block, waves):
    o = ls_ordinates(order)
    g, st, _, src = random_problem(23, 31, p, o, seed=2100 + 10 * p + order, ax=0.0, bx=2.0, by=3.0, materials=4)
    print(f"lanes sweep S{order} p={p} block={block} waves={waves}", flush=True)
    with _ctx(g, o, 1, block, waves) as c:
        for dt in (math.inf, 0.3):
            ref = orc_sweep(g, o, st, src, dt)
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
                for _ in range(2):  # the second sweep reuses the emptied exchange records
                    out = src.copy()
                    c.sweep(out, st, dt, precision=prec)
                    assert c.get_tuning("lanes") == 1
                    assert rel_err(out, ref) <= tol, (dt, prec)
                    np.testing.assert_array_equal(out[0], src[0])
                    np.testing.assert_array_equal(out[-1], src[-1])
                    np.testing.assert_array_equal(out[:, 0], src[:, 0])
                    np.testing.assert_array_equal(out[:, -1], src[:, -1])


@pytest.mark.parametrize("shape", [(1, 1), (1, 40), (40, 1), (3, 200), (130, 7), (64, 64)])
def test_lanes_degenerate_meshes(shape):
    o = ls_ordinates(8)
    g, st, _, src = random_problem(shape[0], shape[1], 2, o, seed=2200 + shape[0], materials=3)
    ref = orc_sweep(g, o, st, src, math.inf)
    with _ctx(g, o, 1) as c:
        out = src.copy()
        c.sweep(out, st, math.inf)
        assert rel_err(out, ref) <= 1e-12


@pytest.mark.parametrize("p", [1, 2])
def test_lanes_bitwise_equals_staggered_kernels(p):
    o = ls_ordinates(12)
    g, st, _, src = random_problem(60, 90, p, o, seed=2300 + p, materials=5)
    outs = {}
    for lanes in (-1, 1):
        with _ctx(g, o, lanes) as c:
            for prec in (tsw.FP64, tsw.FP32):
                out = src.copy()
                c.sweep(out, st, 0.4, precision=prec)
                assert c.get_tuning("lanes") == (1 if lanes == 1 else 0)
                outs[lanes, prec] = out
    for prec in (tsw.FP64, tsw.FP32):
        np.testing.assert_array_equal(outs[1, prec], outs[-1, prec])


def test_lanes_device_arrays_and_in_place():
    import torch
    o = ls_ordinates(8)
    g, st, _, src = random_problem(40, 50, 2, o, seed=2400, materials=2)
    ref = orc_sweep(g, o, st, src, math.inf)
    dev = torch.device("cuda:0")
    with _ctx(g, o, 1) as c:
        d = torch.from_numpy(src).to(dev)
        c.sweep(d, torch.from_numpy(st).to(dev))
        assert rel_err(d.cpu().numpy(), ref) <= 1e-12


def test_lanes_band_stall_reported_and_recovered():
    o = ls_ordinates(8)
    g, st, _, src = random_problem(30, 80, 1, o, seed=2500, materials=3)
    ref = orc_sweep(g, o, st, src, math.inf)
    with _ctx(g, o, 1, 0, 1) as c:  # one wave per workgroup: 12 rows per band, 7 bands
        c.set_tuning("stall_test", 1)
        with pytest.raises(tsw.TswError) as e:
            c.sweep(src.copy(), st)
        assert e.value.code == tsw.TSW_E_HIP
        out = src.copy()
        c.sweep(out, st)
        assert rel_err(out, ref) <= 1e-12


def test_lanes_tuning_keys():
    o = ls_ordinates(4)
    g, st, _, src = random_problem(5, 5, 1, o, seed=1)
    with _ctx(g, o, 0) as c:
        for key, bad in (("lanes", 2), ("lanes_block", 65), ("lanes_waves", 5)):
            with pytest.raises(tsw.TswError):
                c.set_tuning(key, bad)
