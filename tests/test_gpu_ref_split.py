"""Sweeps on the caller's layout by the lane-split kernel (tuning key ref_split, on by default for
tsw_sweep_apply and tsw_sweep_apply_multi with the staggered pattern's default tuning): the oracle's
sweep (FP64 within 1e-12, FP32 within 2e-6, TSW_MIXED as FP64), the same bits as the diagonal-layout
path (ref_split -1), row bands with exchange records (meshes taller than one workgroup's rows), many
right-hand sides against single sweeps, misaligned device arrays through the conversions, and the
forced stall path."""
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, orc_sweep, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, o, ref_split=1):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern("hip-staggered")
    c.set_tuning("ref_split", ref_split)
    return c


# (nx, ny, p): one band; two bands at degree 1 (256 rows per workgroup) and 2-3 (128 rows)
SHAPES = [(17, 23, 1), (11, 300, 1), (21, 40, 2), (13, 150, 2), (9, 31, 3), (12, 170, 3)]


@pytest.mark.parametrize("dt", [math.inf, 0.3])
@pytest.mark.parametrize("shape", SHAPES)
def test_ref_split_sweep_matches_oracle_and_diagonal_path(shape, dt):
    nx, ny, p = shape
    o = ls_ordinates(6)
    g, st, _, src = random_problem(nx, ny, p, o, seed=900 + nx + p, materials=3)
    print("ref split", shape, dt, flush=True)
    ref = orc_sweep(g, o, st, src, dt)
    outs = {}
    for rs in (1, -1):
        with _ctx(g, o, rs) as c:
            for prec in (tsw.FP64, tsw.FP32, tsw.MIXED):
                out = src.copy()
                c.sweep(out, st, dt, precision=prec)
                outs[rs, prec] = out
                if rs == 1:
                    assert c.get_tuning("lane_split") == 1
    for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6), (tsw.MIXED, 1e-12)):
        assert rel_err(outs[1, prec], ref) <= tol, prec
        np.testing.assert_array_equal(outs[1, prec][0], src[0])
        np.testing.assert_array_equal(outs[1, prec][:, -1], src[:, -1])
    # the same arithmetic as the diagonal-layout kernels (TSW_MIXED there refines an FP32 sweep)
    np.testing.assert_array_equal(outs[1, tsw.FP64], outs[-1, tsw.FP64])
    np.testing.assert_array_equal(outs[1, tsw.FP32], outs[-1, tsw.FP32])
    np.testing.assert_array_equal(outs[1, tsw.MIXED], outs[1, tsw.FP64])


def test_ref_split_multi_equals_single_sweeps_and_misaligned_fallback():
    import torch
    o = ls_ordinates(8)
    g, st, _, src = random_problem(14, 160, 2, o, seed=911, materials=4)
    rng = np.random.default_rng(5)
    srcs = [src + rng.uniform(0, 1, size=src.shape) * (k % 2) for k in range(5)]
    dev = torch.device("cuda:0")
    with _ctx(g, o) as c:
        st_d = torch.from_numpy(st).to(dev)
        for prec in (tsw.FP64, tsw.FP32):
            one = []
            for s_ in srcs:
                a = torch.from_numpy(s_).to(dev)
                c.sweep(a, st_d, math.inf, precision=prec)
                one.append(a)
            many = [torch.from_numpy(s_).to(dev) for s_ in srcs]
            c.sweep_multi(many, st_d, math.inf, precision=prec)
            big = torch.zeros(src.size + 1, dtype=torch.float64, device=dev)
            shifted = big[1:].view(src.shape)  # 8 bytes off a 16-byte boundary: the conversions
            shifted.copy_(torch.from_numpy(srcs[0]))
            c.sweep(shifted, st_d, math.inf, precision=prec)
            torch.cuda.synchronize()
            for a, b in zip(one, many):
                assert torch.equal(a, b), prec
            assert torch.equal(one[0], shifted), prec
        assert rel_err(one[0].cpu().numpy(), orc_sweep(g, o, st, srcs[0], math.inf)) <= 2e-6


def test_ref_split_band_stall_is_reported_and_recovered():
    o = ls_ordinates(4)
    g, st, _, src = random_problem(30, 140, 2, o, seed=812, materials=3)
    ref = orc_sweep(g, o, st, src)
    with _ctx(g, o) as c:
        for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
            out = src.copy()
            c.sweep(out, st, precision=prec)
            assert rel_err(out, ref) <= tol
            c.set_tuning("stall_test", 1)
            with pytest.raises(tsw.TswError) as ei:
                c.sweep(src.copy(), st, precision=prec)
            assert ei.value.code == tsw.TSW_E_HIP and "stalled" in str(ei.value)
            for _ in range(2):
                out = src.copy()
                c.sweep(out, st, precision=prec)
                assert rel_err(out, ref) <= tol, prec


def test_ref_split_tuning_key():
    o = ls_ordinates(4)
    g, st, _, src = random_problem(5, 5, 1, o, seed=1)
    with _ctx(g, o) as c:
        assert c.get_tuning("ref_split") == 1
        for bad in (0, 2):
            with pytest.raises(tsw.TswError):
                c.set_tuning("ref_split", bad)
