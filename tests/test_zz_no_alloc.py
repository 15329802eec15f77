"""The allocation-free sweep (tsw_sweep_apply_noalloc, scratch in a per-GPU static array) against
tsw_sweep_apply on the "hip-wavefront-eager" pattern, whose kernels it shares: bitwise equal in FP64, FP32
and TSW_MIXED, on two calls in a row, the second on a smaller problem (the scratch keeps the first call's
contents); halos unchanged; a problem larger than the scratch is refused before any work."""
import ctypes
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, random_problem

pytestmark = pytest.mark.gpu


def _set(c, g, o):
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)


def test_noalloc_sweep_matches_sweep_apply_two_calls():
    import torch
    dev = torch.device("cuda:0")
    cases = [(41, 37, 2, 8, 901), (13, 9, 1, 4, 902)]   # the second call is smaller
    with tsw.Context(0) as ref_ctx, tsw.Context(0) as c:
        ref_ctx.set_pattern("hip-wavefront-eager")
        for nx, ny, p, order, seed in cases:
            o = ls_ordinates(order)
            g, st, _, src = random_problem(nx, ny, p, o, seed=seed)
            _set(ref_ctx, g, o)
            _set(c, g, o)
            d_src = torch.from_numpy(src).to(dev)
            d_st = torch.from_numpy(st).to(dev)
            for prec in (tsw.FP32, tsw.MIXED, tsw.FP64):
                for dt in (math.inf, 0.4):
                    print(f"noalloc {nx}x{ny} p={p} S{order} precision {prec} dt {dt}", flush=True)
                    want = d_src.clone()
                    ref_ctx.sweep(want, d_st, dt, precision=prec)
                    got = d_src.clone()
                    c.sweep_noalloc(got, d_st, dt, precision=prec)
                    assert torch.equal(got, want), (nx, prec, dt)
                    assert torch.equal(got[0], d_src[0]) and torch.equal(got[:, -1], d_src[:, -1])


def test_noalloc_refuses_oversized_problem_and_host_arrays():
    import torch
    dev = torch.device("cuda:0")
    o = ls_ordinates(12)
    nx = ny = 400   # 402^2 * 84 * 9 FP32 values: about 489 MB of scratch, more than the 128 MiB held
    with tsw.Context(0) as c:
        c.problem_set(nx, ny, 2, o.xi, o.eta, o.w, o.quad_off)
        psi = torch.zeros((nx + 2, ny + 2, o.nq, 3, 3), dtype=torch.float64, device=dev)
        st = torch.ones((nx + 2, ny + 2), dtype=torch.float64, device=dev)
        with pytest.raises(tsw.TswError) as ei:
            c.sweep_noalloc(psi, st, precision=tsw.FP32)
        assert ei.value.code == tsw.TSW_E_ARG and "scratch" in str(ei.value)
        assert float(psi.abs().sum()) == 0.0   # nothing ran
        with pytest.raises(TypeError):
            c.sweep_noalloc(np.zeros(psi.shape), np.ones(st.shape))
