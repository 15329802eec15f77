"""Many right-hand sides of one problem (tsw_sweep_apply_multi): one launch over (source, ordinate) items
sharing the material table and element inverses, reading and writing the callers' arrays.  Each source
must come out bit for bit as tsw_sweep_apply sweeps it alone, and match the oracle; meshes taller than
one band (ny > 448) and misaligned device arrays take the one-after-the-other path."""
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, orc_sweep, random_problem, rel_err

pytestmark = pytest.mark.gpu


def _ctx(g, o):
    c = tsw.Context(0)
    c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    c.set_pattern("hip-staggered")
    return c


def _sources(src, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        s = src * rng.uniform(0.5, 2.0)
        s[0] = rng.uniform(0, 1, size=s[0].shape)
        s[:, -1] = rng.uniform(0, 1, size=s[:, -1].shape)
        out.append(s)
    return out


@pytest.mark.parametrize("ref_stage", [0, 1])
@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("order", [4, 8])
def test_multi_bitwise_equals_single_sweeps(p, order, ref_stage):
    o = ls_ordinates(order)
    g, st, _, src = random_problem(29, 23, p, o, seed=3100 + 10 * p + order, materials=5)
    srcs = _sources(src, 5, 7 + p)
    print(f"multi p={p} S{order} ref_stage={ref_stage}", flush=True)
    with _ctx(g, o) as c:
        c.set_tuning("ref_stage", ref_stage)
        for dt in (math.inf, 0.2):
            for prec in (tsw.FP64, tsw.FP32):
                one = []
                for s in srcs:
                    a = s.copy()
                    c.sweep(a, st, dt, precision=prec)
                    one.append(a)
                many = [s.copy() for s in srcs]
                c.sweep_timing(reset=True)
                c.sweep_multi(many, st, dt, precision=prec)
                n, _ = c.sweep_timing(reset=True)
                assert n == len(srcs)
                for a, b in zip(one, many):
                    np.testing.assert_array_equal(a, b)
            ref = orc_sweep(g, o, st, srcs[1], dt)
            mx = [s.copy() for s in srcs]
            c.sweep_multi(mx, st, dt, precision=tsw.MIXED)
            assert rel_err(mx[1], ref) <= 1e-12


def test_multi_one_source_and_oracle():
    o = ls_ordinates(12)
    g, st, _, src = random_problem(40, 37, 2, o, seed=3200, materials=3)
    with _ctx(g, o) as c:
        a = [src.copy()]
        c.sweep_multi(a, st, math.inf)
        assert rel_err(a[0], orc_sweep(g, o, st, src, math.inf)) <= 1e-12
        np.testing.assert_array_equal(a[0][0], src[0])


def test_multi_tall_mesh_and_misaligned_device_arrays():
    """ny = 460 exceeds one band: the sources are swept one after the other, with the same bits."""
    import torch
    o = ls_ordinates(4)
    g, st, _, src = random_problem(12, 460, 1, o, seed=3300, materials=2)
    srcs = _sources(src, 3, 11)
    with _ctx(g, o) as c:
        one = []
        for s in srcs:
            a = s.copy()
            c.sweep(a, st, math.inf)
            one.append(a)
        many = [s.copy() for s in srcs]
        c.sweep_multi(many, st, math.inf)
        for a, b in zip(one, many):
            np.testing.assert_array_equal(a, b)
    g, st, _, src = random_problem(20, 18, 2, o, seed=3301, materials=2)
    srcs = _sources(src, 3, 12)
    dev = torch.device("cuda:0")
    with _ctx(g, o) as c:
        st_d = torch.from_numpy(st).to(dev)
        one = []
        for s in srcs:
            a = torch.from_numpy(s).to(dev)
            c.sweep(a, st_d, math.inf)
            one.append(a.cpu().numpy())
        # one array 8 bytes off a 16-byte boundary sends the call down the one-after-the-other path
        big = torch.zeros(src.size + 1, dtype=torch.float64, device=dev)
        shifted = big[1:].view(src.shape)
        shifted.copy_(torch.from_numpy(srcs[0]))
        many = [shifted] + [torch.from_numpy(s).to(dev) for s in srcs[1:]]
        c.sweep_multi(many, st_d, math.inf)
        for a, b in zip(one, many):
            np.testing.assert_array_equal(a, b.cpu().numpy())
