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
        for key, bad in (("lanes", 2), ("lanes_block", 65), ("lanes_waves", 5), ("lanes_prefetch", 5)):
            with pytest.raises(tsw.TswError):
                c.set_tuning(key, bad)


@pytest.mark.parametrize("p", [0, 1, 2])
def test_lanes_prefetch_depths_give_the_same_bits(p):
    """The deeper prefetch rings (FP64 4, FP32 6) move when inputs arrive, not what is computed."""
    o = ls_ordinates(8)
    g, st, _, src = random_problem(37, 45, p, o, seed=2400 + p, materials=3)
    with _ctx(g, o, 1) as c:
        for prec, deep in ((tsw.FP64, 4), (tsw.FP32, 6)):
            outs = []
            for pf in (0, deep):
                c.set_tuning("lanes_prefetch", pf)
                out = src.copy()
                c.sweep(out, st, math.inf, precision=prec)
                assert c.get_tuning("lanes") == 1
                outs.append(out)
            np.testing.assert_array_equal(outs[0], outs[1])
            if prec == tsw.FP32:  # a depth of the other precision is refused at the launch
                c.set_tuning("lanes_prefetch", 4)
                with pytest.raises(tsw.TswError):
                    c.sweep(src.copy(), st, math.inf, precision=prec)
            c.set_tuning("lanes_prefetch", 0)


@pytest.mark.parametrize("dsa", [0, 1])
def test_lanes_gmres_matches_oracle(dsa):
    from _common import orc_gmres
    o = ls_ordinates(8)
    g, st, ss, src = random_problem(40, 52, 2, o, seed=2600 + dsa, materials=5)
    rc, ref, _, _, _ = orc_gmres(g, o, st, ss, src, abs_tol=1e-13, restart=30)
    assert rc == 0
    with _ctx(g, o, 1) as c:
        c.set_tuning("use_dsa", dsa)
        for prec, tol in ((tsw.FP64, 1e-10), (tsw.FP32, 1e-8)):
            psi = src.copy()
            its, res = c.gmres_solve(psi, st, ss, abs_tol=1e-13 if prec == tsw.FP64 else 1e-11, precision=prec)
            assert c.get_tuning("lanes") == 1
            assert rel_err(psi, ref) <= tol, (prec, its, res)


@pytest.mark.parametrize("graph", [1, 0])
def test_lanes_source_iteration_matches_oracle(graph):
    from _common import orc_si
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(30, 44, 1, o, seed=2700, materials=4)
    ref, its_ref, _ = orc_si(g, o, st, ss, src, abs_tol=1e-12, max_its=500)
    with _ctx(g, o, 1) as c:
        c.set_tuning("si_graph", graph)
        for prec, tol in ((tsw.FP64, 1e-9), (tsw.FP32, 1e-5)):
            c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)  # zero SI state
            psi = src.copy()
            its, err = c.si_solve(psi, st, ss, abs_tol=1e-12, max_its=500, precision=prec)
            assert c.get_tuning("lanes") == 1
            assert rel_err(psi, ref) <= tol, (prec, its, its_ref)
            if prec == tsw.FP64:
                assert abs(its - its_ref) <= 1, (its, its_ref)


def _batch_oracle(P, src):
    from _common import Grid, Ordinates
    g = Grid(P.nx, P.ny, P.p, P.ax, P.bx, P.ay, P.by)
    return orc_sweep(g, Ordinates(P.xi, P.eta, P.w, P.quad_off), P.sigma_t, src)


@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32, tsw.MIXED])
def test_lanes_batch_matches_oracle(precision):
    """A batch of uneven members (DG degrees 0-3: degree 3 stays on the staggered batch kernel) swept by
    the lane engine in place, twice (the records are reused), with random inflow in the halos."""
    from tsw import problems
    batch = problems.batched(seed=11, count=12, n_range=(3, 90), degrees=(0, 1, 2, 3), orders=(2, 4, 8, 12))
    rng = np.random.default_rng(5)
    srcs = []
    for P in batch:
        s = P.source.copy()
        s[0] = rng.uniform(0, 1, size=s[0].shape)
        s[:, -1] = rng.uniform(0, 1, size=s[:, -1].shape)
        srcs.append(s)
    print("lane batch", [(P.nx, P.ny, P.p, P.nq) for P in batch], precision, flush=True)
    tol = {tsw.FP64: 1e-12, tsw.FP32: 2e-6, tsw.MIXED: 1e-10}[precision]
    with tsw.Context(0) as ctx:
        ctx.set_tuning("lanes", 1)
        ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
        for rep in range(2):
            out = [s.copy() for s in srcs]
            ctx.batch_sweep(out, precision=precision)
            for P, s, o in zip(batch, srcs, out):
                e = rel_err(o, _batch_oracle(P, s))
                assert e <= tol, (rep, P.name, P.nx, P.ny, P.p, P.nq, e)
                np.testing.assert_array_equal(o[0], s[0])
                np.testing.assert_array_equal(o[:, -1], s[:, -1])


def test_lanes_batch_bitwise_equals_single_problem_sweeps():
    """A lane member of a batch gets the bits of the same member swept alone by the lane engine."""
    from tsw import problems
    batch = problems.batched(seed=12, count=6, n_range=(20, 70), degrees=(0, 1, 2), orders=(4, 8))
    with tsw.Context(0) as ctx:
        ctx.set_tuning("lanes", 1)
        ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
        for prec in (tsw.FP64, tsw.FP32):
            out = [P.source.copy() for P in batch]
            ctx.batch_sweep(out, precision=prec)
            for P, o in zip(batch, out):
                with tsw.Context(0) as c:
                    c.problem_set(P.nx, P.ny, P.p, P.xi, P.eta, P.w, P.quad_off, P.ax, P.bx, P.ay, P.by)
                    c.set_pattern("hip-staggered")
                    c.set_tuning("lanes", 1)
                    c.set_tuning("lanes_waves", 3)
                    one = P.source.copy()
                    c.sweep(one, P.sigma_t, math.inf, precision=prec)
                    assert c.get_tuning("lanes") == 1
                np.testing.assert_array_equal(o, one)


def test_lanes_full_batch_vs_staggered_batch():
    """The 256-problem batch on the lane engine against the staggered batch kernels."""
    import torch
    from tsw import problems
    batch = problems.batched(seed=2026)
    dev = torch.device("cuda:0")
    srcs = [P.source_device(dev) for P in batch]
    sig = [torch.from_numpy(P.sigma_t).to(dev) for P in batch]
    outs = {}
    for lanes in (-1, 1):
        with tsw.Context(0) as ctx:
            ctx.set_tuning("lanes", lanes)
            ctx.batch_set([problems.as_batch_member(P) for P in batch], sig)
            for prec in (tsw.FP64, tsw.FP32):
                o = [s.clone() for s in srcs]
                ctx.batch_sweep(o, precision=prec)
                outs[lanes, prec] = o
    torch.cuda.synchronize()

    def err(xs, ys):
        num = sum(float((a - b).norm() ** 2) for a, b in zip(xs, ys))
        return math.sqrt(num / sum(float(b.norm() ** 2) for b in ys))

    e64 = err(outs[1, tsw.FP64], outs[-1, tsw.FP64])
    e32 = err(outs[1, tsw.FP32], outs[-1, tsw.FP64])
    print("lane batch vs staggered batch: fp64", e64, "fp32", e32, flush=True)
    assert e64 <= 1e-14
    assert e32 <= 2e-5
