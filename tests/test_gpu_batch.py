"""Batched sweep of many independent problems (uneven size, DG degree, ordinate set, block
materials) in one set of launches, against the oracle member by member."""
import math

import numpy as np
import pytest

import tsw
from tsw import problems
from _common import Grid, Ordinates, orc_sweep, rel_err

pytestmark = pytest.mark.gpu


def oracle_sweep(P, src):
    g = Grid(P.nx, P.ny, P.p, P.ax, P.bx, P.ay, P.by)
    o = Ordinates(P.xi, P.eta, P.w, P.quad_off)
    return orc_sweep(g, o, P.sigma_t, src)


@pytest.mark.parametrize("batch_direct", [-1, 0, 1])
@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32, tsw.MIXED])
def test_small_uneven_batch_matches_oracle(precision, batch_direct):
    batch = problems.batched(seed=7, count=10, n_range=(3, 40), degrees=(0, 1, 2, 3), orders=(2, 4, 6, 8))
    rng = np.random.default_rng(3)
    srcs = []
    for P in batch:  # random inflow in the halos too
        s = P.source.copy()
        s[0] = rng.uniform(0, 1, size=s[0].shape)
        s[:, -1] = rng.uniform(0, 1, size=s[:, -1].shape)
        srcs.append(s)
    print("batch of", len(batch), [(P.nx, P.p, P.nq) for P in batch], flush=True)
    with tsw.Context(0) as ctx:
        ctx.set_tuning("batch_direct", batch_direct)
        ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
        out = [s.copy() for s in srcs]
        ctx.batch_sweep(out, precision=precision)
    # FP32 sweeps with one FP64 refinement step agree with FP64 to the square of the FP32 error
    tol = {tsw.FP64: 1e-12, tsw.FP32: 2e-6, tsw.MIXED: 1e-10}[precision]
    for P, s, o in zip(batch, srcs, out):
        e = rel_err(o, oracle_sweep(P, s))
        print(P.name, P.nx, P.p, P.nq, "rel err", e, flush=True)
        assert e <= tol, (P.name, P.nx, P.p, P.nq)
        np.testing.assert_array_equal(o[0], s[0])  # halos untouched


def test_full_batch_fp32_vs_fp64_and_oracle_members():
    """The 256-problem batch (N 32..256, Q1..Q3, S4..S12, 8x8 block materials): FP32 vs FP64 sweeps,
    and the three smallest members against the oracle."""
    import torch
    batch = problems.batched(seed=2026)
    dev = torch.device("cuda:0")
    srcs = [P.source_device(dev) for P in batch]
    sig = [torch.from_numpy(P.sigma_t).to(dev) for P in batch]
    with tsw.Context(0) as ctx:
        ctx.batch_set([problems.as_batch_member(P) for P in batch], sig)
        o64 = [s.clone() for s in srcs]
        o32 = [s.clone() for s in srcs]
        omx = [s.clone() for s in srcs]
        torch.cuda.synchronize()
        ctx.batch_sweep(o64, precision=tsw.FP64)
        ctx.batch_sweep(o32, precision=tsw.FP32)
        ctx.batch_sweep(omx, precision=tsw.MIXED)

    def err(xs):
        num = sum(float((a - b).norm() ** 2) for a, b in zip(xs, o64))
        return math.sqrt(num / sum(float(b.norm() ** 2) for b in o64))

    e32, emx = err(o32), err(omx)
    print("batch rel err vs FP64: fp32", e32, "mixed", emx, flush=True)
    assert e32 <= 2e-5
    assert emx <= 1e-10  # one FP64 refinement step
    order = np.argsort([P.unknowns for P in batch])[:3]
    for k in order:
        ref = oracle_sweep(batch[k], batch[k].source)
        assert rel_err(o64[k].cpu().numpy(), ref) <= 1e-12


def _random_halo_sources(batch, seed):
    rng = np.random.default_rng(seed)
    srcs = []
    for P in batch:
        s = P.source.copy()
        s[0] = rng.uniform(0, 1, size=s[0].shape)
        s[:, -1] = rng.uniform(0, 1, size=s[:, -1].shape)
        srcs.append(s)
    return srcs


@pytest.mark.parametrize("ref_stage", [0, 1])
@pytest.mark.parametrize("batch_direct", [-1, 1])
@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32, tsw.MIXED])
def test_reference_layout_batch_matches_oracle(precision, batch_direct, ref_stage):
    """batch_ref = 1: the staggered batch kernels read each member's source from, and write psi to, its
    own reference-layout array (storer-wave and direct-store variants, sources loaded by each lane or
    through the LDS stage); TSW_MIXED takes the FP64 sweep."""
    batch = problems.batched(seed=7, count=10, n_range=(3, 40), degrees=(0, 1, 2, 3), orders=(2, 4, 6, 8))
    srcs = _random_halo_sources(batch, 3)
    print("ref-layout batch", precision, batch_direct, flush=True)
    tol = {tsw.FP64: 1e-12, tsw.FP32: 2e-6, tsw.MIXED: 1e-12}[precision]
    with tsw.Context(0) as ctx:
        ctx.set_tuning("batch_direct", batch_direct)
        ctx.set_tuning("lanes", -1)
        ctx.set_tuning("batch_ref", 1)
        ctx.set_tuning("ref_stage", ref_stage)
        ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
        for rep in range(2):
            out = [s.copy() for s in srcs]
            ctx.batch_sweep(out, precision=precision)
            assert ctx.get_tuning("batch_ref") == 1
            for P, s, o in zip(batch, srcs, out):
                e = rel_err(o, oracle_sweep(P, s))
                assert e <= tol, (rep, P.name, P.nx, P.p, P.nq, e)
                np.testing.assert_array_equal(o[0], s[0])
                np.testing.assert_array_equal(o[:, -1], s[:, -1])


@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32])
def test_reference_layout_batch_bitwise_equals_converted_batch(precision):
    """Reading and writing the caller's layout changes where values move, not how they are computed."""
    batch = problems.batched(seed=8, count=12, n_range=(5, 70), degrees=(0, 1, 2, 3), orders=(4, 8, 12))
    srcs = _random_halo_sources(batch, 4)
    outs = {}
    for ref, stage in ((-1, 0), (1, 0), (1, 1)):
        with tsw.Context(0) as ctx:
            ctx.set_tuning("lanes", -1)
            ctx.set_tuning("batch_split", -1)
            ctx.set_tuning("batch_ref", ref)
            ctx.set_tuning("ref_stage", stage)
            ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
            out = [s.copy() for s in srcs]
            ctx.batch_sweep(out, precision=precision)
            assert ctx.get_tuning("batch_ref") == (1 if ref == 1 else 0)
            outs[ref, stage] = out
    for key in ((1, 0), (1, 1)):
        for P, a, b in zip(batch, outs[-1, 0], outs[key]):
            np.testing.assert_array_equal(a, b, err_msg=f"{key} {P.name}")


def test_reference_layout_full_batch_extremes_match_oracle():
    """The 256-problem batch through the reference-layout kernels in FP64; the thickest member (largest
    sigma_t dx) and the smallest Q3 S12 member against the oracle, every member against the converted path."""
    import torch
    batch = problems.batched(seed=2026)
    dev = torch.device("cuda:0")
    srcs = [P.source_device(dev) for P in batch]
    sig = [torch.from_numpy(P.sigma_t).to(dev) for P in batch]
    outs = {}
    for ref in (-1, 1):
        with tsw.Context(0) as ctx:
            ctx.set_tuning("lanes", -1)
            ctx.set_tuning("batch_split", -1)
            ctx.set_tuning("batch_ref", ref)
            ctx.batch_set([problems.as_batch_member(P) for P in batch], sig)
            o = [s.clone() for s in srcs]
            ctx.batch_sweep(o, precision=tsw.FP64)
            assert ctx.get_tuning("batch_ref") == (1 if ref == 1 else 0)
            outs[ref] = o
    torch.cuda.synchronize()
    for k in range(len(batch)):
        assert torch.equal(outs[-1][k], outs[1][k]), batch[k].name
    thick = max(range(len(batch)), key=lambda k: batch[k].sigma_t.max() / batch[k].nx)
    q3 = min((k for k in range(len(batch)) if batch[k].p == 3 and batch[k].nq == 84), key=lambda k: batch[k].nx)
    for k in (thick, q3):
        P = batch[k]
        print("extreme member", P.name, P.nx, P.p, P.nq, flush=True)
        assert rel_err(outs[1][k].cpu().numpy(), oracle_sweep(P, P.source)) <= 1e-12, P.name


@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32])
def test_batch_engines_give_the_same_bits(precision):
    """The four batch engines (staggered kernels through the layout conversions, staggered kernels on the
    caller's layout, ordinate-lane engine for DG degrees 0-2, lane-split kernel on the caller's layout for
    DG degrees 1-3, in one band per member, in bands of 16 rows, and with runs of 3 items per XCD; in a
    persistent launch taking items from a counter and with one workgroup per item) and the autotuner's
    choice agree bit for bit."""
    batch = problems.batched(seed=9, count=12, n_range=(5, 60), degrees=(0, 1, 2, 3), orders=(2, 6, 12))
    srcs = _random_halo_sources(batch, 6)
    outs = {}
    for name, lanes, ref, split, rows, xcd, pers in (
            ("converted", -1, -1, -1, 0, 0, 1), ("ref", -1, 1, -1, 0, 0, 1), ("lanes", 1, -1, -1, 0, 0, 1),
            ("split", -1, -1, 1, 0, 0, 1), ("split16", -1, -1, 1, 16, 0, 1), ("split16xcd", -1, -1, 1, 16, 3, 1),
            ("split16np", -1, -1, 1, 16, 0, 0), ("split16xcdnp", -1, -1, 1, 16, 3, 0), ("lanesnp", 1, -1, -1, 0, 0, 0),
            ("auto", 0, 0, 0, 0, 0, 1)):
        with tsw.Context(0) as ctx:
            ctx.set_tuning("lanes", lanes)
            ctx.set_tuning("batch_ref", ref)
            ctx.set_tuning("batch_split", split)
            ctx.set_tuning("batch_split_rows", rows)
            ctx.set_tuning("batch_xcd_tile", xcd)
            ctx.set_tuning("batch_persistent", pers)
            ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
            engines = [ctx.get_tuning(f"batch_engine_p{p}") for p in range(4)]
            print(name, "engines per DG degree", engines, flush=True)
            present = {P.p for P in batch}
            assert [p for p in range(4) if engines[p] != -1] == sorted(present)
            for p in present:
                if name == "converted":
                    assert engines[p] == 0
                elif name == "ref":
                    assert engines[p] == 1
                elif name.startswith("lanes"):  # degree 3 through the conversions (lanes are built for 0-2)
                    assert engines[p] == (2 if p <= 2 else 0)
                elif name.startswith("split"):  # degree 0 through the conversions (lane split: 1-3)
                    assert engines[p] == (3 if p >= 1 else 0)
                else:
                    assert engines[p] in (0, 1, 2, 3)
            out = [s.copy() for s in srcs]
            ctx.batch_sweep(out, precision=precision)
            outs[name] = out
    for name in ("ref", "lanes", "split", "split16", "split16xcd", "split16np", "split16xcdnp", "lanesnp", "auto"):
        for P, a, b in zip(batch, outs["converted"], outs[name]):
            np.testing.assert_array_equal(a, b, err_msg=f"{name} {P.name}")
    for P, s, o in zip(batch, srcs, outs["auto"]):
        assert rel_err(o, oracle_sweep(P, s)) <= (1e-12 if precision == tsw.FP64 else 2e-6), P.name


def _split_ctx(rows):
    ctx = tsw.Context(0)
    ctx.set_tuning("lanes", -1)
    ctx.set_tuning("batch_ref", -1)
    ctx.set_tuning("batch_split", 1)
    ctx.set_tuning("batch_split_rows", rows)
    return ctx


@pytest.mark.parametrize("rows", [0, 16])
@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32, tsw.MIXED])
def test_split_batch_matches_oracle(precision, rows):
    """batch_split = 1: the lane-split kernel sweeps each member's array in place (DG degrees 1-3; degree 0
    members go through the conversions), in one band per member or in bands of at most 16 rows with
    exchange records between them; two calls, so the records are reused.  TSW_MIXED takes the FP64 sweep."""
    batch = problems.batched(seed=7, count=10, n_range=(3, 40), degrees=(0, 1, 2, 3), orders=(2, 4, 6, 8))
    srcs = _random_halo_sources(batch, 3)
    print("split batch", precision, rows, [(P.nx, P.p, P.nq) for P in batch], flush=True)
    tol = {tsw.FP64: 1e-12, tsw.FP32: 2e-6, tsw.MIXED: 1e-12}[precision]
    with _split_ctx(rows) as ctx:
        ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
        for p in {P.p for P in batch}:
            assert ctx.get_tuning(f"batch_engine_p{p}") == (3 if p >= 1 else 0)
        for rep in range(2):
            out = [s.copy() for s in srcs]
            ctx.batch_sweep(out, precision=precision)
            for P, s, o in zip(batch, srcs, out):
                e = rel_err(o, oracle_sweep(P, s))
                assert e <= tol, (rep, P.name, P.nx, P.p, P.nq, e)
                np.testing.assert_array_equal(o[0], s[0])
                np.testing.assert_array_equal(o[:, -1], s[:, -1])


def test_split_batch_misaligned_array_takes_the_conversions():
    """A device array 8 bytes off a 16-byte boundary sends the call through the layout conversions, where
    the lane-split groups sweep the diagonal layout: the same bits as the in-place call (FP64, FP32), and
    TSW_MIXED (FP32 sweeps refined in FP64 there) within 1e-10 of the oracle."""
    import torch
    batch = problems.batched(seed=14, count=5, n_range=(20, 50), degrees=(1, 2, 3), orders=(4, 8))
    srcs = _random_halo_sources(batch, 9)
    dev = torch.device("cuda:0")
    sig = [torch.from_numpy(P.sigma_t).to(dev) for P in batch]
    with _split_ctx(16) as ctx:
        ctx.batch_set([problems.as_batch_member(P) for P in batch], sig)
        for prec in (tsw.FP64, tsw.FP32):
            aligned = [torch.from_numpy(s).to(dev) for s in srcs]
            ctx.batch_sweep(aligned, precision=prec)
            assert ctx.get_tuning("batch_ref") == 1
            big = torch.zeros(srcs[0].size + 1, dtype=torch.float64, device=dev)
            shifted = big[1:].view(srcs[0].shape)
            shifted.copy_(torch.from_numpy(srcs[0]))
            arrs = [shifted] + [torch.from_numpy(s).to(dev) for s in srcs[1:]]
            ctx.batch_sweep(arrs, precision=prec)
            assert ctx.get_tuning("batch_ref") == 0
            torch.cuda.synchronize()
            for P, a, b in zip(batch, aligned, arrs):
                assert torch.equal(a, b), (prec, P.name)
        # TSW_MIXED: FP64 in place when aligned; through the conversions, FP32 sweeps of the lane-split groups
        # on the diagonal layout with one FP64 refinement (their correction sweeps)
        arrs = [shifted] + [torch.from_numpy(s).to(dev) for s in srcs[1:]]
        shifted.copy_(torch.from_numpy(srcs[0]))
        ctx.batch_sweep(arrs, precision=tsw.MIXED)
        assert ctx.get_tuning("batch_ref") == 0
        for P, s_, a in zip(batch, srcs, arrs):
            assert rel_err(a.cpu().numpy(), oracle_sweep(P, s_)) <= 1e-10, P.name


def test_split_batch_band_stall_is_reported_and_recovered():
    """The stall path of the banded lane-split groups, forced from the host (tuning key stall_test raises
    the stall flag before the sweep, so every bounded wait gives up within one poll interval and every
    workgroup still exits): the call fails with TSW_E_HIP, and the next calls, after every member's
    exchange records are refilled, match the oracle again."""
    batch = problems.batched(seed=13, count=6, n_range=(40, 70), degrees=(1, 2, 3), orders=(4, 8))
    srcs = _random_halo_sources(batch, 7)
    refs = [oracle_sweep(P, s) for P, s in zip(batch, srcs)]
    with _split_ctx(16) as ctx:
        ctx.batch_set([problems.as_batch_member(P) for P in batch], [P.sigma_t for P in batch])
        for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6)):
            out = [s.copy() for s in srcs]
            ctx.batch_sweep(out, precision=prec)
            for P, o, r in zip(batch, out, refs):
                assert rel_err(o, r) <= tol, P.name
            ctx.set_tuning("stall_test", 1)
            with pytest.raises(tsw.TswError) as ei:
                ctx.batch_sweep([s.copy() for s in srcs], precision=prec)
            assert ei.value.code == tsw.TSW_E_HIP and "stalled" in str(ei.value)
            for _ in range(2):
                out = [s.copy() for s in srcs]
                ctx.batch_sweep(out, precision=prec)
                for P, o, r in zip(batch, out, refs):
                    assert rel_err(o, r) <= tol, (prec, P.name)


def test_split_batch_tuning_keys():
    with tsw.Context(0) as ctx:
        for key, bad in (("batch_split", 2), ("batch_split_rows", 8), ("batch_split_rows", 300),
                         ("batch_split_prefetch", 5), ("batch_xcd_tile", -1), ("batch_xcd_tile", 65),
                         ("batch_persistent", 2)):
            with pytest.raises(tsw.TswError):
                ctx.set_tuning(key, bad)
        ctx.set_tuning("batch_split_prefetch", 2)
        assert ctx.get_tuning("batch_split_prefetch") == 2
        ctx.set_tuning("batch_split_prefetch", 0)
        assert ctx.get_tuning("batch_split_prefetch") == 4
