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
