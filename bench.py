#!/usr/bin/env python3
"""Benchmark: sweep-preconditioned GMRES(30) with DSA on the 280x280 Q2 S12 checkerboard lattice, FP64
sweeps versus mixed precision (FP64 Krylov vectors, DSA and residual estimates; FP32 sweeps in the Arnoldi
steps once the residual estimate admits their error), both to the absolute residual 1e-8.

One step = one full transport solve of the lattice (inputs resident in HBM).  `value` is the
mixed-precision solve's angular unknowns swept per second (unknowns per sweep x sweeps / wall time).

The roofline entry is for the FP32 sweep kernel under the operator contract (SweepOperator::Apply:
read the source psi, write psi in place, read sigma): algorithmic bytes per sweep
U*(4+4) + ncell*4, divided by the sweep kernel's average duration, measured with HIP events the
library records on the stream it launches the sweep on (tsw_sweep_timing).  `--copy-probe` adds a
device copy of known size, the calibration for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs.

The CPU baseline is the reference's own sweep (its sources compiled unmodified, OpenMP, SIMD_LEN 16) on a
bounded sample of sweeps of the same lattice, or the oracle's OpenMP sweep where that build is absent.

The line carries a `batched` sub-object: the 256-problem uneven batch (the configuration the
mixed-precision targets are stated for) swept in FP32, mixed (FP32 + one FP64 refinement step) and FP64,
with whole-call and kernel times, errors against FP64 and the FP32 kernels' roofline.
"""
import argparse
import ctypes
import json
import math
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "idc-transport-solver_amd"))

HBM_PEAK_GBS = 8000.0


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def cpu_baseline(P, sweeps):
    """Times the oracle's FP64 wavefront sweep on `sweeps` sweeps of the same problem, with each diagonal's
    cells x ordinates shared among the host threads (OMP_NUM_THREADS, else the CPU count, at most 16:
    one GPU's share of the host)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from _common import Grid, Ordinates, orc_sweep_omp
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(os.cpu_count() or 1, 16)
    g = Grid(P.nx, P.ny, P.p, P.ax, P.bx, P.ay, P.by)
    o = Ordinates(P.xi, P.eta, P.w, P.quad_off)
    src = P.source
    orc_sweep_omp(g, o, P.sigma_t, src, threads=threads)   # warm-up (threads, pages)
    t0 = time.perf_counter()
    for _ in range(sweeps):
        orc_sweep_omp(g, o, P.sigma_t, src, threads=threads)
    dt = time.perf_counter() - t0
    return {"value": P.unknowns * sweeps / dt, "unit": "angular unknowns swept/s", "cores": threads, "kind": "port",
            "sample": f"{sweeps} FP64 sweeps of the {P.nx}x{P.ny} Q{P.p} nq={P.nq} {P.name} problem by the oracle's "
                      f"wavefront sweep on {threads} threads ({dt:.2f} s) on {cpu_model()}"}


REF_LATTICE_DECK = ("ax 0.0\nbx 7.0\nay 0.0\nby 7.0\nnx {n}\nny {n}\ndg_degree 2\npiecewise_cross true\nproblem lattice\n"
                    "moll_radius 0.0\nordinate_type spherical-triangle\nang_order 5\nordinate_sym_reduce true\n"
                    "solve_type source-iteration\nsweep_pattern wavefront\nsweep_solve gaussian-elimination\n"
                    "abs_tol 1e-8\nmax_its 500\ntfinal 1.0\noutput_log false\n")


def cpu_baseline_reference(n, sweeps):
    """The reference's own sweep, compiled unmodified from its sources with its optimisation flags
    (oracle/_ref/libref_fast.so: -O3, OpenMP, SIMD blocking with SIMD_LEN 16 as in
    make/Makefile.workstation.inc:30), on the n x n checkerboard lattice at Q2.  The reference has no
    level-symmetric set, so its T_5 set reduced to the hemisphere (100 ordinates) stands in for S12 (84);
    the unknowns/s figure counts the ordinates actually swept.  Threads: OMP_NUM_THREADS (the host share
    this process runs with).  None when the library is not there."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import tempfile
    import _ref
    if not _ref.available(fast=True):
        return None
    threads = int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1)
    with tempfile.TemporaryDirectory() as tmp:
        path = os.path.join(tmp, "lattice.deck")
        with open(path, "w") as f:
            f.write(REF_LATTICE_DECK.format(n=n))
        old = os.dup(1)  # the reference's Init prints its parameter summary on stdout
        os.dup2(os.open(os.devnull, os.O_WRONLY), 1)
        try:
            _ref.lib(fast=True)
            d = _ref.Deck(path)
            src = d.get_flux(1)
            d.sweep(src)  # warm-up (threads, pages, cell-solve set-up)
            t0 = time.perf_counter()
            for _ in range(sweeps):
                d.sweep(src)
            dt = time.perf_counter() - t0
        finally:
            ctypes.CDLL(None).fflush(None)  # the C library's buffered output goes to /dev/null too
            os.dup2(old, 1)
            os.close(old)
    U = d.nq * d.nx * d.ny * (d.p + 1) ** 2
    return {"value": U * sweeps / dt, "unit": "angular unknowns swept/s", "cores": threads, "kind": "reference",
            "sample": f"{sweeps} FP64 sweeps by the reference's own wavefront sweep (its sources compiled with -O3, "
                      f"OpenMP, SIMD_LEN 16) of the {d.nx}x{d.ny} Q2 checkerboard lattice with T_5 reduced "
                      f"({d.nq} ordinates; the reference has no S12), {dt:.2f} s on {threads} threads of "
                      f"{cpu_model()}"}


def bench_si(args):
    """Source iteration on the CPU-runnable case (64x64, Q1, S8, sigma_t=1, sigma_s=0.5, central source,
    vacuum) to 1e-10: graph-replayed iterations (device-side loop test) against the eager loop."""
    import torch
    import tsw
    from tsw import problems
    P = problems.central_source()
    dev = torch.device("cuda:0")
    d_src = P.source_device(dev)
    d_st = torch.from_numpy(P.sigma_t).to(dev)
    d_ss = torch.from_numpy(P.sigma_s).to(dev)
    work = torch.empty_like(d_src)
    out = {}
    for mode in (1, 0):
        ctx = tsw.Context(0)
        P.set_on(ctx)
        ctx.set_tuning("si_graph", mode)
        its = []
        for rep in range(args.warmup + args.steps):
            if rep == args.warmup:
                torch.cuda.synchronize()
                t0 = time.perf_counter()
            ctx.problem_set(P.nx, P.ny, P.p, P.xi, P.eta, P.w, P.quad_off)   # zero solver state
            work.copy_(d_src)
            it, err = ctx.si_solve(work, d_st, d_ss, abs_tol=1e-10, max_its=1000)
            its.append(it)
        torch.cuda.synchronize()
        wall = (time.perf_counter() - t0) / args.steps
        out["graph" if mode else "eager"] = (wall, its[-1])
        print(f"bench: si {'graph' if mode else 'eager'}: {wall * 1e3:.3f} ms per solve, {its[-1]} iterations",
              file=sys.stderr, flush=True)
        ctx.close()
    U = P.unknowns
    wg, ig = out["graph"]
    we, ie = out["eager"]
    print(json.dumps({"metric": "angular unknowns swept/s (source iteration)", "value": U * ig / wg,
                      "unit": "angular unknowns swept/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                      "ms_per_step": wg * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                      "dtype": "f64", "data": "synthetic",
                      "config": {"workload": "central source 64x64, Q1, S8, sigma_t=1, sigma_s=0.5, SI to 1e-10",
                                 "iterations": ig},
                      "eager": {"value": U * ie / we, "ms_per_step": we * 1e3, "iterations": ie},
                      "speedup_graph_vs_eager": we / wg}), flush=True)


def bench_idc(args):
    """IMEX/IDC line-source pulse: 200x200, Q2, S16, 4 Gauss-Lobatto nodes, 3 corrections, dt = 0.01, FP32
    sweeps with FP64 correction sums; `--idc-steps` of the 100 steps to t = 1 are timed (inputs resident),
    then the adjoint of the ring response (0.8 <= r <= 1.0) w.r.t. the initial condition over the same
    steps, with sigma_s gradients and a checkpoint every 10 steps."""
    import torch
    import tsw
    from tsw import problems
    dev = torch.device("cuda:0")
    LP, psi0_h = problems.line_source()
    n, p = LP.nx, LP.p
    h = (LP.bx - LP.ax) / n
    ctx = tsw.Context(0)
    LP.set_on(ctx)
    ctx.idc_set(4, 3, tsw.IDC_LOBATTO, False)
    psi = torch.from_numpy(psi0_h).to(dev)
    psi0 = psi.clone()
    src = torch.zeros_like(psi)
    st = torch.from_numpy(LP.sigma_t).to(dev)
    ss = torch.from_numpy(LP.sigma_s).to(dev)
    dt, steps = 0.01, args.idc_steps
    ctx.idc_step(psi, src, st, ss, dt, abs_tol=1e-10, precision=tsw.FP32)  # warm-up (tables, tuning)
    psi.copy_(psi0)
    ctx.sweep_timing(reset=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ctx.idc_run(psi, src, st, ss, dt, steps, abs_tol=1e-10, precision=tsw.FP32)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    nsw, swms = ctx.sweep_timing(reset=True)
    U = LP.unknowns
    w_d = torch.from_numpy(LP.w).to(dev)
    phi_mass = float((psi[1:-1, 1:-1, :, 0, 0] * w_d).sum()) * h ** 2   # integral of phi (1 at t=0)
    print(f"bench: idc {steps} steps {wall:.3f}s, {nsw} sweeps ({swms:.1f} ms in sweeps)", file=sys.stderr, flush=True)
    seed = torch.from_numpy(problems.ring_seed(LP)).to(dev)
    J = float((seed * psi).sum())
    grad = torch.zeros_like(psi)
    gss = torch.zeros_like(ss)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ctx.idc_adjoint(seed, grad, src, st, ss, dt, steps, psi0=psi0, checkpoint_every=10, grad_sigma_s=gss,
                    abs_tol=1e-10, precision=tsw.FP32)
    torch.cuda.synchronize()
    wadj = time.perf_counter() - t1
    # first-order check of the initial-condition gradient: J is affine in psi0
    pert = torch.zeros_like(psi0)
    pert[1:-1, 1:-1] = 1e-3 * torch.rand(pert[1:-1, 1:-1].shape, device=dev, dtype=torch.float64)
    x = psi0 + pert
    ctx.idc_run(x, src, st, ss, dt, steps, abs_tol=1e-10, precision=tsw.FP32)
    dJ = float((seed * x).sum()) - J
    pred = float((grad * pert).sum())
    print(f"bench: idc adjoint {wadj:.3f}s; J = {J:.6e}; dJ {dJ:.6e} vs grad.dpsi0 {pred:.6e}", file=sys.stderr,
          flush=True)
    ctx.close()
    print(json.dumps({
        "metric": "angular unknowns swept/s (IDC line source)", "value": U * nsw / wall,
        "unit": "angular unknowns swept/s", "n_gpus": 1, "steps": steps, "warmup": 1,
        "ms_per_step": wall / max(steps, 1) * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 sweep / f64 correction sums", "data": "synthetic (Gaussian pulse)",
        "config": {"workload": "line-source pulse 200x200, Q2, S16 (144), sigma_s=1, sigma_a=0, IDC 4 Lobatto "
                               "nodes / 3 corrections, dt=0.01", "timed_steps": steps, "sweeps": nsw,
                   "mass_phi_t": phi_mass},
        "adjoint": {"seconds": wadj, "checkpoint_every": 10, "J_ring": J, "dJ_fd": dJ, "dJ_grad": pred,
                    "rel_diff": abs(dJ - pred) / max(abs(dJ), 1e-300)}}), flush=True)


def bench_batched(args, steps=None, warmup=None, cpu_sweeps=None):
    """Batched sweep of the seeded 256-problem uneven batch (N 32..256, Q1..Q3, S4..S12, 8x8 block
    materials): one step = one tsw_batch_sweep call sweeping every member in place, inputs resident in HBM.
    Timed in FP32, mixed (TSW_MIXED) and FP64.  The batch engine of each DG degree (staggered kernels
    through layout conversions or on the caller's layout, the ordinate-lane engine, or the lane-split
    staggered kernel on the caller's layout) is the autotuner's
    choice at tsw_batch_set unless forced.  TSW_MIXED runs the FP64 sweep on the engines that work on the
    caller's layout (their FP32 sweep reads and writes the same FP64 arrays and computes no faster), and
    FP32 sweeps with one FP64 refinement on converted members.  value = unknowns/s of the faster of the
    mixed and FP64 calls (both meet 1e-8 against the FP64 sweep; the errors are reported).  Returns the
    JSON line (dict)."""
    steps = args.steps if steps is None else steps
    warmup = args.warmup if warmup is None else warmup
    cpu_sweeps = args.cpu_sweeps if cpu_sweeps is None else cpu_sweeps
    import torch
    import tsw
    from tsw import problems
    batch = problems.batched(seed=2026)
    dev = torch.device("cuda:0")
    srcs = [P.source_device(dev) for P in batch]
    sig = [torch.from_numpy(P.sigma_t).to(dev) for P in batch]
    ctx = tsw.Context(0)
    ctx.set_tuning("prefetch", args.prefetch)
    ctx.set_tuning("batch_direct", args.batch_direct)
    ctx.set_tuning("lanes", args.lanes)
    ctx.set_tuning("batch_ref", args.batch_ref)
    ctx.set_tuning("batch_split", args.batch_split)
    t0 = time.perf_counter()
    ctx.batch_set([problems.as_batch_member(P) for P in batch], sig)
    t_set = time.perf_counter() - t0
    U = sum(P.unknowns for P in batch)
    ncell = sum(P.nx * P.ny for P in batch)
    names = [n for n in args.batched_precisions.split(",") if n]
    prec_of = {"fp32": tsw.FP32, "mixed": tsw.MIXED, "fp64": tsw.FP64}
    res, outs = {}, {}
    for name in names:
        prec = prec_of[name]
        work = [s.clone() for s in srcs]
        for _ in range(warmup):
            ctx.batch_sweep(work, precision=prec)
        ctx.sweep_timing(reset=True)
        wall = 0.0
        for _ in range(steps):
            for w, s in zip(work, srcs):   # restore the sources (outside the timed call)
                w.copy_(s)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.batch_sweep(work, precision=prec)   # returns when the device work is done
            wall += time.perf_counter() - t0
        n, ms = ctx.sweep_timing(reset=True)
        res[name] = (wall / steps, ms / max(n, 1), n // max(steps, 1))
        outs[name] = work
        print(f"bench: batched {name}: call {wall / steps * 1e3:.3f} ms, {res[name][2]} batched sweeps per call "
              f"of {ms / max(n, 1):.3f} ms kernels each", file=sys.stderr, flush=True)
    tuned = {k: ctx.get_tuning(k) for k in ("batch_direct", "prefetch", "lanes", "batch_ref", "batch_split",
                                             "batch_split_prefetch", "batch_engine_p1", "batch_engine_p2",
                                             "batch_engine_p3", "batch_split_rows_p1", "batch_split_rows_p2",
                                             "batch_split_rows_p3")}
    ctx.close()
    engine_names = {0: "staggered, layout conversions", 1: "staggered, caller's layout", 2: "ordinate lanes",
                    3: "lane-split staggered, caller's layout"}
    engines = {f"Q{p}": engine_names.get(tuned[f"batch_engine_p{p}"], "none") for p in (1, 2, 3)}
    print(f"bench: batched engines {engines} (tsw_batch_set {t_set * 1e3:.0f} ms)",
          file=sys.stderr, flush=True)

    def rel(a_list, b_list):
        num = sum(float((a - b).norm() ** 2) for a, b in zip(a_list, b_list))
        return math.sqrt(num / sum(float(b.norm() ** 2) for b in b_list))

    errs = {n: rel(outs[n], outs["fp64"]) for n in names if n != "fp64"} if "fp64" in outs else {}
    if errs:
        print("bench: batched rel err vs FP64: " + ", ".join(f"{n} {e:.3e}" for n, e in errs.items()),
              file=sys.stderr, flush=True)
    # one batched sweep moves the sources in and psi out of the caller's FP64 arrays (16 B per unknown) and
    # reads one material index per cell
    b_alg = U * 16 + ncell * 4
    accurate = [n for n in ("mixed", "fp64") if n in res]
    best = min(accurate, key=lambda n: res[n][0]) if accurate else names[0]
    cbest = res[best][0]
    line = {
        "metric": "angular unknowns swept/s, mixed vs FP64 sweep at 1e-8 rel err; %HBM roofline",
        "value": U / cbest, "unit": "angular unknowns swept/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": cbest * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if best == "fp64" else "mixed (see config)",
        "data": "synthetic (seeded batch generated in-process)",
        "config": {"workload": "batched sweep: 256 uneven problems, N 32..256, Q1..Q3, S4..S12, 8x8 block "
                               "materials, seed 2026, in place on the caller's FP64 arrays; value = the faster "
                               "of the mixed and FP64 calls",
                   "value_from": best, "engines": engines, "unknowns_per_batch_sweep": U,
                   "batch_set_ms": t_set * 1e3},
        "mixed_rel_err_vs_fp64": errs.get("mixed"),
    }
    for n in names:
        line[n] = {"value": U / res[n][0], "ms_per_step": res[n][0] * 1e3, "sweep_kernels_ms": res[n][1],
                   "kernel_gbs": b_alg / (res[n][1] * 1e-3) / 1e9}
        if n in errs:
            line[n]["rel_err_vs_fp64"] = errs[n]
    if "mixed" in res and "fp64" in res:
        line["speedup_mixed_vs_fp64"] = res["fp64"][0] / res["mixed"][0]
    if "fp64" in res:
        t64 = res["fp64"][1] * 1e-3
        traffic, traffic_as = traffic_from_profiles("batched256 fp64")
        line["roofline"] = {"bound": "hbm", "achieved": b_alg / t64 / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": b_alg / t64 / 1e9 / HBM_PEAK_GBS,
                            "kernel": "one FP64 batched sweep: every engine's launches on four streams, timed "
                                      "from fork to join",
                            "algorithmic_bytes_per_launch": b_alg, "traffic": traffic,
                            "traffic_measured_as": traffic_as, "tuning": tuned}
    if cpu_sweeps > 0:
        mid = sorted(batch, key=lambda P: P.unknowns)[len(batch) // 2]
        line["cpu_baseline"] = cpu_baseline(mid, cpu_sweeps)
    return line


def traffic_from_profiles(family, band_rows=None):
    """HBM bytes per launch of the kernel family `family` (the sweep template without its prefetch depth
    and store variant, which move the same data) from the committed rocprofv3 --pmc summary, with the
    configuration it was measured in; the figure for `band_rows` where the summary has one per band
    height; None when no summary exists."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(family)
    except (OSError, ValueError):
        return None, None
    if not entry:
        return None, None
    entry = entry.get("by_band_rows", {}).get(str(band_rows), entry)
    return entry.get("hbm_bytes_per_launch"), entry.get("measured_as")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--pattern", default="hip-staggered")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--cpu-sweeps", type=int, default=3, help="oracle sweeps timed for cpu_baseline (0: skip)")
    ap.add_argument("--roofline-sweeps", type=int, default=20, help="operator-contract sweeps timed per precision")
    ap.add_argument("--copy-probe", action="store_true", help="run a 1 GiB device copy (PMC calibration)")
    ap.add_argument("--sweeps-only", action="store_true", help="skip the solves (profiling runs)")
    ap.add_argument("--workload", default="lattice", choices=["lattice", "batched", "si", "idc"])
    ap.add_argument("--batched-steps", type=int, default=5,
                    help="lattice workload: timed calls per precision of the 256-problem batch reported in its "
                         "'batched' sub-object (0: skip)")
    ap.add_argument("--idc-steps", type=int, default=10, help="IDC steps timed by --workload idc (100 reach t=1)")
    ap.add_argument("--band-rows", type=int, default=0, help="rows per workgroup band (0 automatic, -1 none)")
    ap.add_argument("--prefetch", type=int, default=0, help="prefetch depth of the staggered sweep (0 default)")
    ap.add_argument("--lane-split", type=int, default=0,
                    help="1: lane-split sweep kernel; 0: the autotuner chooses (with automatic bands and prefetch)")
    ap.add_argument("--batch-direct", type=int, default=0,
                    help="batch groups: -1 storer wave, 0 or 1 direct stores where the tables allow")
    ap.add_argument("--lanes", type=int, default=0,
                    help="batched workload: ordinate-lane engine forced (1), forbidden (-1) or autotuned (0)")
    ap.add_argument("--batched-precisions", default="fp32,mixed,fp64",
                    help="batched workload: the precisions timed (comma-separated fp32, mixed, fp64)")
    ap.add_argument("--batch-ref", type=int, default=0,
                    help="batched workload: staggered kernels on the caller's layout forced (1), forbidden (-1) "
                         "or autotuned (0)")
    ap.add_argument("--batch-split", type=int, default=0,
                    help="batched workload: lane-split kernel on the caller's layout forced (1), forbidden (-1) "
                         "or autotuned (0)")
    ap.add_argument("--dsa", type=int, default=1, help="diffusion synthetic acceleration of the GMRES (1 on)")
    ap.add_argument("--dsa-cycles", type=int, default=1, help="multigrid V-cycles per DSA application")
    ap.add_argument("--band-scan", default="", help="comma-separated band_rows values: time contract sweeps "
                                                     "for each and report them (tuning aid)")
    ap.add_argument("--prefetch-scan", default="", help="comma-separated prefetch depths crossed with --band-scan")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: every box in this setup has a single MI355X, and the sweep has no "
              f"multi-GPU path (asked for --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batched":
        print(json.dumps(bench_batched(args)), flush=True)
        return
    if args.workload == "si":
        bench_si(args)
        return
    if args.workload == "idc":
        bench_idc(args)
        return

    import torch
    import tsw
    from tsw import problems

    P = problems.lattice()
    dev = torch.device("cuda:0")
    d_src = P.source_device(dev)
    d_st = torch.from_numpy(P.sigma_t).to(dev)
    d_ss = torch.from_numpy(P.sigma_s).to(dev)
    work = torch.empty_like(d_src)
    phi64 = torch.zeros((P.nx + 2, P.ny + 2, P.p + 1, P.p + 1), dtype=torch.float64, device=dev)
    phi32 = torch.zeros_like(phi64)
    ctx = tsw.Context(0)
    P.set_on(ctx)
    ctx.set_pattern(args.pattern)
    ctx.set_tuning("band_rows", args.band_rows)
    ctx.set_tuning("prefetch", args.prefetch)
    ctx.set_tuning("lane_split", args.lane_split)

    def solve(precision, phi):
        work.copy_(d_src)
        torch.cuda.synchronize()
        return ctx.gmres_solve(work, d_st, d_ss, phi_out=phi, abs_tol=args.tol, restart=30, precision=precision)

    def timed(precision, phi):
        for _ in range(args.warmup):
            solve(precision, phi)
        ctx.sweep_timing(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        its = []
        for _ in range(args.steps):
            its.append(solve(precision, phi))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        n_sweeps, sweep_ms = ctx.sweep_timing(reset=True)
        return wall, n_sweeps, sweep_ms, its

    print(f"bench: lattice {P.nx}x{P.ny} Q{P.p} S12 nq={P.nq} U={P.unknowns}", file=sys.stderr, flush=True)
    U = P.unknowns
    ncell = P.nx * P.ny

    def contract_sweeps(precision, ref_split=-1, wall=False):
        """Average kernel time (ms) of in-place operator-contract sweeps of the lattice source: the kernel on
        the diagonal layout (ref_split -1, the roofline kernel; the call converts layouts around it) or the
        lane-split kernel on the caller's layout (ref_split 1, the default of tsw_sweep_apply).  With `wall`,
        also the average wall time of the whole call."""
        ctx.set_tuning("ref_split", ref_split)
        ctx.sweep(work, d_st, precision=precision)  # warm-up (builds tables)
        ctx.sweep_timing(reset=True)
        t_call = 0.0
        for _ in range(args.roofline_sweeps):
            work.copy_(d_src)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            ctx.sweep(work, d_st, precision=precision)
            torch.cuda.synchronize()
            t_call += time.perf_counter() - t0
        n, ms = ctx.sweep_timing(reset=True)
        ctx.set_tuning("ref_split", 1)
        k = ms / max(n, 1)
        return (k, t_call / args.roofline_sweeps * 1e3) if wall else k

    if args.band_scan or args.prefetch_scan:
        brs = [int(v) for v in args.band_scan.split(",")] if args.band_scan else [args.band_rows]
        pfs = [int(v) for v in args.prefetch_scan.split(",")] if args.prefetch_scan else [args.prefetch]
        for br in brs:
            for pf in pfs:
                ctx.set_tuning("band_rows", br)
                ctx.set_tuning("prefetch", pf)
                a32, a64 = contract_sweeps(tsw.FP32), contract_sweeps(tsw.FP64)
                print(f"bench: band_rows {br} prefetch {pf}: contract sweep fp32 {a32:.4f} ms fp64 {a64:.4f} ms",
                      file=sys.stderr, flush=True)
        ctx.set_tuning("band_rows", args.band_rows)
        ctx.set_tuning("prefetch", args.prefetch)
    t_sw32 = contract_sweeps(tsw.FP32)
    tuned32 = (ctx.get_tuning("band_rows"), ctx.get_tuning("prefetch"))
    store32 = 2 if ctx.get_tuning("one_wave") else 0  # kernel template argument: storer wave or one wave
    split32 = ctx.get_tuning("lane_split")  # 1: the lane-split kernel
    t_sw64 = contract_sweeps(tsw.FP64)
    tuned64 = (ctx.get_tuning("band_rows"), ctx.get_tuning("prefetch"))
    split64 = ctx.get_tuning("lane_split")
    print(f"bench: sweep tuning (band_rows, prefetch): fp32 {tuned32} fp64 {tuned64}; lane split fp32 {split32} "
          f"fp64 {split64}", file=sys.stderr, flush=True)
    print(f"bench: contract sweep fp32 {t_sw32:.3f} ms fp64 {t_sw64:.3f} ms", file=sys.stderr, flush=True)
    # whole tsw_sweep_apply calls (SweepOperator::Apply) on device arrays: in place on the caller's layout by
    # the lane-split kernel (the default) and through the layout conversions
    apply_calls = {}
    for name, prec in (("fp64", tsw.FP64), ("fp32", tsw.FP32), ("mixed", tsw.MIXED)):
        k1, w1 = contract_sweeps(prec, 1, wall=True)
        k0, w0 = contract_sweeps(prec, -1, wall=True)
        apply_calls[name] = {"call_ms": w1, "kernel_ms": k1, "through_conversions_call_ms": w0}
    print("bench: tsw_sweep_apply calls (caller's layout / through conversions, ms): " +
          ", ".join(f"{n} {v['call_ms']:.3f} / {v['through_conversions_call_ms']:.3f}" for n, v in apply_calls.items()),
          file=sys.stderr, flush=True)
    if args.copy_probe:
        a = torch.empty(1 << 27, dtype=torch.float64, device=dev)  # 1 GiB
        b = torch.empty_like(a)
        a.fill_(1.0)                # writes 1 GiB
        torch.cuda.synchronize()
        torch.add(a, 1.0, out=b)     # reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        s_ = float(b.sum())          # reads 1 GiB
        torch.cuda.synchronize()
        print(f"bench: probes: fill 1 GiB, add 1 GiB -> 1 GiB, sum 1 GiB ({s_:.3e})", file=sys.stderr, flush=True)
    if args.sweeps_only:
        ctx.close()
        return

    ctx.set_tuning("use_dsa", 0)
    w64n, n64n, ms64n, its64n = timed(tsw.FP64, phi64)
    print(f"bench: fp64 without DSA: wall {w64n:.3f}s sweeps {n64n} its {its64n}", file=sys.stderr, flush=True)
    ctx.set_tuning("use_dsa", args.dsa)
    ctx.set_tuning("dsa_cycles", args.dsa_cycles)
    w64, n64, ms64, its64 = timed(tsw.FP64, phi64)
    print(f"bench: fp64 wall {w64:.3f}s sweeps {n64} sweep-ms {ms64:.1f} its {its64}", file=sys.stderr, flush=True)
    w32, n32, ms32, its32 = timed(tsw.FP32, phi32)
    print(f"bench: mixed wall {w32:.3f}s sweeps {n32} sweep-ms {ms32:.1f} its {its32}", file=sys.stderr, flush=True)

    inner = (slice(1, -1), slice(1, -1))
    err = float((phi32[inner] - phi64[inner]).norm() / phi64[inner].norm())
    value = U * n32 / w32
    value64 = U * n64 / w64
    bytes32 = U * 8 + ncell * 4
    bytes64 = U * 16 + ncell * 8
    achieved = bytes32 / (t_sw32 * 1e-3) / 1e9
    achieved64 = bytes64 / (t_sw64 * 1e-3) / 1e9

    family32 = "k_diag_sweep_split<float, 3, 1>" if split32 else "k_diag_sweep<float, 3, true, 1>"
    kernel32 = (f"k_diag_sweep_split<float, 3, 1, {tuned32[1]}>" if split32
                else f"k_diag_sweep<float, 3, true, 1, {tuned32[1]}, {store32}>")
    traffic32, traffic32_as = traffic_from_profiles(family32, tuned32[0])
    line = {
        "metric": "angular unknowns swept/s, mixed vs FP64 sweep at 1e-8 rel err; %HBM roofline",
        "value": value,
        "unit": "angular unknowns swept/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": w32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 sweep / f64 Krylov",
        "data": "synthetic (lattice cross sections and source generated in-process)",
        "config": {"workload": "checkerboard lattice 280x280 cells, Q2 DG, S12 (84 ordinates), Peierls GMRES(30) "
                               "with DSA to abs residual 1e-8; mixed = FP32 sweeps in the Arnoldi steps once the FP64 residual "
                               "estimate is <= 1e-3, FP64 sweeps otherwise, FP64 Krylov vectors and DSA",
                   "sweep_pattern": args.pattern, "band_rows": args.band_rows, "prefetch": args.prefetch,
                   "tuned_band_rows_prefetch": {"fp32": tuned32, "fp64": tuned64},
                   "lane_split": {"fp32": split32, "fp64": split64}, "dsa": bool(args.dsa),
                   "dsa_cycles": args.dsa_cycles,
                   "unknowns_per_sweep": U},
        "fp64": {"value": value64, "ms_per_step": w64 / args.steps * 1e3, "sweeps_per_solve": n64 / args.steps,
                 "gmres_its": [i for i, _ in its64], "solve_sweep_ms": ms64 / max(n64, 1),
                 "contract_sweep_ms": t_sw64, "contract_sweep_gbs": achieved64},
        "fp64_no_dsa": {"value": U * n64n / w64n, "ms_per_step": w64n / args.steps * 1e3,
                        "sweeps_per_solve": n64n / args.steps, "gmres_its": [i for i, _ in its64n]},
        "mixed": {"sweeps_per_solve": n32 / args.steps, "gmres_its": [i for i, _ in its32],
                  "solve_sweep_ms": ms32 / max(n32, 1), "contract_sweep_ms": t_sw32},
        "speedup_mixed_vs_fp64": (w64 / w32),
        "mixed_rel_err_vs_fp64": err,
        "sweep_apply_calls": apply_calls,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "kernel": f"{kernel32} (FP32 in-place operator-contract sweep, band_rows {tuned32[0]})",
                     "algorithmic_bytes_per_launch": bytes32,
                     "traffic": traffic32,
                     "traffic_measured_as": traffic32_as},
    }
    if args.cpu_sweeps > 0:
        ref = cpu_baseline_reference(P.nx, args.cpu_sweeps)
        line["cpu_baseline"] = ref if ref else cpu_baseline(P, args.cpu_sweeps)
    ctx.close()
    del work, d_src, d_st, d_ss, phi64, phi32
    if args.batched_steps > 0:
        # the batched configuration, where the mixed-precision targets are stated: the 256-problem batch
        # timed in FP32, mixed and FP64 (its own sub-object; `value` stays the lattice solve's)
        torch.cuda.empty_cache()
        b = bench_batched(args, steps=args.batched_steps, warmup=1, cpu_sweeps=0)
        line["batched"] = {k: b[k] for k in ("value", "unit", "ms_per_step", "steps", "warmup", "dtype", "config",
                                              "mixed_rel_err_vs_fp64", "fp32", "mixed", "fp64",
                                              "speedup_mixed_vs_fp64", "roofline") if k in b}
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
