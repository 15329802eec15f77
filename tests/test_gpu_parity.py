"""GPU parity: the HIP path behind the C-ABI against the CPU oracle on the same seeded inputs.

Bars: FP64 operator results within 1e-12 relative of the oracle; FP32 sweeps within 2e-6 relative of
the FP64 oracle sweep; converged FP64 solves within the solver tolerance (1e-11 relative at
abs_tol 1e-13); mixed-precision GMRES (FP32 sweeps, FP64 refinement to 1e-10 absolute) within 1e-8."""
import math

import numpy as np
import pytest

import tsw
from _common import (Grid, ls_ordinates, orc_gmres, orc_lmv, orc_pmv, orc_si, orc_smv, orc_sweep,
                     random_ordinates, random_problem, rel_err)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = tsw.Context(0)
    yield c
    c.close()


def setup(ctx, g, o):
    ctx.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
    ctx.set_pattern("hip-wavefront")


PATTERNS = ["hip-wavefront", "hip-wavefront-eager", "hip-staggered", "hip-staggered-ge"]


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("dt", [math.inf, 0.1])
def test_sweep_fp64_matches_oracle(ctx, pattern, p, dt):
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(7, 5, p, o, seed=10 + p, ax=-1.0, bx=2.5, ay=0.0, by=1.0)
    print(f"sweep fp64 p={p} pattern={pattern} dt={dt}", flush=True)
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    ref = orc_sweep(g, o, st, src, dt)
    out = src.copy()
    ctx.sweep(out, st, dt)
    assert rel_err(out, ref) <= 1e-12
    np.testing.assert_array_equal(out[0], src[0])       # halos untouched
    np.testing.assert_array_equal(out[:, -1], src[:, -1])


@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32])
@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("materials", [1, 3, 64])
def test_sweep_material_table_matches_oracle(ctx, p, materials, precision):
    """Piecewise-constant sigma_t with few distinct values: hip-staggered multiplies by precomputed
    element-matrix inverses; results must equal the oracle's elimination."""
    o = ls_ordinates(8)
    g, st, ss, src = random_problem(11, 13, p, o, seed=70 + p, materials=materials)
    setup(ctx, g, o)
    ctx.set_pattern("hip-staggered")
    for dt in (math.inf, 0.2):
        ref = orc_sweep(g, o, st, src, dt)
        out = src.copy()
        ctx.sweep(out, st, dt, precision=precision)
        assert rel_err(out, ref) <= (1e-12 if precision == tsw.FP64 else 2e-6)


@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32, tsw.MIXED])
def test_sweep_direct_store_layout(ctx, precision):
    """Q3 with 64 materials and 200 rows in one band per ordinate: the FP64 inverse table (128 KB) fits in
    LDS only without the storer's staging buffer, so the direct-store kernel runs; against the oracle."""
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(40, 200, 3, o, seed=70, materials=64)
    setup(ctx, g, o)
    ctx.set_pattern("hip-staggered")
    ctx.set_tuning("band_rows", -1)
    try:
        ref = orc_sweep(g, o, st, src, 0.5)
        out = src.copy()
        ctx.sweep(out, st, 0.5, precision=precision)
        assert ctx.get_tuning("band_rows") == 0
    finally:
        ctx.set_tuning("band_rows", 0)
    e = rel_err(out, ref)
    print(f"direct-store layout precision {precision}: rel err {e:.3e}", flush=True)
    assert e <= {tsw.FP64: 1e-12, tsw.FP32: 2e-6, tsw.MIXED: 1e-10}[precision]


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("shape", [(1, 1), (1, 9), (9, 1), (12, 12), (3, 70), (70, 3)])
def test_sweep_degenerate_shapes(ctx, shape, pattern):
    o = random_ordinates(3, np.random.default_rng(1))
    g, st, ss, src = random_problem(shape[0], shape[1], 1, o, seed=3)
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    ref = orc_sweep(g, o, st, src)
    out = src.copy()
    ctx.sweep(out, st)
    assert rel_err(out, ref) <= 1e-12


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_sweep_fp32_within_tolerance(ctx, p, pattern):
    o = ls_ordinates(8)
    g, st, ss, src = random_problem(9, 8, p, o, seed=20 + p)
    src[1:-1, 1:-1] = np.abs(src[1:-1, 1:-1])
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    ref = orc_sweep(g, o, st, src)
    out = src.copy()
    ctx.sweep(out, st, precision=tsw.FP32)
    assert rel_err(out, ref) <= 2e-6


@pytest.mark.parametrize("pattern", PATTERNS)
@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("dt", [math.inf, 0.1])
def test_sweep_mixed_refined_to_fp64(ctx, p, pattern, dt):
    """TSW_MIXED: FP32 sweeps with one FP64 refinement step reach FP64 accuracy (the FP32 error squared),
    with inflow halos and finite dt; two steps do at least as well."""
    o = ls_ordinates(8)
    g, st, ss, src = random_problem(9, 8, p, o, seed=30 + p)
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    ref = orc_sweep(g, o, st, src, dt)
    errs = []
    for steps in (1, 2):
        ctx.set_tuning("refine_steps", steps)
        assert ctx.get_tuning("refine_steps") == steps
        out = src.copy()
        ctx.sweep(out, st, dt, precision=tsw.MIXED)
        np.testing.assert_array_equal(out[0], src[0])
        errs.append(rel_err(out, ref))
    ctx.set_tuning("refine_steps", 1)
    print(f"mixed sweep p={p} {pattern} dt={dt}: rel err {errs}", flush=True)
    assert errs[0] <= 1e-10 and errs[1] <= 1e-12
    with pytest.raises(tsw.TswError):
        ctx.set_tuning("refine_steps", 0)


def test_sweep_device_memory_equals_host(ctx):
    import torch
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(10, 6, 2, o, seed=5)
    setup(ctx, g, o)
    host = src.copy()
    ctx.sweep(host, st)
    d_psi = torch.from_numpy(src.copy()).cuda()
    d_st = torch.from_numpy(st.copy()).cuda()
    torch.cuda.synchronize()
    ctx.sweep(d_psi, d_st)
    np.testing.assert_array_equal(d_psi.cpu().numpy(), host)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_pmv_smv_lmv_match_oracle(ctx, p):
    o = ls_ordinates(4)
    g, st, ss, src = random_problem(6, 7, p, o, seed=30 + p)
    setup(ctx, g, o)
    rng = np.random.default_rng(p)
    phi0 = rng.normal(size=g.phi_shape())
    for all_cells in (0, 1):
        ref = orc_pmv(g, o, src, phi0, alpha=0.7, beta=-1.3, all_cells=all_cells)
        out = ctx.pmv(src, phi0.copy(), alpha=0.7, beta=-1.3, all_cells=all_cells)
        assert rel_err(out, ref) <= 1e-13
        ref = orc_smv(g, o, ss, phi0, src, alpha=1.1, beta=0.4, all_cells=all_cells)
        out = ctx.smv(ss, phi0, src.copy(), alpha=1.1, beta=0.4, all_cells=all_cells)
        assert rel_err(out, ref) <= 1e-13
    y0 = rng.normal(size=src.shape)
    ref = orc_lmv(g, o, st, src, y0, alpha=0.9, beta=0.5)
    out = ctx.lmv(st, src, y0.copy(), alpha=0.9, beta=0.5)
    assert rel_err(out, ref) <= 1e-12


@pytest.mark.parametrize("pattern", ["hip-wavefront", "hip-staggered"])
@pytest.mark.parametrize("precision", [tsw.FP64, tsw.FP32])
def test_source_iteration_matches_oracle(ctx, precision, pattern):
    o = ls_ordinates(8)
    g = Grid(16, 16, 1)
    st = np.ones((18, 18))
    ss = np.full((18, 18), 0.5)
    src = np.zeros(g.psi_shape(o.nq))
    src[1 + 6:1 + 10, 1 + 6:1 + 10, :, 0, 0] = 1.0 / (4 * math.pi)  # central block source, vacuum
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    ref, it_ref, err_ref = orc_si(g, o, st, ss, src, abs_tol=1e-11, max_its=200)
    out = src.copy()
    it, err = ctx.si_solve(out, st, ss, abs_tol=1e-11, max_its=200, precision=precision)
    if precision == tsw.FP64:
        assert it == it_ref
        assert rel_err(out, ref) <= 1e-12
    else:
        assert rel_err(out, ref) <= 1e-5


@pytest.mark.parametrize("materials", [None, 2])
@pytest.mark.parametrize("pattern", ["hip-wavefront", "hip-staggered"])
@pytest.mark.parametrize("p", [1, 2])
def test_gmres_fp64_matches_oracle(ctx, p, pattern, materials):
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(10, 9, p, o, seed=40 + p, materials=materials)
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    rc, ref_psi, ref_phi, its_ref, res_ref = orc_gmres(g, o, st, ss, src, abs_tol=1e-13, restart=30)
    assert rc == 0
    out = src.copy()
    phi = np.zeros(g.phi_shape())
    its, res = ctx.gmres_solve(out, st, ss, phi_out=phi, abs_tol=1e-13, restart=30)
    assert res <= 1e-13
    assert rel_err(phi, ref_phi) <= 1e-11
    assert rel_err(out, ref_psi) <= 1e-11


@pytest.mark.parametrize("dt", [math.inf, 0.1])
@pytest.mark.parametrize("materials", [None, 3])
@pytest.mark.parametrize("pattern", ["hip-wavefront", "hip-staggered"])
@pytest.mark.parametrize("refine", [0, 1])
def test_gmres_mixed_precision_refinement(ctx, pattern, materials, dt, refine):
    """Mixed-precision GMRES: the relaxed strategy (FP32 sweeps in the late Arnoldi steps) and FP32 sweeps
    inside FP64 refinement (gmres_refine; with finite dt its residual must carry 1/dt)."""
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(12, 12, 2, o, seed=50, materials=materials)
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    ctx.set_tuning("gmres_refine", refine)
    assert ctx.get_tuning("gmres_refine") == refine
    rc, ref_psi, ref_phi, _, _ = orc_gmres(g, o, st, ss, src, dt=dt, abs_tol=1e-13, restart=30)
    out = src.copy()
    phi = np.zeros(g.phi_shape())
    its, res = ctx.gmres_solve(out, st, ss, phi_out=phi, dt=dt, abs_tol=1e-10, restart=30, precision=tsw.FP32)
    ctx.set_tuning("gmres_refine", 0)
    print(f"mixed gmres {pattern} refine={refine}: its {its} res {res:.3e} phi err {rel_err(phi, ref_phi):.3e}",
          flush=True)
    assert res <= 1e-10
    assert rel_err(phi, ref_phi) <= 1e-8
    assert rel_err(out, ref_psi) <= 1e-8


def test_errors_are_reported(ctx):
    o = ls_ordinates(4)
    with pytest.raises(tsw.TswError) as e:
        ctx.problem_set(4, 4, 4, o.xi, o.eta, o.w, o.quad_off)
    assert e.value.code == tsw.TSW_E_UNSUPPORTED
    bad = o.quad_off.copy()
    bad[1], bad[2] = bad[2], bad[1]
    with pytest.raises(tsw.TswError) as e:
        ctx.problem_set(4, 4, 1, o.xi, o.eta, o.w, bad)
    assert e.value.code == tsw.TSW_E_ARG
    with pytest.raises(tsw.TswError) as e:
        ctx.set_pattern("no-such-pattern")
    assert e.value.code == tsw.TSW_E_ARG
    fresh = tsw.Context(0)
    psi = np.zeros((3, 3, o.nq, 2, 2))
    with pytest.raises(tsw.TswError) as e:
        fresh.sweep(psi, np.ones((3, 3)))
    assert e.value.code == tsw.TSW_E_STATE
    fresh.close()


@pytest.mark.parametrize("pattern", ["hip-wavefront", "hip-staggered", "hip-staggered-ge"])
def test_full_size_lattice_sweep_properties(ctx, pattern):
    """280x280, Q2, S12 (the lattice workload's sweep): L (L^{-1} q) = q with the GPU Lmv, FP64 and FP32,
    and agreement of the FP32 and FP64 sweeps."""
    import torch
    from tsw import problems
    lat = problems.lattice()
    o = ls_ordinates(12)
    g, _, ss, src = random_problem(280, 280, 2, o, seed=60, bx=7.0, by=7.0)
    st = lat.sigma_t  # the lattice's two materials (table path for hip-staggered)
    setup(ctx, g, o)
    ctx.set_pattern(pattern)
    d_src = torch.from_numpy(src).cuda()
    d_st = torch.from_numpy(st).cuda()
    psi = d_src.clone()
    torch.cuda.synchronize()
    ctx.sweep(psi, d_st)
    back = torch.zeros_like(psi)
    ctx.lmv(d_st, psi, back)
    inner = (slice(1, -1), slice(1, -1))
    r = (back[inner] - d_src[inner]).norm() / d_src[inner].norm()
    assert float(r) <= 1e-12
    psi32 = d_src.clone()
    torch.cuda.synchronize()
    ctx.sweep(psi32, d_st, precision=tsw.FP32)
    r32 = (psi32[inner] - psi[inner]).norm() / psi[inner].norm()
    assert float(r32) <= 3e-6
    psim = d_src.clone()
    torch.cuda.synchronize()
    ctx.sweep(psim, d_st, precision=tsw.MIXED)
    rm = (psim[inner] - psi[inner]).norm() / psi[inner].norm()
    print(f"{pattern}: FP32 rel err {float(r32):.3e}, mixed {float(rm):.3e}", flush=True)
    assert float(rm) <= 1e-10
