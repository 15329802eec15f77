"""The HIP library against the reference's own outputs (tests/golden/ref_fixtures.npz, written from the
reference sources compiled unmodified): sweeps on every pattern with degree-0 and degree-p cross sections
(FP64 to 1e-12, FP32 to 2e-6), source iteration (the reference's iteration count, its angular flux to
1e-11) on lattice decks with both kinds of cross section, and IDC steps (IDCIntegrator::Step) for the input
deck's configuration (update form, 2 Gauss-Radau nodes, 2 corrections) and the other node families and
forms (FP64 to 1e-11: the reference's implicit solves stop at 1e-13)."""
import math
import os

import numpy as np
import pytest

import tsw
from _common import REPO, rel_err

pytestmark = pytest.mark.gpu
FIX = os.path.join(REPO, "tests", "golden", "ref_fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _ctx(fx, name, pattern="hip-staggered"):
    nx, ny, p, nq, sd = (int(v) for v in fx[name + "/dims"])
    ax, bx, ay, by = (float(v) for v in fx[name + "/box"])
    c = tsw.Context(0)
    c.problem_set(nx, ny, p, fx[name + "/xi"], fx[name + "/eta"], fx[name + "/w"], fx[name + "/quad"], ax, bx, ay, by)
    c.set_pattern(pattern)
    st, ss = fx[name + "/sigma_t"], fx[name + "/sigma_s"]
    if sd:
        c.set_xs_degree(sd)
    else:
        st, ss = st[..., 0, 0].copy(), ss[..., 0, 0].copy()
    return c, st, ss


@pytest.mark.parametrize("name", ["sweep_pwc", "sweep_ho", "sweep_ho3"])
@pytest.mark.parametrize("pattern", ["hip-staggered", "hip-staggered-ge", "hip-wavefront"])
def test_sweeps_match_reference(fx, name, pattern):
    print(f"reference fixture {name} on {pattern}", flush=True)
    c, st, _ = _ctx(fx, name, pattern)
    with c:
        for key, dt in (("steady", math.inf), ("dt03", 0.3)):
            for prec, tol in ((tsw.FP64, 1e-12), (tsw.FP32, 2e-6), (tsw.MIXED, 1e-10)):
                out = fx[name + "/src"].copy()
                c.sweep(out, st, dt, precision=prec)
                assert rel_err(out, fx[name + "/" + key]) <= tol, (key, prec)


@pytest.mark.parametrize("name", ["si_pwc", "si_ho"])
@pytest.mark.parametrize("graph", [1, 0])
def test_source_iteration_matches_reference(fx, name, graph):
    print(f"reference fixture {name} SI graph={graph}", flush=True)
    c, st, ss = _ctx(fx, name)
    with c:
        c.set_tuning("si_graph", graph)
        out = fx[name + "/src"].copy()
        its, err = c.si_solve(out, st, ss, abs_tol=float(fx[name + "/tol"]), max_its=5000)
    assert its == int(fx[name + "/its"])
    assert abs(err - float(fx[name + "/err"])) <= 1e-4 * float(fx[name + "/err"])
    assert rel_err(out, fx[name + "/psi"]) <= 1e-11


IDC_NODES = {"gauss-radau": tsw.IDC_RADAU, "gauss-lobatto": tsw.IDC_LOBATTO, "gauss-legendre": tsw.IDC_LEGENDRE,
             "chebyshev-radau": tsw.IDC_CHEBYSHEV_RADAU, "chebyshev": tsw.IDC_CHEBYSHEV}


@pytest.mark.parametrize("name", ["idc_radau_update", "idc_lobatto_error", "idc_legendre_update",
                                  "idc_chebradau_error_colloc", "idc_chebyshev_update"])
def test_idc_step_matches_reference(fx, name):
    print(f"reference fixture {name}", flush=True)
    update, nn, nc, colloc = (int(v) for v in fx[name + "/config"])
    c, st, ss = _ctx(fx, name)
    with c:
        c.idc_set(nn, nc, IDC_NODES[str(fx[name + "/nodes"])], bool(colloc),
                  form=tsw.IDC_UPDATE if update else tsw.IDC_ERROR)
        for prec, tol in ((tsw.FP64, 1e-11), (tsw.FP32, 1e-9)):
            out = fx[name + "/psi0"].copy()
            c.idc_step(out, fx[name + "/src"], st, ss, float(fx[name + "/dt"]), abs_tol=1e-13, precision=prec)
            assert rel_err(out, fx[name + "/psi1"]) <= tol, (name, prec)
