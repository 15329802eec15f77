"""The relaxed mixed-precision GMRES's admission factor (tuning key fp32_safety): every setting meets the
solve's bar against the oracle, the final FP64 residual check guarding the result."""
import math

import numpy as np
import pytest

import tsw
from _common import ls_ordinates, orc_gmres, random_problem, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("safety", [1, 10, 1000])
def test_relaxed_gmres_safety_factor_meets_bar(safety):
    o = ls_ordinates(6)
    g, st, ss, src = random_problem(12, 12, 2, o, seed=52, materials=3)
    rc, ref_psi, ref_phi, _, _ = orc_gmres(g, o, st, ss, src, abs_tol=1e-13, restart=30)
    with tsw.Context(0) as c:
        c.problem_set(g.nx, g.ny, g.p, o.xi, o.eta, o.w, o.quad_off, g.ax, g.bx, g.ay, g.by)
        c.set_pattern("hip-staggered")
        c.set_tuning("fp32_safety", safety)
        assert c.get_tuning("fp32_safety") == safety
        for bad in (0, 1001):
            with pytest.raises(tsw.TswError):
                c.set_tuning("fp32_safety", bad)
        out = src.copy()
        phi = np.zeros(g.phi_shape())
        its, res = c.gmres_solve(out, st, ss, phi_out=phi, abs_tol=1e-10, restart=30, precision=tsw.FP32)
    print("fp32_safety", safety, "its", its, "res", res, flush=True)
    assert res <= 1e-10
    assert rel_err(phi, ref_phi) <= 1e-8
    assert rel_err(out, ref_psi) <= 1e-8
