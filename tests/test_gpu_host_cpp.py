"""The C++ host layer (SweepOperator / ImplicitSolver factory over the C-ABI) driven by a C++ program,
its outputs compared with the oracle on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

from _common import REPO, Grid, Ordinates, orc_si, orc_sweep, rel_err

pytestmark = pytest.mark.gpu


def test_host_cpp_driver_matches_oracle(tmp_path):
    exe = os.path.join(REPO, "tests", "cpp", "host_api_test")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    print("running tests/cpp/host_api_test", flush=True)
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    import tsw
    xi, eta, w, off = tsw.level_symmetric(6)
    o = Ordinates(xi, eta, w, off)
    g = Grid(12, 10, 2, 0.0, 1.5, 0.0, 1.0)
    load = lambda name, shape: np.fromfile(os.path.join(tmp_path, name), dtype=np.float64).reshape(shape)
    st = load("sigma_t.bin", (14, 12))
    ss = load("sigma_s.bin", (14, 12))
    src = load("source.bin", g.psi_shape(o.nq))
    assert rel_err(load("sweep.bin", src.shape), orc_sweep(g, o, st, src)) <= 1e-12
    ref_si, _, _ = orc_si(g, o, st, ss, src, abs_tol=1e-13, max_its=1000)
    assert rel_err(load("si.bin", src.shape), ref_si) <= 1e-12
    assert rel_err(load("gmres.bin", src.shape), ref_si) <= 1e-10
    assert rel_err(load("gmres_dsa.bin", src.shape), ref_si) <= 1e-10
    # two IDC steps of the C++ IDCIntegrator against the oracle's restatement
    import idc
    from _common import OracleIdcOps
    psi0 = src.copy()
    src0 = src.copy()
    OracleIdcOps.zero_halo(psi0)
    OracleIdcOps.zero_halo(src0)
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    ref = idc.idc_step(ops, idc.idc_step(ops, psi0, src0, 0.05), src0, 0.05)
    assert rel_err(load("idc.bin", src.shape), ref) <= 1e-10
