"""The checkerboard-lattice solve at full size on the GPU: 280x280 cells on [0,7]^2, Q2, S12 (84 ordinates),
absorber blocks as in LatticeProblem.cpp:231-255, q = 1 in the centre block, vacuum; sweep-preconditioned
Peierls GMRES(30) (PETScPeierlsSolver.cpp:423-572) to the absolute residual 1e-8.

- FP64 sweeps with DSA, mixed precision (FP32 sweeps in the relaxed Arnoldi steps) with DSA, and FP64
  without DSA reach the same scalar flux (within 1e-6 relative: each stops at its own residual 1e-8);
- the reported residuals are <= 1e-8, and for the mixed solve the reported residual is the FP64 one;
- the returned psi is the transport solution of the returned phi: sum_q w_q psi_q equals phi to the
  solve tolerance (psi = L^{-1}(Q + S phi), so P psi - phi is the unpreconditioned Peierls residual)."""
import pytest

import tsw
from tsw import problems

pytestmark = pytest.mark.gpu


def test_lattice_gmres_dsa_full_size_fp64_mixed_no_dsa():
    import torch
    dev = torch.device("cuda:0")
    P = problems.lattice()
    src = P.source_device(dev)
    st = torch.from_numpy(P.sigma_t).to(dev)
    ss = torch.from_numpy(P.sigma_s).to(dev)
    w = torch.from_numpy(P.w).to(dev)
    inner = (slice(1, -1), slice(1, -1))
    runs = {}
    with tsw.Context(0) as c:
        P.set_on(c)
        for name, prec, dsa in (("fp64", tsw.FP64, 1), ("mixed", tsw.FP32, 1), ("fp64_no_dsa", tsw.FP64, 0)):
            print(f"lattice 280x280 Q2 S12 GMRES(30) to 1e-8: {name}", flush=True)
            c.set_tuning("use_dsa", dsa)
            psi = src.clone()
            phi = torch.zeros((P.nx + 2, P.ny + 2, P.p + 1, P.p + 1), dtype=torch.float64, device=dev)
            its, res = c.gmres_solve(psi, st, ss, phi_out=phi, abs_tol=1e-8, restart=30, precision=prec)
            assert res <= 1e-8, (name, its, res)
            pphi = torch.einsum("ijqde,q->ijde", psi, w)
            fixed = float((pphi[inner] - phi[inner]).norm() / phi[inner].norm())
            assert fixed <= 1e-9, (name, fixed)
            runs[name] = (phi, its)
            del psi
    ref = runs["fp64"][0][inner]
    for name in ("mixed", "fp64_no_dsa"):
        err = float((runs[name][0][inner] - ref).norm() / ref.norm())
        assert err <= 1e-6, (name, err)
    assert runs["fp64"][1] <= runs["fp64_no_dsa"][1]
