"""The oracle and the library's host-side pieces against the reference's own outputs
(tests/golden/ref_fixtures.npz, written by tests/golden/make_ref_fixtures.py from the reference sources
compiled unmodified, oracle/build_ref.sh):

* ordinate sets (tsw_ordinates) equal the reference's OrdinateSet, order included, for T_1..T_8 and
  Chebyshev-Legendre 2..8, full and symmetry-reduced: bit for bit for Chebyshev-Legendre and power-of-two
  T_N (midpoint subdivision), and within 8 ulps (directions) and 1e-13 relative (weights, spherical
  excesses of small triangles) for the row tessellation of other T_N;
* the oracle's sweep equals the reference's SweepOperator::Apply on the same inputs (degree-0 and
  degree-p cross sections, steady and with dt);
* the oracle's source iteration takes the reference's iteration count and reaches its angular flux;
* the numpy deck path builds the cross sections the reference's Init builds;
* the oracle's IDC step (oracle/idc.py over the oracle's GMRES) equals the reference's IDCIntegrator::Step
  for the input deck's configuration (update form, 2 Gauss-Radau nodes, 2 corrections) and the other node
  families and forms (the reference's implicit solves are source iterations to 1e-13).
Where oracle/_ref/libref.so is present the live reference is also run on fresh seeded inputs."""
import math
import os

import numpy as np
import pytest

import tsw
import _ref
from _common import (REPO, Grid, Ordinates, orc_reflect_boundaries, orc_reflect_map, orc_si, orc_sweep,
                     orc_si_reflect, orc_sweep_reflect, rel_err, xs_degree)

FIX = os.path.join(REPO, "tests", "golden", "ref_fixtures.npz")


@pytest.fixture(scope="module")
def fx():
    with np.load(FIX, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def _case(fx, name):
    nx, ny, p, nq, sd = (int(v) for v in fx[name + "/dims"])
    ax, bx, ay, by = (float(v) for v in fx[name + "/box"])
    g = Grid(nx, ny, p, ax, bx, ay, by)
    o = Ordinates(fx[name + "/xi"], fx[name + "/eta"], fx[name + "/w"], fx[name + "/quad"].astype(np.int32))
    st, ss = fx[name + "/sigma_t"], fx[name + "/sigma_s"]
    if sd == 0:
        st, ss = st[..., 0, 0].copy(), ss[..., 0, 0].copy()
    return g, o, sd, st, ss


@pytest.mark.parametrize("kind,orders", [("spherical-triangle", range(1, 9)), ("chebyshev-legendre", (2, 4, 6, 8))])
@pytest.mark.parametrize("reduce", [0, 1])
def test_ordinate_sets_equal_reference(fx, kind, orders, reduce):
    for order in orders:
        key = f"ords/{kind}/{order}/{reduce}"
        xi, eta, w, off = tsw.ordinates(kind, order, bool(reduce))
        np.testing.assert_array_equal(off, fx[key + "/quad"], err_msg=key)
        if kind == "chebyshev-legendre" or order & (order - 1) == 0:
            np.testing.assert_array_equal(xi, fx[key + "/xi"], err_msg=key)
            np.testing.assert_array_equal(eta, fx[key + "/eta"], err_msg=key)
            np.testing.assert_array_equal(w, fx[key + "/w"], err_msg=key)
        else:   # same directions in the same positions, to rounding
            for a, b in ((xi, fx[key + "/xi"]), (eta, fx[key + "/eta"])):
                assert np.all(np.abs(a - b) <= 8 * np.spacing(np.abs(b))), key
            np.testing.assert_allclose(w, fx[key + "/w"], rtol=1e-13, atol=0, err_msg=key)


@pytest.mark.parametrize("name", ["sweep_pwc", "sweep_ho", "sweep_ho3"])
def test_oracle_sweep_equals_reference(fx, name):
    g, o, sd, st, _ = _case(fx, name)
    src = fx[name + "/src"]
    with xs_degree(sd):
        for key, dt in (("steady", math.inf), ("dt03", 0.3)):
            out = orc_sweep(g, o, st, src, dt)
            assert rel_err(out, fx[name + "/" + key]) <= 1e-14, (name, key)


REFLECT = ["reflect_xmin", "reflect_xmax_ymin", "reflect_all", "reflect_ymax_q3", "reflect_xmin_ymax_ho"]


@pytest.mark.parametrize("name", REFLECT)
def test_oracle_reflecting_sweep_equals_reference(fx, name):
    """Reflecting edges: the quadrants one after the other with the ghosts reflected after each, bit for
    bit the reference's SweepOperator::Apply (ghost cells included) with cell-mean cross sections."""
    g, o, sd, st, _ = _case(fx, name)
    src, mask = fx[name + "/src"], int(fx[name + "/mask"])
    with xs_degree(sd):
        for key, dt in (("steady", math.inf), ("dt03", 0.3)):
            out = orc_sweep_reflect(g, o, st, src, mask, dt)
            if sd == 0:
                np.testing.assert_array_equal(out, fx[name + "/" + key], err_msg=f"{name} {key}")
            else:  # degree-p cross sections: the sweep itself agrees to rounding (test_oracle_sweep_equals_reference)
                assert rel_err(out, fx[name + "/" + key]) <= 1e-14, (name, key)


def test_oracle_reflecting_source_iteration_equals_reference(fx):
    """SISolver::Solve on a lattice deck reflecting on three edges: the reference's iteration count and
    last change, its angular flux bit for bit."""
    name = "si_reflect"
    g, o, _, st, ss = _case(fx, name)
    psi, its, err = orc_si_reflect(g, o, st, ss, fx[name + "/src"], int(fx[name + "/mask"]),
                                   abs_tol=float(fx[name + "/tol"]), max_its=5000)
    assert its == int(fx[name + "/its"])
    assert float(f"{err:.4e}") == float(fx[name + "/err"])
    np.testing.assert_array_equal(psi, fx[name + "/psi"])


def test_oracle_periodic_source_iteration_equals_reference(fx):
    """SISolver::Solve on a lattice deck with periodic boundaries (halos synchronized from the opposite side
    before each iteration's source copy, which then skips them): iteration count, last change and angular
    flux bit for bit."""
    name = "si_periodic"
    g, o, _, st, ss = _case(fx, name)
    psi, its, err = orc_si_reflect(g, o, st, ss, fx[name + "/src"], 0, abs_tol=float(fx[name + "/tol"]),
                                   max_its=5000, periodic=True)
    assert its == int(fx[name + "/its"])
    assert float(f"{err:.4e}") == float(fx[name + "/err"])
    np.testing.assert_array_equal(psi, fx[name + "/psi"])


def test_oracle_reflect_map_and_ghosts_closed_forms(fx):
    """The mirror map flips one direction cosine and is an involution; reflected ghosts of a field that is
    itself mirror-symmetric reproduce the field's own halo values."""
    g, o, _, _, _ = _case(fx, "reflect_all")
    m = orc_reflect_map(o).reshape(2, o.nq)
    for dim, (a, b) in enumerate(((o.xi, o.eta), (o.eta, o.xi))):
        np.testing.assert_allclose(a[m[dim]], -a, rtol=0, atol=1e-14)  # the reference's matching tolerance
        np.testing.assert_allclose(b[m[dim]], b, rtol=0, atol=1e-14)
        np.testing.assert_array_equal(m[dim][m[dim]], np.arange(o.nq))
    # an unreduced set has +-mu pairs with equal (xi, eta): no unique map from (xi, eta)
    xi, eta, w, off = tsw.ordinates("chebyshev-legendre", 4, False)
    with pytest.raises(ValueError):
        orc_reflect_map(Ordinates(xi, eta, w, off))
    # psi(q, i, j) = f(i) with f(0) = f(1) (x-min mirror plane), coefficient d odd in x zero: x-min ghosts
    # reflect onto themselves
    rng = np.random.default_rng(3)
    psi = np.zeros(g.psi_shape(o.nq))
    col = rng.uniform(size=(g.ny + 2, o.nq))
    col = 0.5 * (col + col[:, m[0]])  # symmetric under the x mirror
    psi[:, :, :, 0, :] = col[None, :, :, None]
    out = orc_reflect_boundaries(g, o, psi, 1, m.ravel())
    np.testing.assert_array_equal(out[0, 1:-1], psi[0, 1:-1])


@pytest.mark.parametrize("name", ["si_pwc", "si_ho"])
def test_oracle_source_iteration_equals_reference(fx, name):
    g, o, sd, st, ss = _case(fx, name)
    with xs_degree(sd):
        psi, its, err = orc_si(g, o, st, ss, fx[name + "/src"], abs_tol=float(fx[name + "/tol"]), max_its=5000)
    assert its == int(fx[name + "/its"])
    assert abs(err - float(fx[name + "/err"])) <= 1e-4 * float(fx[name + "/err"])   # printed with 4 decimals
    assert rel_err(psi, fx[name + "/psi"]) <= 1e-12


IDC_CASES = ["idc_radau_update", "idc_lobatto_error", "idc_legendre_update", "idc_chebradau_error_colloc",
             "idc_chebyshev_update"]
IDC_KINDS = {"gauss-radau": "radau", "gauss-lobatto": "lobatto", "gauss-legendre": "legendre",
             "chebyshev-radau": "chebyshev-radau", "chebyshev": "chebyshev"}


@pytest.mark.parametrize("name", IDC_CASES)
def test_oracle_idc_step_equals_reference(fx, name):
    import idc
    from _common import OracleIdcOps
    g, o, _, st, ss = _case(fx, name)
    update, nn, nc, colloc = (int(v) for v in fx[name + "/config"])
    ops = OracleIdcOps(g, o, st, ss, tol=1e-14)
    out = idc.idc_step(ops, fx[name + "/psi0"], fx[name + "/src"], float(fx[name + "/dt"]), num_nodes=nn,
                       num_corr=nc, kind=IDC_KINDS[str(fx[name + "/nodes"])], collocation=bool(colloc),
                       form="update" if update else "error")
    assert rel_err(out, fx[name + "/psi1"]) <= 1e-12, name


def test_deck_cross_sections_equal_reference_init(fx):
    """The numpy Init (tsw/deck.py) projects the lattice cross sections as the reference's Init does, at
    degree 0 and at the flux degree."""
    from tsw import deck
    for name in ("sweep_pwc", "sweep_ho", "si_ho"):
        nx, ny, p, nq, sd = (int(v) for v in fx[name + "/dims"])
        text = ("ax 0.0\nbx 7.0\nay 0.0\nby 7.0\n" f"nx {nx}\nny {ny}\ndg_degree {p}\n"
                f"piecewise_cross {'true' if sd == 0 else 'false'}\nproblem lattice\nmoll_radius 0.0\n"
                "ordinate_type spherical-triangle\nang_order 2\nordinate_sym_reduce true\n")
        P, ex = deck.init(dict(line.split() for line in text.strip().split("\n")))
        st = ex[f"sigma_t_dg{sd}_nx{nx:07d}.cs2"]
        ss = ex[f"sigma_s_dg{sd}_nx{nx:07d}.cs2"]
        np.testing.assert_allclose(st, fx[name + "/sigma_t"], rtol=0, atol=1e-13, err_msg=name)
        np.testing.assert_allclose(ss, fx[name + "/sigma_s"], rtol=0, atol=1e-13, err_msg=name)


@pytest.mark.skipif(not _ref.available(), reason="oracle/_ref/libref.so is built only where the reference is")
def test_live_reference_sweeps_and_si(tmp_path):
    """Fresh inputs through the compiled reference and the oracle: sweeps with random sources and random
    degree-p cross sections, and a scattering solve."""
    from _common import random_poly_sigma
    deck = ("ax 0.0\nbx 1.0\nay 0.0\nby 1.5\nnx 7\nny 6\ndg_degree {p}\npiecewise_cross {pwc}\nproblem lattice\n"
            "moll_radius 0.0\nordinate_type chebyshev-legendre\nang_order 4\nordinate_sym_reduce true\n"
            "solve_type source-iteration\nsweep_pattern wavefront\nsweep_solve gaussian-elimination\n"
            "abs_tol 1e-13\nmax_its 5000\ntfinal 1.0\noutput_log false\n")
    for p, pwc in ((1, "true"), (2, "false"), (3, "false")):
        path = tmp_path / f"d{p}.deck"
        path.write_text(deck.format(p=p, pwc=pwc))
        d = _ref.Deck(path)
        g = Grid(d.nx, d.ny, d.p, *d.box)
        o = Ordinates(d.xi, d.eta, d.w, d.quad.astype(np.int32))
        rng = np.random.default_rng(100 + p)
        sd = d.xs_degree
        st = random_poly_sigma(g, rng, rel=0.5) if sd else rng.uniform(0.3, 3.0, size=(d.nx + 2, d.ny + 2))
        ss = st * 0.5
        d.set_xs(0, st if sd else st[..., None, None])
        d.set_xs(1, ss if sd else ss[..., None, None])
        src = rng.uniform(-0.5, 1.0, size=d.psi_shape())
        with xs_degree(sd):
            for dt in (math.inf, 0.2):
                assert rel_err(orc_sweep(g, o, st, src, dt), d.sweep(src, dt)) <= 1e-14, (p, dt)
            psi, its, err = orc_si(g, o, st, ss, src, abs_tol=1e-13, max_its=5000)
        ref_psi, ref_its, _ = d.solve(src, tmp_path / f"si{p}.log")
        assert its == ref_its and rel_err(psi, ref_psi) <= 1e-12, p
