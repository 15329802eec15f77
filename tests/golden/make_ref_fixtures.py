#!/usr/bin/env python3
"""Writes tests/golden/ref_fixtures.npz from the reference's own code (oracle/_ref/libref.so, built by
oracle/build_ref.sh from the unmodified reference sources; run in the container that holds the reference):

* ordinate sets: OrdinateSet(order, reduce, type) for spherical-triangle T_1..T_8 and chebyshev-legendre
  orders 2..8 (even), full and symmetry-reduced: xi, eta, w and the quadrant starts, in the reference's
  order (SphericalTriangle.cpp:180-351, OrdinateSet.cpp:441-520,1065-1190);
* sweeps: SweepOperator::Apply (sweep_pattern wavefront, sweep_solve gaussian-elimination) on seeded random
  sources with inflow halos, for decks whose cross sections Init projects at degree 0 and at the flux
  degree (piecewise_cross false), steady and with dt;
* reflecting boundaries: SweepOperator::Apply on decks with reflect_x_min / _x_max / _y_min / _y_max
  (quadrants in sequence with OrdinateFlux::ReflectBoundaries after each, SweepPatternUniCartWavefront.cpp:
  180-245), the mask (1 x-min, 2 x-max, 4 y-min, 8 y-max) stored beside them;
* source iteration with reflecting edges: SISolver::Solve on a lattice deck reflecting on three edges, and
  with periodic boundaries (deck key periodic);
* source iteration: SISolver::Solve on two lattice decks (steady, abs_tol 1e-10 and 1e-12): the angular
  flux, the iteration count and the last change (printed with four decimals);
* IDC steps: IDCIntegrator::Step on a lattice deck from a seeded initial condition, for the input deck's
  configuration (idc_type update, gauss-radau, 2 nodes, 2 corrections) and the other node families and
  forms, each implicit solve a source iteration to 1e-13.
Each case also stores the inputs the reference's Init built (cross-section coefficients, ordinates, mesh),
so the tests need only the fixture.  Usage: python tests/golden/make_ref_fixtures.py
"""
import math
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import _ref  # noqa: E402

LATTICE = ("ax 0.0\nbx 7.0\nay 0.0\nby 7.0\nnx {n}\nny {m}\ndg_degree {p}\npiecewise_cross {pwc}\nproblem lattice\n"
           "moll_radius 0.0\nordinate_type {otype}\nang_order {order}\nordinate_sym_reduce true\n"
           "solve_type source-iteration\nsweep_pattern wavefront\nsweep_solve gaussian-elimination\n"
           "abs_tol {tol}\nmax_its 5000\ntfinal 1.0\noutput_log false\n")

SWEEPS = {  # name: deck parameters, seed
    "sweep_pwc": (dict(n=9, m=8, p=1, pwc="true", otype="spherical-triangle", order=2, tol="1e-10"), 11),
    "sweep_ho": (dict(n=8, m=9, p=2, pwc="false", otype="chebyshev-legendre", order=4, tol="1e-10"), 12),
    "sweep_ho3": (dict(n=6, m=5, p=3, pwc="false", otype="spherical-triangle", order=2, tol="1e-10"), 13),
}
REFLECT_SWEEPS = {  # name: deck parameters, reflecting-edge mask, seed
    "reflect_xmin": (dict(n=9, m=7, p=1, pwc="true", otype="chebyshev-legendre", order=4, tol="1e-10"), 1, 21),
    "reflect_xmax_ymin": (dict(n=8, m=10, p=2, pwc="true", otype="spherical-triangle", order=2, tol="1e-10"), 6, 22),
    "reflect_all": (dict(n=11, m=9, p=1, pwc="true", otype="chebyshev-legendre", order=6, tol="1e-10"), 15, 23),
    "reflect_ymax_q3": (dict(n=6, m=7, p=3, pwc="true", otype="spherical-triangle", order=1, tol="1e-10"), 8, 24),
    "reflect_xmin_ymax_ho": (dict(n=7, m=6, p=2, pwc="false", otype="chebyshev-legendre", order=4, tol="1e-10"), 9, 25),
}
REFLECT_KEYS = ("reflect_x_min", "reflect_x_max", "reflect_y_min", "reflect_y_max")
REFLECT_SOLVES = {  # source iteration with reflecting edges: deck parameters, mask
    "si_reflect": (dict(n=12, m=10, p=1, pwc="true", otype="spherical-triangle", order=2, tol="1e-10"), 13),
}
PERIODIC_SOLVES = {  # source iteration with periodic boundaries (deck key periodic)
    "si_periodic": dict(n=12, m=11, p=1, pwc="true", otype="chebyshev-legendre", order=4, tol="1e-10"),
}
IDC_DECK = LATTICE.format(n=10, m=9, p=1, pwc="true", otype="chebyshev-legendre", order=4, tol="1e-13") + \
    "cfl 0.5\nidc_type {form}\nidc_nodes_type {nodes}\nidc_num_nodes {nn}\nidc_num_corr {nc}\n" \
    "idc_use_collocation {colloc}\n"
IDC_STEPS = {  # name: (form, nodes, number of nodes, corrections, collocation)
    "idc_radau_update": ("update", "gauss-radau", 2, 2, "false"),
    "idc_lobatto_error": ("error", "gauss-lobatto", 3, 2, "false"),
    "idc_legendre_update": ("update", "gauss-legendre", 3, 1, "false"),
    "idc_chebradau_error_colloc": ("error", "chebyshev-radau", 3, 2, "true"),
    "idc_chebyshev_update": ("update", "chebyshev", 2, 1, "false"),
}
IDC_DT = 0.1
SOLVES = {
    "si_pwc": dict(n=14, m=14, p=1, pwc="true", otype="spherical-triangle", order=2, tol="1e-10"),
    "si_ho": dict(n=10, m=10, p=2, pwc="false", otype="chebyshev-legendre", order=4, tol="1e-12"),
}


def main():
    if not _ref.available():
        sys.exit("oracle/_ref/libref.so is missing: run oracle/build_ref.sh")
    out = {}
    tmp = tempfile.mkdtemp()
    # Reflecting decks first: the reference builds an ordinate set's mirror map only when the set is first
    # constructed while reflecting boundaries are on (OrdinateSet.cpp:499-504) and reuses built sets.
    for name, (params, mask, seed) in REFLECT_SWEEPS.items():
        path = os.path.join(tmp, name + ".deck")
        with open(path, "w") as f:
            f.write(LATTICE.format(**params))
            for b, key in enumerate(REFLECT_KEYS):
                if mask & (1 << b):
                    f.write(key + " true\n")
        d = _ref.Deck(path)
        rng = np.random.default_rng(seed)
        src = rng.uniform(-0.5, 1.0, size=d.psi_shape())
        _store_inputs(out, name, d)
        out[name + "/mask"] = np.array(mask)
        out[name + "/src"] = src
        out[name + "/steady"] = d.sweep(src)
        out[name + "/dt03"] = d.sweep(src, dt=0.3)
    for name, (params, mask) in REFLECT_SOLVES.items():
        path = os.path.join(tmp, name + ".deck")
        with open(path, "w") as f:
            f.write(LATTICE.format(**params))
            for b, key in enumerate(REFLECT_KEYS):
                if mask & (1 << b):
                    f.write(key + " true\n")
        d = _ref.Deck(path)
        _store_inputs(out, name, d)
        src = d.get_flux(1)
        out[name + "/mask"] = np.array(mask)
        out[name + "/src"] = src
        out[name + "/tol"] = np.array(float(params["tol"]))
        psi, its, err = d.solve(src, os.path.join(tmp, name + ".log"))
        out[name + "/psi"] = psi
        out[name + "/its"] = np.array(its)
        out[name + "/err"] = np.array(err)
        print(name, "iterations", its, "last change", err)
    for name, params in PERIODIC_SOLVES.items():
        path = os.path.join(tmp, name + ".deck")
        with open(path, "w") as f:
            f.write(LATTICE.format(**params) + "periodic true\n")
        d = _ref.Deck(path)
        _store_inputs(out, name, d)
        src = d.get_flux(1)
        out[name + "/src"] = src
        out[name + "/tol"] = np.array(float(params["tol"]))
        psi, its, err = d.solve(src, os.path.join(tmp, name + ".log"))
        out[name + "/psi"] = psi
        out[name + "/its"] = np.array(its)
        out[name + "/err"] = np.array(err)
        print(name, "iterations", its, "last change", err)
    for kind, orders in (("spherical-triangle", range(1, 9)), ("chebyshev-legendre", (2, 4, 6, 8))):
        for order in orders:
            for red in (0, 1):
                xi, eta, w, quad = _ref.ordinate_set(kind, order, red)
                key = f"ords/{kind}/{order}/{red}"
                out[key + "/xi"], out[key + "/eta"], out[key + "/w"], out[key + "/quad"] = xi, eta, w, quad
    for name, (params, seed) in SWEEPS.items():
        path = os.path.join(tmp, name + ".deck")
        with open(path, "w") as f:
            f.write(LATTICE.format(**params))
        d = _ref.Deck(path)
        rng = np.random.default_rng(seed)
        src = rng.uniform(-0.5, 1.0, size=d.psi_shape())
        _store_inputs(out, name, d)
        out[name + "/src"] = src
        out[name + "/steady"] = d.sweep(src)
        out[name + "/dt03"] = d.sweep(src, dt=0.3)
    for name, params in SOLVES.items():
        path = os.path.join(tmp, name + ".deck")
        with open(path, "w") as f:
            f.write(LATTICE.format(**params))
        d = _ref.Deck(path)
        _store_inputs(out, name, d)
        src = d.get_flux(1)
        out[name + "/src"] = src
        out[name + "/tol"] = np.array(float(params["tol"]))
        psi, its, err = d.solve(src, os.path.join(tmp, name + ".log"))
        out[name + "/psi"] = psi
        out[name + "/its"] = np.array(its)
        out[name + "/err"] = np.array(err)
        print(name, "iterations", its, "last change", err)
    for name, (form, nodes, nn, nc, colloc) in IDC_STEPS.items():
        path = os.path.join(tmp, name + ".deck")
        with open(path, "w") as f:
            f.write(IDC_DECK.format(form=form, nodes=nodes, nn=nn, nc=nc, colloc=colloc))
        d = _ref.Deck(path)
        _store_inputs(out, name, d)
        rng = np.random.default_rng(77)
        psi0 = np.zeros(d.psi_shape())
        psi0[1:-1, 1:-1] = rng.uniform(0.0, 1.0, size=psi0[1:-1, 1:-1].shape)
        src = d.get_flux(1)
        out[name + "/psi0"] = psi0
        out[name + "/src"] = src
        out[name + "/config"] = np.array([1 if form == "update" else 0, nn, nc, 1 if colloc == "true" else 0])
        out[name + "/nodes"] = np.array(nodes)
        out[name + "/dt"] = np.array(IDC_DT)
        out[name + "/psi1"] = d.idc_step(psi0, src, IDC_DT)
        print(name, "step norm", float(np.linalg.norm(out[name + "/psi1"])))
    np.savez_compressed(os.path.join(HERE, "ref_fixtures.npz"), **out)
    print("wrote", os.path.join(HERE, "ref_fixtures.npz"), len(out), "arrays")


def _store_inputs(out, name, d):
    out[name + "/dims"] = np.array([d.nx, d.ny, d.p, d.nq, d.xs_degree])
    out[name + "/box"] = np.array(d.box)
    out[name + "/xi"], out[name + "/eta"], out[name + "/w"], out[name + "/quad"] = d.xi, d.eta, d.w, d.quad
    out[name + "/sigma_t"] = d.get_xs(0)
    out[name + "/sigma_s"] = d.get_xs(1)


if __name__ == "__main__":
    main()
