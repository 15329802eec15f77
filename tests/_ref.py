"""ctypes view of oracle/_ref/libref.so: the reference's own sources compiled unmodified with
oracle/ref_harness.cpp (oracle/build_ref.sh).  Test infrastructure: used by tests/golden/make_ref_fixtures.py
to write fixtures and by the CPU tests for live comparisons where the library exists (it is built only
where the reference tree and MKL are present)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(REPO, "oracle", "_ref", "libref.so")            # -O2 build: fixtures and parity
LIB_FAST = os.path.join(REPO, "oracle", "_ref", "libref_fast.so")  # the reference's -O3 flags: CPU timing
_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_int64)


def available(fast=False):
    return os.path.exists(LIB_FAST if fast else LIB)


def lib(fast=False):
    """Loads one of the two builds (the first call decides which for the process)."""
    global _lib
    if _lib is None:
        L = ctypes.CDLL(LIB_FAST if fast else LIB)
        L.ref_last_error.restype = ctypes.c_char_p
        L.ref_init.argtypes = [ctypes.c_char_p]
        L.ref_dims.argtypes = [_lp, _dp]
        L.ref_ordinates.argtypes = [_dp, _dp, _dp, _lp]
        L.ref_ordinate_set.argtypes = [ctypes.c_char_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int64, _lp, _dp, _dp,
                                       _dp, _lp]
        L.ref_get_flux.argtypes = [ctypes.c_int, _dp]
        L.ref_set_flux.argtypes = [ctypes.c_int, _dp]
        L.ref_get_xs.argtypes = [ctypes.c_int, _dp]
        L.ref_set_xs.argtypes = [ctypes.c_int, _dp]
        L.ref_sweep.argtypes = [ctypes.c_double]
        L.ref_solve.argtypes = [ctypes.c_double, ctypes.c_char_p]
        L.ref_idc_step.argtypes = [ctypes.c_double]
        _lib = L
    return _lib


def _check(rc):
    if rc != 0:
        raise RuntimeError("reference: " + lib().ref_last_error().decode())


def ordinate_set(kind, order, reduce):
    """(xi, eta, w, quad[5]) of the reference's OrdinateSet(order, reduce, kind)."""
    n = ctypes.c_int64(0)
    _check(lib().ref_ordinate_set(kind.encode(), order, int(reduce), 0, ctypes.byref(n), None, None, None, None))
    xi, eta, w = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
    quad = np.zeros(5, dtype=np.int64)
    _check(lib().ref_ordinate_set(kind.encode(), order, int(reduce), n.value, ctypes.byref(n), xi.ctypes.data_as(_dp),
                                  eta.ctypes.data_as(_dp), w.ctypes.data_as(_dp), quad.ctypes.data_as(_lp)))
    return xi, eta, w, quad


class Deck:
    """The reference's Init on a deck file, and operations on the objects it built."""

    def __init__(self, path):
        _check(lib().ref_init(str(path).encode()))
        dims = np.zeros(5, dtype=np.int64)
        box = np.zeros(4)
        _check(lib().ref_dims(dims.ctypes.data_as(_lp), box.ctypes.data_as(_dp)))
        self.nx, self.ny, self.p, self.nq, self.xs_degree = (int(v) for v in dims)
        self.box = tuple(float(v) for v in box)
        self.xi, self.eta, self.w = np.zeros(self.nq), np.zeros(self.nq), np.zeros(self.nq)
        self.quad = np.zeros(5, dtype=np.int64)
        _check(lib().ref_ordinates(self.xi.ctypes.data_as(_dp), self.eta.ctypes.data_as(_dp),
                                   self.w.ctypes.data_as(_dp), self.quad.ctypes.data_as(_lp)))

    def psi_shape(self):
        return (self.nx + 2, self.ny + 2, self.nq, self.p + 1, self.p + 1)

    def xs_shape(self):
        return (self.nx + 2, self.ny + 2, self.xs_degree + 1, self.xs_degree + 1)

    def get_flux(self, which):
        a = np.zeros(self.psi_shape())
        _check(lib().ref_get_flux(which, a.ctypes.data_as(_dp)))
        return a

    def set_flux(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.psi_shape()
        _check(lib().ref_set_flux(which, a.ctypes.data_as(_dp)))

    def get_xs(self, which):
        a = np.zeros(self.xs_shape())
        _check(lib().ref_get_xs(which, a.ctypes.data_as(_dp)))
        return a

    def set_xs(self, which, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.xs_shape()
        _check(lib().ref_set_xs(which, a.ctypes.data_as(_dp)))

    def sweep(self, src, dt=float("inf")):
        """L^{-1} src by the reference's SweepOperator (source object, in place)."""
        self.set_flux(1, src)
        _check(lib().ref_sweep(dt))
        return self.get_flux(1)

    def solve(self, src, log_path, dt=float("inf")):
        """The deck's ImplicitSolver on src; returns (psi, iterations, last error) from its log."""
        self.set_flux(1, src)
        _check(lib().ref_solve(dt, str(log_path).encode()))
        its, err = None, None
        with open(log_path) as f:
            for line in f:
                if line.strip().startswith("Iterations:"):
                    parts = line.split()
                    its, err = int(parts[1]), float(parts[4])
        return self.get_flux(1), its, err

    def idc_step(self, psi0, src, dt):
        """One step of the deck's IDC integrator from psi0 (ghost cells zero) with source src."""
        self.set_flux(0, psi0)
        self.set_flux(1, src)
        _check(lib().ref_idc_step(dt))
        return self.get_flux(0)
