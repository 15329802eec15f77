"""Python view of libtsw.so through its C-ABI (include/tsw.h), for tests, benchmarks and the
autograd wrapper.  No compute happens here: every call goes to the HIP library, and a missing
library is an error (there is no CPU fallback).

Arrays are numpy float64 arrays in the reference's host layouts (TSW_MEM_HOST), or torch tensors on
the GPU in the same layouts (TSW_MEM_DEVICE, passed by data pointer).
"""
from __future__ import annotations

import ctypes
import math
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
PKG_ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(PKG_ROOT, "libtsw.so")
HEADER = os.path.join(os.path.dirname(PKG_ROOT), "include", "tsw.h")

TSW_OK = 0
TSW_E_ARG = -1
TSW_E_HIP = -2
TSW_E_STATE = -3
TSW_E_NOCONV = -4
TSW_E_UNSUPPORTED = -5
FP64 = 0
FP32 = 1
MIXED = 2  # sweeps: FP32 with FP64 iterative refinement; solvers: as FP32
REFLECT_X_MIN, REFLECT_X_MAX, REFLECT_Y_MIN, REFLECT_Y_MAX = 1, 2, 4, 8
PERIODIC = 0x100
MEM_HOST = 0
MEM_DEVICE = 1
IDC_LOBATTO = 0
IDC_LEGENDRE = 1
IDC_RADAU = 2
IDC_CHEBYSHEV = 3
IDC_CHEBYSHEV_RADAU = 4
IDC_ERROR = 0
IDC_UPDATE = 1
# the reference's idc_nodes_type / idc_type strings (Quadrule.hpp:47-70, idc.cpp:44-50)
IDC_NODES_TYPES = {"gauss-lobatto": IDC_LOBATTO, "gauss-legendre": IDC_LEGENDRE, "gauss-radau": IDC_RADAU,
                   "chebyshev": IDC_CHEBYSHEV, "chebyshev-radau": IDC_CHEBYSHEV_RADAU}
IDC_TYPES = {"error": IDC_ERROR, "update": IDC_UPDATE}


class TswProblem(ctypes.Structure):
    """tsw_problem (include/tsw.h)."""
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("ax", ctypes.c_double), ("bx", ctypes.c_double),
                ("ay", ctypes.c_double), ("by", ctypes.c_double), ("p", ctypes.c_int), ("nq", ctypes.c_int),
                ("xi", ctypes.POINTER(ctypes.c_double)), ("eta", ctypes.POINTER(ctypes.c_double)),
                ("w", ctypes.POINTER(ctypes.c_double)), ("quad_off", ctypes.c_int * 5)]


class TswError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"tsw error {code}: {msg}")
        self.code = code


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p


def lib():
    """Loads libtsw.so (built in-tree by `make -C idc-transport-solver_amd`).

    When torch is importable it is imported first: torch ships its own HIP runtime with the same
    SONAME (libamdhip64.so.7), and loading it first makes libtsw.so bind to that copy, so a process
    holds one HIP runtime and device pointers move freely between torch and libtsw."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        c, d, i, ll = ctypes.c_void_p, ctypes.c_double, ctypes.c_int, ctypes.c_longlong
        L.tsw_create.argtypes = [i, ctypes.POINTER(c)]
        L.tsw_destroy.argtypes = [c]
        L.tsw_destroy.restype = None
        L.tsw_last_error.argtypes = [c]
        L.tsw_last_error.restype = ctypes.c_char_p
        L.tsw_problem_set.argtypes = [c, i, i, d, d, d, d, i, i, _dp, _dp, _dp, _ip]
        L.tsw_set_pattern.argtypes = [c, ctypes.c_char_p]
        L.tsw_set_tuning.argtypes = [c, ctypes.c_char_p, i]
        L.tsw_set_xs_degree.argtypes = [c, i]
        L.tsw_get_tuning.argtypes = [c, ctypes.c_char_p, _ip]
        L.tsw_sweep_apply.argtypes = [c, _vp, _vp, d, i, i]
        L.tsw_sweep_apply_noalloc.argtypes = [c, _vp, _vp, d, i]
        L.tsw_sweep_apply_multi.argtypes = [c, i, ctypes.POINTER(_vp), _vp, d, i, i]
        L.tsw_set_boundaries.argtypes = [c, i, _vp]
        L.tsw_reflect_boundaries.argtypes = [c, _vp, i]
        L.tsw_synchronize_halos.argtypes = [c, _vp, i]
        L.tsw_pmv.argtypes = [c, d, d, _vp, _vp, i, i]
        L.tsw_smv.argtypes = [c, d, d, _vp, _vp, _vp, i, i]
        L.tsw_lmv.argtypes = [c, d, d, _vp, _vp, _vp, i]
        L.tsw_si_solve.argtypes = [c, _vp, _vp, _vp, d, d, i, i, i, _ip, _dp]
        L.tsw_gmres_solve.argtypes = [c, _vp, _vp, _vp, d, d, i, i, i, i, _vp, _ip, _dp]
        L.tsw_sweep_vjp.argtypes = [c, _vp, _vp, d, _vp, _vp, _vp, i, i]
        L.tsw_solve_vjp.argtypes = [c, _vp, _vp, _vp, d, _vp, _vp, _vp, _vp, d, i, i, i, i, _ip, _dp]
        L.tsw_idc_set.argtypes = [c, i, i, i, i]
        L.tsw_idc_set_type.argtypes = [c, i]
        L.tsw_idc_nodes.argtypes = [i, i, i, _ip, _dp, _dp]
        L.tsw_idc_step.argtypes = [c, _vp, _vp, _vp, _vp, d, d, i, i, i]
        L.tsw_idc_run.argtypes = [c, _vp, _vp, _vp, _vp, d, i, d, i, i, i]
        L.tsw_idc_adjoint.argtypes = [c, _vp, _vp, _vp, _vp, d, i, i, _vp, _vp, _vp, _vp, d, i, i, i]
        L.tsw_ordinates_level_symmetric.argtypes = [i, _dp, _dp, _dp, _ip]
        L.tsw_ordinates.argtypes = [ctypes.c_char_p, i, i, i, _ip, _dp, _dp, _dp, _ip]
        L.tsw_gauss_legendre.argtypes = [i, _dp, _dp]
        L.tsw_sweep_timing.argtypes = [c, i, ctypes.POINTER(ll), _dp]
        L.tsw_batch_set.argtypes = [c, i, ctypes.POINTER(TswProblem), ctypes.POINTER(_vp), d, i]
        L.tsw_batch_sweep.argtypes = [c, ctypes.POINTER(_vp), i, i]
        _lib = L
    return _lib


def declared_symbols():
    """Names of the functions include/tsw.h declares."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|void|const char \*)\s*\*?\s*(tsw_\w+)\s*\(", text, re.M)))


def level_symmetric(order):
    """(xi, eta, w, quad_off) of the 2D level-symmetric S_order set."""
    nq = order * (order + 2) // 2
    xi, eta, w = np.zeros(nq), np.zeros(nq), np.zeros(nq)
    off = np.zeros(5, dtype=np.int32)
    rc = lib().tsw_ordinates_level_symmetric(order, xi.ctypes.data_as(_dp), eta.ctypes.data_as(_dp),
                                             w.ctypes.data_as(_dp), off.ctypes.data_as(_ip))
    if rc != TSW_OK:
        raise TswError(rc, f"no level-symmetric set of order {order}")
    return xi, eta, w, off


def idc_nodes(nodes_type, num_nodes):
    """(stage nodes on [0, 1], integration matrix W) of an IDC configuration (host-only)."""
    S = ctypes.c_int(0)
    rc = lib().tsw_idc_nodes(nodes_type, num_nodes, 0, ctypes.byref(S), None, None)
    if rc != TSW_OK:
        raise TswError(rc, f"no IDC nodes of type {nodes_type} and count {num_nodes}")
    nodes, W = np.zeros(S.value), np.zeros((S.value, S.value))
    rc = lib().tsw_idc_nodes(nodes_type, num_nodes, S.value, ctypes.byref(S), nodes.ctypes.data_as(_dp),
                             W.ctypes.data_as(_dp))
    if rc != TSW_OK:
        raise TswError(rc, f"no IDC nodes of type {nodes_type} and count {num_nodes}")
    return nodes, W


def ordinates(ordinate_type, order, symmetric_reduce=True):
    """(xi, eta, w, quad_off) of a named 2D ordinate set: "spherical-triangle", "chebyshev-legendre"
    (the reference's sets, in its order) or "level-symmetric"."""
    n = ctypes.c_int(0)
    name = ordinate_type.encode()
    rc = lib().tsw_ordinates(name, order, int(bool(symmetric_reduce)), 0, ctypes.byref(n), None, None, None, None)
    if rc != TSW_OK:
        raise TswError(rc, f"no ordinate set '{ordinate_type}' of order {order}")
    xi, eta, w = np.zeros(n.value), np.zeros(n.value), np.zeros(n.value)
    off = np.zeros(5, dtype=np.int32)
    rc = lib().tsw_ordinates(name, order, int(bool(symmetric_reduce)), n.value, ctypes.byref(n),
                             xi.ctypes.data_as(_dp), eta.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                             off.ctypes.data_as(_ip))
    if rc != TSW_OK:
        raise TswError(rc, f"no ordinate set '{ordinate_type}' of order {order}")
    return xi, eta, w, off


def gauss_legendre(n):
    """n-point Gauss-Legendre nodes (ascending) and weights on [-1, 1]."""
    x, w = np.zeros(n), np.zeros(n)
    rc = lib().tsw_gauss_legendre(n, x.ctypes.data_as(_dp), w.ctypes.data_as(_dp))
    if rc != TSW_OK:
        raise TswError(rc, f"no Gauss-Legendre rule with {n} points")
    return x, w


def _addr(a):
    """Data address of a numpy array or torch tensor, and the matching TSW_MEM_* value."""
    if isinstance(a, np.ndarray):
        if a.dtype != np.float64 or not a.flags["C_CONTIGUOUS"]:
            raise TypeError("arrays must be C-contiguous float64")
        return a.ctypes.data, MEM_HOST
    if a.dtype.__str__() != "torch.float64" or not a.is_contiguous():
        raise TypeError("tensors must be contiguous float64")
    return a.data_ptr(), (MEM_DEVICE if a.is_cuda else MEM_HOST)


def _addrs(*arrays):
    out, mems = [], set()
    for a in arrays:
        if a is None:
            out.append(None)
            continue
        p, m = _addr(a)
        out.append(p)
        mems.add(m)
    if len(mems) > 1:
        raise TypeError("all arrays of one call must be on the host or all on the device")
    mem = mems.pop() if mems else MEM_HOST
    if mem == MEM_DEVICE:
        # the library works on its own stream: finish what torch queued on the inputs first
        import torch
        torch.cuda.current_stream().synchronize()
    return out, mem


class Context:
    """One tsw_ctx: a device, a stream, one problem (mesh + ordinates) at a time."""

    def __init__(self, device=0):
        self._L = lib()
        h = ctypes.c_void_p()
        rc = self._L.tsw_create(device, ctypes.byref(h))
        if rc != TSW_OK:
            raise TswError(rc, f"tsw_create({device}) failed (no HIP device?)")
        self._h = h
        self.nx = self.ny = self.p = self.nq = None
        self.xs_degree = 0

    def close(self):
        if getattr(self, "_h", None):
            self._L.tsw_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc):
        if rc != TSW_OK:
            raise TswError(rc, self._L.tsw_last_error(self._h).decode())

    # sizes of the reference-layout arrays of the current problem
    def _sizes(self):
        if self.nx is None:
            return None
        cells = (self.nx + 2) * (self.ny + 2)
        N = (self.p + 1) ** 2
        return {"psi": cells * self.nq * N, "phi": cells * N, "cell": cells,
                "xs": cells * (N if self.xs_degree else 1)}

    def _expect(self, **arrays):
        """Refuses arrays whose element count does not match the problem (device arrays are passed by
        pointer, so a short buffer would otherwise be read out of bounds)."""
        sz = self._sizes()
        if sz is None:
            return
        for kind, a in arrays.items():
            if a is None:
                continue
            n = a.size if isinstance(a, np.ndarray) else a.numel()
            want = sz[kind.rstrip("0123456789")]
            if n != want:
                raise ValueError(f"{kind} array has {n} elements, the problem needs {want}")

    def problem_set(self, nx, ny, p, xi, eta, w, quad_off, ax=0.0, bx=1.0, ay=0.0, by=1.0):
        xi = np.ascontiguousarray(xi, dtype=np.float64)
        eta = np.ascontiguousarray(eta, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        off = np.ascontiguousarray(quad_off, dtype=np.int32)
        self._check(self._L.tsw_problem_set(self._h, nx, ny, ax, bx, ay, by, p, xi.size, xi.ctypes.data_as(_dp),
                                            eta.ctypes.data_as(_dp), w.ctypes.data_as(_dp),
                                            off.ctypes.data_as(_ip)))
        self.nx, self.ny, self.p, self.nq = nx, ny, p, xi.size
        self.xs_degree = 0

    def set_xs_degree(self, degree):
        """0: cross sections are cell means [i][j]; p: Legendre coefficients [i][j][u][v] of degree p."""
        self._check(self._L.tsw_set_xs_degree(self._h, int(degree)))
        self.xs_degree = int(degree)

    def set_pattern(self, name):
        self._check(self._L.tsw_set_pattern(self._h, name.encode()))

    def set_tuning(self, key, value):
        self._check(self._L.tsw_set_tuning(self._h, key.encode(), int(value)))

    def get_tuning(self, key):
        v = ctypes.c_int(0)
        self._check(self._L.tsw_get_tuning(self._h, key.encode(), ctypes.byref(v)))
        return v.value

    def sweep(self, psi, sigma_t, dt=math.inf, precision=FP64):
        self._expect(psi=psi, xs=sigma_t)
        (pp, sp), mem = _addrs(psi, sigma_t)
        self._check(self._L.tsw_sweep_apply(self._h, pp, sp, dt, precision, mem))
        return psi

    def set_boundaries(self, reflect_mask, reflect_map=None):
        """tsw_set_boundaries: reflecting edges (REFLECT_X_MIN | ...), optional mirror map [2][nq]."""
        if reflect_map is None:
            self._check(self._L.tsw_set_boundaries(self._h, int(reflect_mask), None))
        else:
            m = np.ascontiguousarray(reflect_map, dtype=np.int32).ravel()
            self._check(self._L.tsw_set_boundaries(self._h, int(reflect_mask), m.ctypes.data))

    def synchronize_halos(self, psi):
        """tsw_synchronize_halos: reflecting edges, then the periodic copy."""
        self._expect(psi=psi)
        (pp,), mem = _addrs(psi)
        self._check(self._L.tsw_synchronize_halos(self._h, pp, mem))
        return psi

    def reflect_boundaries(self, psi):
        """tsw_reflect_boundaries: the ghost cells of the reflecting edges from psi's interior."""
        self._expect(psi=psi)
        (pp,), mem = _addrs(psi)
        self._check(self._L.tsw_reflect_boundaries(self._h, pp, mem))
        return psi

    def sweep_multi(self, psi_list, sigma_t, dt=math.inf, precision=FP64):
        """tsw_sweep_apply_multi: every array of psi_list swept in place (same layouts as sweep)."""
        for a in psi_list:
            self._expect(psi=a, xs=sigma_t)
        ptrs, mem = _addrs(*psi_list, sigma_t)
        parr = (_vp * len(psi_list))(*ptrs[:-1])
        self._check(self._L.tsw_sweep_apply_multi(self._h, len(psi_list), parr, ptrs[-1], dt, precision, mem))
        return psi_list

    def sweep_noalloc(self, psi, sigma_t, dt=math.inf, precision=FP64):
        """tsw_sweep_apply_noalloc: device tensors only; no device allocation inside the call."""
        self._expect(psi=psi, xs=sigma_t)
        (pp, sp), mem = _addrs(psi, sigma_t)
        if mem != MEM_DEVICE:
            raise TypeError("the allocation-free sweep takes device tensors")
        self._check(self._L.tsw_sweep_apply_noalloc(self._h, pp, sp, dt, precision))
        return psi

    def pmv(self, psi, phi, alpha=1.0, beta=0.0, all_cells=False):
        self._expect(psi=psi, phi=phi)
        (a, b), mem = _addrs(psi, phi)
        self._check(self._L.tsw_pmv(self._h, alpha, beta, a, b, int(all_cells), mem))
        return phi

    def smv(self, sigma_s, phi, psi, alpha=1.0, beta=1.0, all_cells=False):
        self._expect(xs=sigma_s, phi=phi, psi=psi)
        (a, b, c), mem = _addrs(sigma_s, phi, psi)
        self._check(self._L.tsw_smv(self._h, alpha, beta, a, b, c, int(all_cells), mem))
        return psi

    def lmv(self, sigma_t, x, y, alpha=1.0, beta=0.0):
        self._expect(xs=sigma_t, psi=x, psi2=y)
        (a, b, c), mem = _addrs(sigma_t, x, y)
        self._check(self._L.tsw_lmv(self._h, alpha, beta, a, b, c, mem))
        return y

    def si_solve(self, source, sigma_t, sigma_s, dt=math.inf, abs_tol=1e-8, max_its=500, precision=FP64):
        self._expect(psi=source, xs=sigma_t, xs2=sigma_s)
        (a, b, c), mem = _addrs(source, sigma_t, sigma_s)
        it, err = ctypes.c_int(0), ctypes.c_double(0.0)
        self._check(self._L.tsw_si_solve(self._h, a, b, c, dt, abs_tol, max_its, precision, mem, ctypes.byref(it),
                                         ctypes.byref(err)))
        return it.value, err.value

    def gmres_solve(self, source, sigma_t, sigma_s, phi_out=None, dt=math.inf, abs_tol=1e-8, max_its=3000,
                    restart=30, precision=FP64):
        self._expect(psi=source, xs=sigma_t, xs2=sigma_s, phi=phi_out)
        (a, b, c, d), mem = _addrs(source, sigma_t, sigma_s, phi_out)
        it, res = ctypes.c_int(0), ctypes.c_double(0.0)
        self._check(self._L.tsw_gmres_solve(self._h, a, b, c, dt, abs_tol, max_its, restart, precision, mem, d,
                                            ctypes.byref(it), ctypes.byref(res)))
        return it.value, res.value

    def sweep_vjp(self, seed, grad_src, sigma_t, psi=None, grad_sigma_t=None, dt=math.inf, precision=FP64):
        """grad_src <- d<seed, psi>/d(source, inflow); grad_sigma_t <- d/dsigma_t (needs the forward psi)."""
        self._expect(psi=seed, psi2=grad_src, psi3=psi, xs=sigma_t, cell=grad_sigma_t)
        (a, b, c, d, e), mem = _addrs(psi, sigma_t, seed, grad_src, grad_sigma_t)
        self._check(self._L.tsw_sweep_vjp(self._h, a, b, dt, c, d, e, precision, mem))
        return grad_src

    def solve_vjp(self, seed, grad_src, sigma_t, sigma_s, psi=None, grad_sigma_t=None, grad_sigma_s=None,
                  dt=math.inf, abs_tol=1e-10, max_its=3000, restart=30, precision=FP64):
        """Adjoint of the scattering solve; returns (iterations, residual) of the adjoint GMRES."""
        self._expect(psi=seed, psi2=grad_src, psi3=psi, xs=sigma_t, xs2=sigma_s, cell3=grad_sigma_t,
                     cell4=grad_sigma_s)
        (a, b, c, d, e, f, h), mem = _addrs(psi, sigma_t, sigma_s, seed, grad_src, grad_sigma_t, grad_sigma_s)
        it, res = ctypes.c_int(0), ctypes.c_double(0.0)
        self._check(self._L.tsw_solve_vjp(self._h, a, b, c, dt, d, e, f, h, abs_tol, max_its, restart, precision,
                                          mem, ctypes.byref(it), ctypes.byref(res)))
        return it.value, res.value

    def idc_set(self, num_nodes=4, num_corrections=3, nodes_type=IDC_LOBATTO, use_collocation=False, form=None):
        self._check(self._L.tsw_idc_set(self._h, num_nodes, num_corrections, nodes_type, int(use_collocation)))
        if form is not None:
            self._check(self._L.tsw_idc_set_type(self._h, form))

    def idc_step(self, psi, source, sigma_t, sigma_s, dt, abs_tol=1e-12, max_its=3000, precision=FP64):
        """psi^n -> psi^{n+1} in place (zero halos); source holds Q and the inflow."""
        self._expect(psi=psi, psi2=source, xs=sigma_t, xs2=sigma_s)
        (a, b, c, d), mem = _addrs(psi, source, sigma_t, sigma_s)
        self._check(self._L.tsw_idc_step(self._h, a, b, c, d, dt, abs_tol, max_its, precision, mem))
        return psi

    def idc_run(self, psi, source, sigma_t, sigma_s, dt, nsteps, abs_tol=1e-12, max_its=3000, precision=FP64):
        self._expect(psi=psi, psi2=source, xs=sigma_t, xs2=sigma_s)
        (a, b, c, d), mem = _addrs(psi, source, sigma_t, sigma_s)
        self._check(self._L.tsw_idc_run(self._h, a, b, c, d, dt, nsteps, abs_tol, max_its, precision, mem))
        return psi

    def idc_adjoint(self, seed, grad_psi0, source, sigma_t, sigma_s, dt, nsteps, psi0=None, checkpoint_every=10,
                    grad_sigma_t=None, grad_sigma_s=None, abs_tol=1e-12, max_its=3000, precision=FP64):
        """grad_psi0 <- dJ/dpsi^0 of J = <seed, psi^nsteps>; optional cross-section gradients (need psi0)."""
        self._expect(psi=seed, psi2=grad_psi0, psi3=source, psi4=psi0, xs=sigma_t, xs2=sigma_s,
                     cell3=grad_sigma_t, cell4=grad_sigma_s)
        (a, b, c, d, e, f, h, k), mem = _addrs(psi0, source, sigma_t, sigma_s, seed, grad_psi0, grad_sigma_t,
                                               grad_sigma_s)
        self._check(self._L.tsw_idc_adjoint(self._h, a, b, c, d, dt, nsteps, checkpoint_every, e, f, h, k, abs_tol,
                                            max_its, precision, mem))
        return grad_psi0

    def batch_set(self, problems, sigma_t, dt=math.inf):
        """problems: list of dicts with nx, ny, p, xi, eta, w, quad_off and optional ax, bx, ay, by;
        sigma_t: list of cell arrays (numpy or device tensors)."""
        self._batch_keep = []
        arr = (TswProblem * len(problems))()
        for k, pr in enumerate(problems):
            xi = np.ascontiguousarray(pr["xi"], dtype=np.float64)
            eta = np.ascontiguousarray(pr["eta"], dtype=np.float64)
            w = np.ascontiguousarray(pr["w"], dtype=np.float64)
            self._batch_keep += [xi, eta, w]
            arr[k].nx, arr[k].ny, arr[k].p, arr[k].nq = pr["nx"], pr["ny"], pr["p"], xi.size
            arr[k].ax, arr[k].bx = pr.get("ax", 0.0), pr.get("bx", 1.0)
            arr[k].ay, arr[k].by = pr.get("ay", 0.0), pr.get("by", 1.0)
            arr[k].xi = xi.ctypes.data_as(_dp)
            arr[k].eta = eta.ctypes.data_as(_dp)
            arr[k].w = w.ctypes.data_as(_dp)
            for m in range(5):
                arr[k].quad_off[m] = int(pr["quad_off"][m])
        ptrs, mem = _addrs(*sigma_t)
        parr = (_vp * len(ptrs))(*ptrs)
        self._check(self._L.tsw_batch_set(self._h, len(problems), arr, parr, dt, mem))
        self._batch_shapes = [((pr["nx"] + 2) * (pr["ny"] + 2) * np.asarray(pr["xi"]).size * (pr["p"] + 1) ** 2)
                              for pr in problems]

    def batch_sweep(self, psi_list, precision=FP64):
        for a, n in zip(psi_list, self._batch_shapes):
            m = a.size if isinstance(a, np.ndarray) else a.numel()
            if m != n:
                raise ValueError(f"batch psi array has {m} elements, its problem needs {n}")
        ptrs, mem = _addrs(*psi_list)
        parr = (_vp * len(ptrs))(*ptrs)
        self._check(self._L.tsw_batch_sweep(self._h, parr, precision, mem))
        return psi_list

    def sweep_timing(self, reset=False):
        n, ms = ctypes.c_longlong(0), ctypes.c_double(0.0)
        self._check(self._L.tsw_sweep_timing(self._h, int(reset), ctypes.byref(n), ctypes.byref(ms)))
        return n.value, ms.value
