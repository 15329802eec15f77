"""Scalar and angular flux files in the reference's single-rank binary format, and the norms and
difference its post-processor reports (main_postproc.cpp:394-436).

File layout (little-endian, as the reference writes it with fwrite):
  mesh header  per dimension int64 nx, float64 ax, float64 bx      (SpatialDomain.cpp:385-407)
  [angular]    ordinate-set name, null-terminated, "#R" appended when symmetry-reduced, then int64
               angular order                                          (OrdinateFlux.cpp:2855-2864)
  density      int64 DG degree, int64 byte count, int64 value count, then the float64 values with
               halo cells, [i][j][d][e] or [i][j][q][d][e]           (DensityFunction.cpp:2350-2358)
"""
from __future__ import annotations

import math
import struct
from dataclasses import dataclass

import numpy as np

from . import gauss_legendre, ordinates


@dataclass(frozen=True)
class Mesh:
    nx: int
    ny: int
    ax: float
    bx: float
    ay: float
    by: float

    @property
    def dx(self):
        return (self.bx - self.ax) / self.nx

    @property
    def dy(self):
        return (self.by - self.ay) / self.ny


@dataclass(frozen=True)
class OrdinateInfo:
    type: str
    order: int
    symmetric_reduce: bool


def _write_header(f, mesh: Mesh):
    f.write(struct.pack("<qdd", mesh.nx, mesh.ax, mesh.bx))
    f.write(struct.pack("<qdd", mesh.ny, mesh.ay, mesh.by))


def _write_values(f, degree, values):
    v = np.ascontiguousarray(values, dtype="<f8")
    f.write(struct.pack("<qqq", degree, v.size * 8, v.size))
    f.write(v.tobytes())


def _read_exact(f, n, path):
    b = f.read(n)
    if len(b) != n:
        raise RuntimeError(f"Failed to read field file '{path}': truncated")
    return b


def _read_header(f, path):
    nx, ax, bx = struct.unpack("<qdd", _read_exact(f, 24, path))
    ny, ay, by = struct.unpack("<qdd", _read_exact(f, 24, path))
    if nx <= 0 or ny <= 0:
        raise RuntimeError(f"Failed to read field file '{path}': bad mesh")
    return Mesh(nx, ny, ax, bx, ay, by)


def _read_values(f, path, shape_of):
    degree, nbytes, count = struct.unpack("<qqq", _read_exact(f, 24, path))
    if degree < 0 or degree > 64:
        raise RuntimeError(f"Failed to read field file '{path}': bad DG degree")
    shape = shape_of(degree)
    if count != int(np.prod(shape)) or nbytes != 8 * count:
        raise RuntimeError(f"Failed to read field file '{path}': sizes do not match the mesh")
    return degree, np.frombuffer(_read_exact(f, 8 * count, path), dtype="<f8").reshape(shape).astype(np.float64)


def write_density(path, values, mesh: Mesh):
    """Writes a scalar flux [nx+2][ny+2][p+1][p+1] (RKDG::DensityFunction::WriteToDisk)."""
    v = np.asarray(values, dtype=np.float64)
    if v.ndim != 4 or v.shape[:2] != (mesh.nx + 2, mesh.ny + 2) or v.shape[2] != v.shape[3]:
        raise ValueError("values must be [nx+2][ny+2][p+1][p+1]")
    with open(path, "wb") as f:
        _write_header(f, mesh)
        _write_values(f, v.shape[2] - 1, v)


def read_density(path):
    """(values [nx+2][ny+2][p+1][p+1], Mesh) of a DensityFunction file."""
    with open(path, "rb") as f:
        mesh = _read_header(f, path)
        _, v = _read_values(f, path, lambda p: (mesh.nx + 2, mesh.ny + 2, p + 1, p + 1))
    return v, mesh


def write_ordinate_flux(path, values, mesh: Mesh, info: OrdinateInfo):
    """Writes an angular flux [nx+2][ny+2][nq][p+1][p+1] (RKDG::OrdinateFlux::WriteToDisk)."""
    v = np.asarray(values, dtype=np.float64)
    if v.ndim != 5 or v.shape[:2] != (mesh.nx + 2, mesh.ny + 2):
        raise ValueError("values must be [nx+2][ny+2][nq][p+1][p+1]")
    name = info.type + ("#R" if info.symmetric_reduce else "")
    with open(path, "wb") as f:
        _write_header(f, mesh)
        f.write(name.encode() + b"\0")
        f.write(struct.pack("<q", info.order))
        _write_values(f, v.shape[3] - 1, v)


def read_ordinate_flux(path):
    """(values [nx+2][ny+2][nq][p+1][p+1], Mesh, OrdinateInfo) of an OrdinateFlux file; the ordinate
    count comes from rebuilding the named set (KeyError for a set that is not available)."""
    with open(path, "rb") as f:
        mesh = _read_header(f, path)
        name = bytearray()
        while True:
            c = _read_exact(f, 1, path)
            if c == b"\0":
                break
            name += c
        name = name.decode()
        reduced = False
        if "#" in name:
            reduced = name.endswith("R")
            name = name[:-2]
        (order,) = struct.unpack("<q", _read_exact(f, 8, path))
        if name not in ("spherical-triangle", "chebyshev-legendre", "level-symmetric"):
            raise KeyError(f"ordinate type '{name}' is not available")
        nq = ordinates(name, order, reduced)[0].size
        _, v = _read_values(f, path, lambda p: (mesh.nx + 2, mesh.ny + 2, nq, p + 1, p + 1))
    return v, mesh, OrdinateInfo(name, order, reduced)


# ---------------------------------------------------------------------------------------------------
def legendre_table(p, x):
    """P_0..P_p at the points x, shape x.shape + (p+1,)."""
    x = np.asarray(x, dtype=np.float64)
    out = np.empty(x.shape + (p + 1,))
    out[..., 0] = 1.0
    if p >= 1:
        out[..., 1] = x
    for k in range(2, p + 1):
        out[..., k] = ((2 * k - 1) * x * out[..., k - 1] - (k - 1) * out[..., k - 2]) / k
    return out


def l1_norm(values, mesh: Mesh):
    """Integral of |f| by the (p+1)-point Gauss rule per axis and cell (DensityFunction.cpp:1082-1170)."""
    p = values.shape[2] - 1
    g, w = gauss_legendre(p + 1)
    L = legendre_table(p, g)
    c = values[1:-1, 1:-1]
    f = np.einsum("ijde,ld,me->ijlm", c, L, L)
    return float(np.sum(np.abs(f) * w[None, None, :, None] * w[None, None, None, :]) * mesh.dx * mesh.dy / 4.0)


def l2_norm(values, mesh: Mesh):
    """sqrt of sum c^2 dx dy / ((2d+1)(2e+1)) over interior cells (DensityFunction.cpp:1175-1232)."""
    p = values.shape[2] - 1
    k = 2 * np.arange(p + 1) + 1.0
    return math.sqrt(float(np.sum(values[1:-1, 1:-1] ** 2 / np.outer(k, k))) * mesh.dx * mesh.dy)


def linf_norm(values, mesh: Mesh, tol=1e-3):
    """Per cell, max |f| on uniform grids of 2, 4, 8, ... points per axis (end points included) until
    two successive maxima agree to tol relative; the largest over cells (DensityFunction.cpp:1239-1400)."""
    p = values.shape[2] - 1
    c = values[1:-1, 1:-1].reshape(-1, p + 1, p + 1)
    if p == 0:
        return float(np.max(np.abs(c[:, 0, 0]))) if c.size else -1.0
    best = np.zeros(c.shape[0])
    prev = np.full(c.shape[0], np.inf)
    todo = np.arange(c.shape[0])
    part = 1
    while todo.size and part < 4096:
        part *= 2
        L = legendre_table(p, -1.0 + (2.0 / (part - 1.0)) * np.arange(part))
        f = np.abs(np.einsum("cde,ld,me->clm", c[todo], L, L)).reshape(todo.size, -1).max(axis=1)
        with np.errstate(invalid="ignore", divide="ignore"):
            go = np.abs(f - prev[todo]) / f > tol
        best[todo] = f
        prev[todo] = f
        todo = todo[go]
    return float(best.max()) if best.size else -1.0


def _projection(p_in, p_out, k):
    """T[s, a, d]: coefficient a (degree p_out) of P_d (degree p_in) restricted to subcell s of k."""
    g, w = gauss_legendre(p_out + 1)
    La = legendre_table(p_out, g)
    T = np.zeros((k, p_out + 1, p_in + 1))
    for s in range(k):
        Ld = legendre_table(p_in, (g + 2 * s + 1) / k - 1.0)
        T[s] = np.einsum("l,la,ld->ad", w * 1.0, La, Ld) * ((2 * np.arange(p_out + 1) + 1) / 2.0)[:, None]
    return T


def compute_difference(a, mesh_a: Mesh, b, mesh_b: Mesh):
    """(a - b, Mesh): coefficient difference on a common mesh at the higher DG degree, or both projected
    exactly onto the lcm(nx) x lcm(ny) mesh (DensityFunction.cpp:525-700)."""
    for u, v in ((mesh_a.ax, mesh_b.ax), (mesh_a.bx, mesh_b.bx), (mesh_a.ay, mesh_b.ay), (mesh_a.by, mesh_b.by)):
        if abs(u - v) > 1e-14:
            raise ValueError("domains are not compatible for computing a difference")
    pa, pb = a.shape[2] - 1, b.shape[2] - 1
    R = max(pa, pb)
    if (mesh_a.nx, mesh_a.ny) == (mesh_b.nx, mesh_b.ny):
        out = np.zeros((mesh_a.nx + 2, mesh_a.ny + 2, R + 1, R + 1))
        out[1:-1, 1:-1, :pa + 1, :pa + 1] += a[1:-1, 1:-1]
        out[1:-1, 1:-1, :pb + 1, :pb + 1] -= b[1:-1, 1:-1]
        return out, mesh_a
    m = Mesh(math.lcm(mesh_a.nx, mesh_b.nx), math.lcm(mesh_a.ny, mesh_b.ny), mesh_a.ax, mesh_a.bx, mesh_a.ay, mesh_a.by)
    out = np.zeros((m.nx + 2, m.ny + 2, R + 1, R + 1))
    for src, ms, sign in ((a, mesh_a, 1.0), (b, mesh_b, -1.0)):
        kx, ky = m.nx // ms.nx, m.ny // ms.ny
        p = src.shape[2] - 1
        Tx, Ty = _projection(p, R, kx), _projection(p, R, ky)
        c = src[1:-1, 1:-1]  # [nx][ny][d][e]
        # fine cell (I*kx + s, J*ky + t) takes Tx[s] c[I, J] Ty[t]^T
        fine = np.einsum("sad,IJde,tbe->IsJtab", Tx, c, Ty).reshape(m.nx, m.ny, R + 1, R + 1)
        out[1:-1, 1:-1] += sign * fine
    return out, m
