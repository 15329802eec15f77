// Diagonal-major ("sweep-order") layout of per-ordinate arrays.
//
// For an ordinate in quadrant Q, cell (i,j) has sweep coordinates c (column counted from the inflow
// edge in x) and r (row counted from the inflow edge in y):
//   c = xneg ? nx - i : i - 1,   r = yneg ? ny - j : j - 1,   s = c + r (its diagonal).
// Cells are numbered diagonal by diagonal, and along a diagonal by increasing r:
//   pos(s, r) = base(s) + r - rlo(s),  base(s) = cells on diagonals < s,  rlo(s) = max(0, s - nx + 1).
// A staggered sweep solves diagonal s at step s with one lane per row r, so the lanes of a wavefront
// touch consecutive positions.  Per-ordinate arrays are stored [q][k][pos] (k the DG coefficient),
// per-quadrant cell arrays [Q][pos].  Every ordinate of a quadrant shares the numbering.
#pragma once

#include <hip/hip_runtime.h>

namespace tsw {

struct DiagMap {
    int nx, ny;
    __host__ __device__ long long base(int s) const {
        const long long a = nx < ny ? nx : ny, b = nx < ny ? ny : nx, n1 = (long long)nx + ny - 1;
        if (s <= a) return (long long)s * (s + 1) / 2;
        if (s <= b) return a * (a + 1) / 2 + (long long)(s - a) * a;
        return a * (a + 1) / 2 + (b - a) * a + (long long)(s - b) * n1 - ((long long)s * (s - 1) / 2 - b * (b - 1) / 2);
    }
    __host__ __device__ int rlo(int s) const { return s - nx + 1 > 0 ? s - nx + 1 : 0; }
    __host__ __device__ long long pos(int s, int r) const { return base(s) + r - rlo(s); }
    __host__ __device__ long long cells() const { return (long long)nx * ny; }
    // sweep coordinates of mesh cell (i,j), 1 <= i <= nx, 1 <= j <= ny
    __host__ __device__ void local(int i, int j, bool xneg, bool yneg, int &c, int &r) const {
        c = xneg ? nx - i : i - 1;
        r = yneg ? ny - j : j - 1;
    }
    __host__ __device__ void mesh(int c, int r, bool xneg, bool yneg, int &i, int &j) const {
        i = xneg ? nx - c : c + 1;
        j = yneg ? ny - r : r + 1;
    }
    __host__ __device__ long long pos_of_cell(int i, int j, bool xneg, bool yneg) const {
        int c, r;
        local(i, j, xneg, yneg, c, r);
        return pos(c + r, r);
    }
};

// quadrant index 0..3 -> signs (0: ++, 1: +-, 2: -+, 3: --), as OrdinateSet::SetSigns
__host__ __device__ inline bool quad_xneg(int Q) { return Q >= 2; }
__host__ __device__ inline bool quad_yneg(int Q) { return (Q & 1) != 0; }

}  // namespace tsw
