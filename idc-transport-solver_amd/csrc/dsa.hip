// Diffusion synthetic acceleration for the Peierls GMRES: y = x + C^{-1}(sigma_s x) applied to the
// cell-mean coefficient of x, as the reference's DSA preconditioner computes y = x + D^{-1} sigma_s x
// (PETScPeierlsSolver.cpp:588-640; the reference builds it only in 1D).  C is the cell-centred
// finite-volume diffusion operator
//     (C d)_c V_c = sum_faces G_f (d_c - d_n) + (sigma_t + 1/dt - sigma_s) V_c d_c,
// D = 1 / (3 (sigma_t + 1/dt)), face conductance G = face length / (h/(2 D_c) + h/(2 D_n)), vacuum
// (Marshak) faces G = face length / (h/(2 D_c) + 2).  C^{-1} is approximated by a fixed number of
// multigrid V-cycles (red-black Gauss-Seidel smoothing, tile-wise on the fine levels, 2x2 aggregation
// with piecewise-constant transfers, coarse operators P^T C P), a fixed linear operator, so GMRES stays
// valid.
//
// Level arrays: cells c = i*ny + j; cx[i*ny + j], i in 0..nx, conductance of the face left of cell i
// (i = 0 and nx are domain faces); cy[i*(ny+1) + j], j in 0..ny, likewise in y; sv[c] = absorption
// times volume.
#include "tsw_internal.hpp"

namespace tsw {

__global__ void k_dsa_rhs(const double *__restrict__ x, const double *__restrict__ sigma_s, double *__restrict__ b,
                          int nx, int ny, int N, double vol) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny) return;
    const int i = c / ny + 1, j = c % ny + 1;
    const size_t h = (size_t)i * (ny + 2) + j;
    b[c] = vol * sigma_s[h] * x[h * N];
}

__global__ void k_dsa_add(const double *__restrict__ d, double *__restrict__ y, int nx, int ny, int N) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny) return;
    const int i = c / ny + 1, j = c % ny + 1;
    y[((size_t)i * (ny + 2) + j) * N] += d[c];
}

// Smoothing of the fine levels: two red-black Gauss-Seidel sweeps inside each 16x16 tile of cells, the
// tile's one-cell halo held at its input values (Gauss-Seidel within a tile, Jacobi between tiles).
// Input and output are distinct arrays, so the result does not depend on the order the tiles run in,
// and the smoother is a fixed linear operator.  in null: zero start; xc non-null: the input is
// in + xc prolongated (piecewise constant over 2x2 aggregates of the coarse level, row length nyc).
// c0 is the colour swept first.  One thread per tile cell.
constexpr int kTile = 16;

__global__ __launch_bounds__(kTile * kTile) void k_mg_smooth_tile(MgLevel L, const double *__restrict__ in,
                                                                  double *__restrict__ out,
                                                                  const double *__restrict__ xc, int nyc, int c0) {
    __shared__ double xs[kTile + 2][kTile + 2];
    const int nx = L.nx, ny = L.ny;
    const int ti = blockIdx.x * kTile - 1, tj = blockIdx.y * kTile - 1;  // mesh index of halo row/column 0
    for (int k = threadIdx.x; k < (kTile + 2) * (kTile + 2); k += blockDim.x) {
        const int li = k / (kTile + 2), lj = k - li * (kTile + 2);
        const int i = ti + li, j = tj + lj;
        double v = 0.0;
        if (in && i >= 0 && i < nx && j >= 0 && j < ny) {
            v = in[i * ny + j];
            if (xc) v += xc[(i / 2) * nyc + j / 2];
        }
        xs[li][lj] = v;
    }
    const int li = threadIdx.x / kTile + 1, lj = threadIdx.x % kTile + 1;
    const int i = ti + li, j = tj + lj;
    const bool ok = i < nx && j < ny;
    double bb = 0.0, cw = 0.0, ce = 0.0, cs = 0.0, cn = 0.0, dg = 1.0;
    if (ok) {
        const double gw = L.cx[i * ny + j], ge = L.cx[(i + 1) * ny + j];
        const double gs = L.cy[i * (ny + 1) + j], gn = L.cy[i * (ny + 1) + j + 1];
        bb = L.b[i * ny + j];
        dg = L.sv[i * ny + j] + gw + ge + gs + gn;
        // a domain face couples to no neighbour (its conductance is in the diagonal only)
        cw = i > 0 ? gw : 0.0;
        ce = i < nx - 1 ? ge : 0.0;
        cs = j > 0 ? gs : 0.0;
        cn = j < ny - 1 ? gn : 0.0;
    }
    __syncthreads();
    for (int sweep = 0; sweep < 2; ++sweep)
        for (int h = 0; h < 2; ++h) {
            if (ok && ((i + j) & 1) == (h == 0 ? c0 : 1 - c0))
                xs[li][lj] = (bb + cw * xs[li - 1][lj] + ce * xs[li + 1][lj] + cs * xs[li][lj - 1] +
                              cn * xs[li][lj + 1]) / dg;
            __syncthreads();
        }
    if (ok) out[i * ny + j] = xs[li][lj];
}

// Coarse right-hand side: the sum of the fine residuals b - C x over each 2x2 aggregate, the residuals
// formed on the fly.
__global__ void k_mg_restrict_residual(const double *__restrict__ x, const double *__restrict__ b,
                                       const double *__restrict__ cx, const double *__restrict__ cy,
                                       const double *__restrict__ sv, double *__restrict__ bc, int nxf, int nyf,
                                       int nxc, int nyc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nxc * nyc) return;
    const int I = c / nyc, J = c % nyc;
    double acc = 0.0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj) {
            const int i = 2 * I + di, j = 2 * J + dj;
            if (i < nxf && j < nyf) {
                const int f = i * nyf + j;
                const double gw = cx[i * nyf + j], ge = cx[(i + 1) * nyf + j];
                const double gs = cy[i * (nyf + 1) + j], gn = cy[i * (nyf + 1) + j + 1];
                double ax = (sv[f] + gw + ge + gs + gn) * x[f];
                if (i > 0) ax -= gw * x[f - nyf];
                if (i < nxf - 1) ax -= ge * x[f + nyf];
                if (j > 0) ax -= gs * x[f - 1];
                if (j < nyf - 1) ax -= gn * x[f + 1];
                acc += b[f] - ax;
            }
        }
    bc[c] = acc;
}

// The V-cycle from level l0 down, in one workgroup with every level's arrays in LDS (for levels small
// enough that launches would dominate): descend (two red-black sweeps, residual restricted to the next
// level), coarsest level (x = A^{-1} b with its dense inverse when the level carries one, else 16
// sweeps), ascend (prolongation added, two sweeps in reverse colour order).
// Gauss-Seidel in place, phases separated by block barriers.  Level l0 starts from zero, or from its y
// when from_y; its result goes to lv[l0].y.  LDS per level: cx, cy, sv, b, x (mg_fused_lds_doubles).
struct MgLds {
    int nx, ny;
    double *cx, *cy, *sv, *b, *x;
};

__device__ void mg_rbgs_lds(const MgLds &L, int colour) {
    const int n = L.nx * L.ny;
    for (int c = threadIdx.x; c < n; c += blockDim.x) {
        const int i = c / L.ny, j = c % L.ny;
        if (((i + j) & 1) != colour) continue;
        const double gw = L.cx[i * L.ny + j], ge = L.cx[(i + 1) * L.ny + j];
        const double gs = L.cy[i * (L.ny + 1) + j], gn = L.cy[i * (L.ny + 1) + j + 1];
        double acc = L.b[c];
        if (i > 0) acc += gw * L.x[c - L.ny];
        if (i < L.nx - 1) acc += ge * L.x[c + L.ny];
        if (j > 0) acc += gs * L.x[c - 1];
        if (j < L.ny - 1) acc += gn * L.x[c + 1];
        L.x[c] = acc / (L.sv[c] + gw + ge + gs + gn);
    }
    __syncthreads();
}

__global__ __launch_bounds__(1024) void k_mg_fused(const MgLevel *__restrict__ lv, int l0, int nlev, int from_y) {
    extern __shared__ __align__(16) unsigned char mg_smem[];
    __shared__ MgLds Ls[kMgMaxLevels];
    double *lds = reinterpret_cast<double *>(mg_smem);
    const int nl = nlev - l0;
    if (threadIdx.x == 0) {
        size_t off = 0;
        for (int k = 0; k < nl; ++k) {
            const int nx = lv[l0 + k].nx, ny = lv[l0 + k].ny;
            MgLds &L = Ls[k];
            L.nx = nx;
            L.ny = ny;
            L.cx = lds + off;
            off += (size_t)(nx + 1) * ny;
            L.cy = lds + off;
            off += (size_t)nx * (ny + 1);
            L.sv = lds + off;
            off += (size_t)nx * ny;
            L.b = lds + off;
            off += (size_t)nx * ny;
            L.x = lds + off;
            off += (size_t)nx * ny;
        }
    }
    __syncthreads();
    for (int k = 0; k < nl; ++k) {
        const MgLevel G = lv[l0 + k];
        const MgLds L = Ls[k];
        const int n = L.nx * L.ny;
        for (int c = threadIdx.x; c < (L.nx + 1) * L.ny; c += blockDim.x) L.cx[c] = G.cx[c];
        for (int c = threadIdx.x; c < L.nx * (L.ny + 1); c += blockDim.x) L.cy[c] = G.cy[c];
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            L.sv[c] = G.sv[c];
            if (k == 0) {
                L.b[c] = G.b[c];
                L.x[c] = from_y ? G.y[c] : 0.0;
            }
        }
    }
    __syncthreads();
    const int last = nl - 1;
    for (int k = 0; k <= last; ++k) {
        const MgLds L = Ls[k];
        const int n = L.nx * L.ny;
        if (k > 0) {
            for (int c = threadIdx.x; c < n; c += blockDim.x) L.x[c] = 0.0;
            __syncthreads();
        }
        const double *inv = k == last ? lv[l0 + k].inv : nullptr;
        if (inv) {  // coarsest level solved exactly: x = A^{-1} b, one row per thread
            for (int c = threadIdx.x; c < n; c += blockDim.x) {
                double acc = 0.0;
                for (int m = 0; m < n; ++m) acc += inv[(size_t)c * n + m] * L.b[m];
                L.x[c] = acc;
            }
            __syncthreads();
        } else {
            const int sweeps = k == last ? 16 : 2;
            for (int it = 0; it < sweeps; ++it) {
                mg_rbgs_lds(L, 0);
                mg_rbgs_lds(L, 1);
            }
        }
        if (k == last) break;
        const MgLds C = Ls[k + 1];
        const int nc = C.nx * C.ny;
        for (int c = threadIdx.x; c < nc; c += blockDim.x) {
            const int I = c / C.ny, J = c % C.ny;
            double acc = 0.0;
            for (int di = 0; di < 2; ++di)
                for (int dj = 0; dj < 2; ++dj) {
                    const int i = 2 * I + di, j = 2 * J + dj;
                    if (i < L.nx && j < L.ny) {
                        const int f = i * L.ny + j;
                        const double gw = L.cx[i * L.ny + j], ge = L.cx[(i + 1) * L.ny + j];
                        const double gs = L.cy[i * (L.ny + 1) + j], gn = L.cy[i * (L.ny + 1) + j + 1];
                        double ax = (L.sv[f] + gw + ge + gs + gn) * L.x[f];
                        if (i > 0) ax -= gw * L.x[f - L.ny];
                        if (i < L.nx - 1) ax -= ge * L.x[f + L.ny];
                        if (j > 0) ax -= gs * L.x[f - 1];
                        if (j < L.ny - 1) ax -= gn * L.x[f + 1];
                        acc += L.b[f] - ax;
                    }
                }
            C.b[c] = acc;
        }
        __syncthreads();
    }
    for (int k = last - 1; k >= 0; --k) {
        const MgLds L = Ls[k], C = Ls[k + 1];
        const int n = L.nx * L.ny;
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const int i = c / L.ny, j = c % L.ny;
            L.x[c] += C.x[(i / 2) * C.ny + j / 2];
        }
        __syncthreads();
        for (int it = 0; it < 2; ++it) {
            mg_rbgs_lds(L, 1);
            mg_rbgs_lds(L, 0);
        }
    }
    const MgLds L0 = Ls[0];
    double *y = lv[l0].y;
    for (int c = threadIdx.x; c < L0.nx * L0.ny; c += blockDim.x) y[c] = L0.x[c];
}

namespace {
inline dim3 grid1(int n) { return dim3((unsigned)((n + 255) / 256)); }
}

size_t mg_fused_lds_doubles(const MgLevel *lv, int l0, int nlev) {
    size_t n = 0;
    for (int l = l0; l < nlev; ++l)
        n += (size_t)(lv[l].nx + 1) * lv[l].ny + (size_t)lv[l].nx * (lv[l].ny + 1) + 3 * (size_t)lv[l].nx * lv[l].ny;
    return n;
}

void launch_mg_fused(const MgLevel *lv_dev, int l0, int nlev, size_t lds_bytes, bool from_y, hipStream_t s) {
    if (nlev - l0 > kMgMaxLevels) throw Error(TSW_E_ARG, "too many multigrid levels for the fused kernel");
    if (lds_bytes > 64 * 1024)
        TSW_HIP(hipFuncSetAttribute((const void *)k_mg_fused, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes));
    hipLaunchKernelGGL(k_mg_fused, dim3(1), dim3(1024), lds_bytes, s, lv_dev, l0, nlev, (int)from_y);
    TSW_HIP(hipGetLastError());
}

void launch_dsa_rhs(const Geom &g, const double *x, const double *sigma_s, double *b, hipStream_t s) {
    hipLaunchKernelGGL(k_dsa_rhs, grid1(g.nx * g.ny), dim3(256), 0, s, x, sigma_s, b, g.nx, g.ny, g.N(), g.dx * g.dy);
    TSW_HIP(hipGetLastError());
}
void launch_dsa_add(const Geom &g, const double *d, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_dsa_add, grid1(g.nx * g.ny), dim3(256), 0, s, d, y, g.nx, g.ny, g.N());
    TSW_HIP(hipGetLastError());
}
void launch_mg_smooth(const MgLevel &L, const double *in, double *out, const MgLevel *coarse, int c0, hipStream_t s) {
    const dim3 grid((unsigned)((L.nx + kTile - 1) / kTile), (unsigned)((L.ny + kTile - 1) / kTile));
    hipLaunchKernelGGL(k_mg_smooth_tile, grid, dim3(kTile * kTile), 0, s, L, in, out,
                       coarse ? (const double *)coarse->y : nullptr, coarse ? coarse->ny : 0, c0);
    TSW_HIP(hipGetLastError());
}
void launch_mg_restrict_residual(const MgLevel &f, const MgLevel &c, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_restrict_residual, grid1(c.nx * c.ny), dim3(256), 0, s, f.x, f.b, f.cx, f.cy, f.sv, c.b,
                       f.nx, f.ny, c.nx, c.ny);
    TSW_HIP(hipGetLastError());
}

}  // namespace tsw
