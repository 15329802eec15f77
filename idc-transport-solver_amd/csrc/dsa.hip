// Diffusion synthetic acceleration for the Peierls GMRES: y = x + C^{-1}(sigma_s x) applied to the
// cell-mean coefficient of x, as the reference's DSA preconditioner computes y = x + D^{-1} sigma_s x
// (PETScPeierlsSolver.cpp:588-640; the reference builds it only in 1D).  C is the cell-centred
// finite-volume diffusion operator
//     (C d)_c V_c = sum_faces G_f (d_c - d_n) + (sigma_t + 1/dt - sigma_s) V_c d_c,
// D = 1 / (3 (sigma_t + 1/dt)), face conductance G = face length / (h/(2 D_c) + h/(2 D_n)), vacuum
// (Marshak) faces G = face length / (h/(2 D_c) + 2).  C^{-1} is approximated by a fixed number of
// multigrid V-cycles (red-black Gauss-Seidel smoothing, 2x2 aggregation with piecewise-constant
// transfers, coarse operators P^T C P), a fixed linear operator, so GMRES stays valid.
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

__global__ void k_mg_zero(double *__restrict__ x, int n) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) x[c] = 0.0;
}

// one colour of a red-black Gauss-Seidel sweep
__global__ void k_mg_rbgs(double *__restrict__ x, const double *__restrict__ b, const double *__restrict__ cx,
                          const double *__restrict__ cy, const double *__restrict__ sv, int nx, int ny, int colour) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny) return;
    const int i = c / ny, j = c % ny;
    if (((i + j) & 1) != colour) return;
    const double gw = cx[i * ny + j], ge = cx[(i + 1) * ny + j];
    const double gs = cy[i * (ny + 1) + j], gn = cy[i * (ny + 1) + j + 1];
    double acc = b[c];
    if (i > 0) acc += gw * x[c - ny];
    if (i < nx - 1) acc += ge * x[c + ny];
    if (j > 0) acc += gs * x[c - 1];
    if (j < ny - 1) acc += gn * x[c + 1];
    x[c] = acc / (sv[c] + gw + ge + gs + gn);
}

__global__ void k_mg_residual(const double *__restrict__ x, const double *__restrict__ b, double *__restrict__ r,
                              const double *__restrict__ cx, const double *__restrict__ cy,
                              const double *__restrict__ sv, int nx, int ny) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nx * ny) return;
    const int i = c / ny, j = c % ny;
    const double gw = cx[i * ny + j], ge = cx[(i + 1) * ny + j];
    const double gs = cy[i * (ny + 1) + j], gn = cy[i * (ny + 1) + j + 1];
    double ax = (sv[c] + gw + ge + gs + gn) * x[c];
    if (i > 0) ax -= gw * x[c - ny];
    if (i < nx - 1) ax -= ge * x[c + ny];
    if (j > 0) ax -= gs * x[c - 1];
    if (j < ny - 1) ax -= gn * x[c + 1];
    r[c] = b[c] - ax;
}

// coarse right-hand side: sum of the fine residuals of each 2x2 aggregate
__global__ void k_mg_restrict(const double *__restrict__ rf, double *__restrict__ bc, int nxf, int nyf, int nxc,
                              int nyc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nxc * nyc) return;
    const int I = c / nyc, J = c % nyc;
    double acc = 0.0;
    for (int di = 0; di < 2; ++di)
        for (int dj = 0; dj < 2; ++dj) {
            const int i = 2 * I + di, j = 2 * J + dj;
            if (i < nxf && j < nyf) acc += rf[i * nyf + j];
        }
    bc[c] = acc;
}

__global__ void k_mg_prolong_add(const double *__restrict__ xc, double *__restrict__ xf, int nxf, int nyf, int nyc) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nxf * nyf) return;
    const int i = c / nyf, j = c % nyf;
    xf[c] += xc[(i / 2) * nyc + j / 2];
}

// The V-cycle below level l0 in one workgroup, for levels small enough that launches would dominate:
// descend (zero x, two red-black sweeps, residual, restrict), coarsest level (16 sweeps), ascend (prolong,
// two sweeps in reverse colour order).  Phases are separated by block barriers; lv is a device array.
__device__ void mg_rbgs_block(const MgLevel &L, int colour) {
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

__global__ __launch_bounds__(1024) void k_mg_fused(const MgLevel *__restrict__ lv, int l0, int nlev) {
    const int last = nlev - 1;
    for (int l = l0; l <= last; ++l) {
        const MgLevel L = lv[l];
        const int n = L.nx * L.ny;
        for (int c = threadIdx.x; c < n; c += blockDim.x) L.x[c] = 0.0;
        __syncthreads();
        const int sweeps = l == last ? 16 : 2;
        for (int k = 0; k < sweeps; ++k) {
            mg_rbgs_block(L, 0);
            mg_rbgs_block(L, 1);
        }
        if (l == last) break;
        const MgLevel C = lv[l + 1];
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const int i = c / L.ny, j = c % L.ny;
            const double gw = L.cx[i * L.ny + j], ge = L.cx[(i + 1) * L.ny + j];
            const double gs = L.cy[i * (L.ny + 1) + j], gn = L.cy[i * (L.ny + 1) + j + 1];
            double ax = (L.sv[c] + gw + ge + gs + gn) * L.x[c];
            if (i > 0) ax -= gw * L.x[c - L.ny];
            if (i < L.nx - 1) ax -= ge * L.x[c + L.ny];
            if (j > 0) ax -= gs * L.x[c - 1];
            if (j < L.ny - 1) ax -= gn * L.x[c + 1];
            L.r[c] = L.b[c] - ax;
        }
        __syncthreads();
        const int nc = C.nx * C.ny;
        for (int c = threadIdx.x; c < nc; c += blockDim.x) {
            const int I = c / C.ny, J = c % C.ny;
            double acc = 0.0;
            for (int di = 0; di < 2; ++di)
                for (int dj = 0; dj < 2; ++dj) {
                    const int i = 2 * I + di, j = 2 * J + dj;
                    if (i < L.nx && j < L.ny) acc += L.r[i * L.ny + j];
                }
            C.b[c] = acc;
        }
        __syncthreads();
    }
    for (int l = last - 1; l >= l0; --l) {
        const MgLevel L = lv[l], C = lv[l + 1];
        const int n = L.nx * L.ny;
        for (int c = threadIdx.x; c < n; c += blockDim.x) {
            const int i = c / L.ny, j = c % L.ny;
            L.x[c] += C.x[(i / 2) * C.ny + j / 2];
        }
        __syncthreads();
        for (int k = 0; k < 2; ++k) {
            mg_rbgs_block(L, 1);
            mg_rbgs_block(L, 0);
        }
    }
}

namespace {
inline dim3 grid1(int n) { return dim3((unsigned)((n + 255) / 256)); }
}

void launch_mg_fused(const MgLevel *lv_dev, int l0, int nlev, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_fused, dim3(1), dim3(1024), 0, s, lv_dev, l0, nlev);
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
void launch_mg_zero(double *x, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_zero, grid1(n), dim3(256), 0, s, x, n);
    TSW_HIP(hipGetLastError());
}
void launch_mg_rbgs(const MgLevel &L, double *x, const double *b, int colour, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_rbgs, grid1(L.nx * L.ny), dim3(256), 0, s, x, b, L.cx, L.cy, L.sv, L.nx, L.ny, colour);
    TSW_HIP(hipGetLastError());
}
void launch_mg_residual(const MgLevel &L, const double *x, const double *b, double *r, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_residual, grid1(L.nx * L.ny), dim3(256), 0, s, x, b, r, L.cx, L.cy, L.sv, L.nx, L.ny);
    TSW_HIP(hipGetLastError());
}
void launch_mg_restrict(const MgLevel &f, const MgLevel &c, const double *rf, double *bc, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_restrict, grid1(c.nx * c.ny), dim3(256), 0, s, rf, bc, f.nx, f.ny, c.nx, c.ny);
    TSW_HIP(hipGetLastError());
}
void launch_mg_prolong_add(const MgLevel &f, const MgLevel &c, const double *xc, double *xf, hipStream_t s) {
    hipLaunchKernelGGL(k_mg_prolong_add, grid1(f.nx * f.ny), dim3(256), 0, s, xc, xf, f.nx, f.ny, c.ny);
    TSW_HIP(hipGetLastError());
}

}  // namespace tsw
