// Transport operators around the sweep (Pmv, Smv, Lmv) and the FP64 vector kernels used by the
// Krylov and source-iteration drivers.  All are HBM-bound streaming kernels over the reference
// layouts; every thread owns whole coefficient vectors so loads of a wavefront are contiguous.
#include <algorithm>
#include <cmath>

#include "cell_solve.hpp"
#include "tsw_internal.hpp"

namespace tsw {

namespace {
constexpr int kBlock = 256;
constexpr double kInv4Pi = 0.079577471545947667884;  // 1/(4 pi)

inline int grid_for(size_t n) { return (int)((n + kBlock - 1) / kBlock); }

// Cell (i,j) of a linear index over the chosen domain (interior or halo-inclusive).
struct Domain {
    int i0, j0, ni, nj;
    __host__ __device__ size_t count() const { return (size_t)ni * nj; }
};
inline Domain domain(const Geom &g, int all_cells) {
    return all_cells ? Domain{0, 0, g.nx + 2, g.ny + 2} : Domain{1, 1, g.nx, g.ny};
}
}  // namespace

// ---- Moments from the ordinate-lane sweep's block partials: y = beta xb + alpha sum_b mom[b] on the
// interior (blocks summed in order), halos of y untouched.  The Pmv of a GMRES matvec whose sweep wrote
// the partials instead of psi.
__global__ __launch_bounds__(kBlock) void k_mom_sum(const double *__restrict__ mom, int nblk, size_t stride,
                                                    const double *xb, double *y, double alpha, double beta, int nx,
                                                    int ny, int N, const int *skip) {
    if (skip && *skip) return;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nx * ny * N) return;
    const size_t cell = t / N;
    const int k = (int)(t - cell * N);
    const int i = (int)(cell / ny) + 1, j = (int)(cell % ny) + 1;
    const size_t at = ((size_t)i * (ny + 2) + j) * N + k;
    double acc = mom[at];
    for (int b = 1; b < nblk; ++b) acc += mom[(size_t)b * stride + at];
    y[at] = beta * xb[at] + alpha * acc;
}

void launch_mom_sum(const Geom &g, int nblk, const double *mom, const double *xb, double *y, double alpha,
                    double beta, hipStream_t s, const int *skip) {
    const size_t n = (size_t)g.nx * g.ny * g.N();
    hipLaunchKernelGGL(k_mom_sum, dim3(grid_for(n)), dim3(kBlock), 0, s, mom, nblk, g.nphi(), xb, y, alpha, beta,
                       g.nx, g.ny, g.N(), skip);
    TSW_HIP(hipGetLastError());
}

// ---- Pmv: phi = beta*phi + alpha*sum_q w_q psi_q (TransportOperator.cpp:140-158) ----
// TI: storage type of psi; TO: type of phi, w and the accumulation.
template <typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void k_pmv(const TI *__restrict__ psi, TO *__restrict__ phi,
                                                const TO *__restrict__ w, TO alpha, TO beta, Domain dom, int ny,
                                                int nq, int N) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= dom.count() * N) return;
    const int k = (int)(t % N);
    const size_t c = t / N;
    const int i = dom.i0 + (int)(c / dom.nj), j = dom.j0 + (int)(c % dom.nj);
    const size_t cell = (size_t)i * (ny + 2) + j;
    TO acc = beta == TO(0) ? TO(0) : phi[cell * N + k] * beta;
    const TI *x = psi + cell * nq * N + k;
    for (int q = 0; q < nq; ++q) acc += alpha * w[q] * TO(x[(size_t)q * N]);
    phi[cell * N + k] = acc;
}

template <typename TI, typename TO>
void launch_pmv(const Geom &g, const TI *psi, TO *phi, const TO *w, TO alpha, TO beta, int all_cells,
                hipStream_t s) {
    const Domain dom = domain(g, all_cells);
    const size_t n = dom.count() * g.N();
    hipLaunchKernelGGL((k_pmv<TI, TO>), dim3(grid_for(n)), dim3(kBlock), 0, s, psi, phi, w, alpha, beta, dom, g.ny,
                       g.nq, g.N());
    TSW_HIP(hipGetLastError());
}

// ---- Smv: psi_q = beta*psi_q + alpha*sigma_s*phi/(4 pi) (TransportOperator.cpp:369-392) ----
template <typename T>
__global__ __launch_bounds__(kBlock) void k_smv(const T *__restrict__ sigma_s, const T *__restrict__ phi,
                                                T *__restrict__ psi, T alpha, T beta, Domain dom, int ny, int nq,
                                                int N) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per_cell = (size_t)nq * N;
    if (t >= dom.count() * per_cell) return;
    const size_t c = t / per_cell;
    const int r = (int)(t - c * per_cell);
    const int k = r % N;
    const int i = dom.i0 + (int)(c / dom.nj), j = dom.j0 + (int)(c % dom.nj);
    const size_t cell = (size_t)i * (ny + 2) + j;
    const T val = alpha * (sigma_s[cell] * phi[cell * N + k] * T(kInv4Pi));
    T *y = psi + cell * per_cell + r;
    *y = beta == T(0) ? val : beta * *y + val;
}

// Smv with a degree-p sigma_s (TransportOperator.cpp:375-381): the projection of sigma_s phi / (4 pi),
// val(d,e) = alpha (2d+1)(2e+1)/(16 pi) sum_{u,v} sigma(i,j,d,e,u,v) phi(u,v).
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_smv_xs(const T *__restrict__ xs, const T *__restrict__ phi,
                                                   T *__restrict__ psi, T alpha, T beta, Domain dom, int ny,
                                                   int nq) {
    constexpr int N = P * P;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t per_cell = (size_t)nq * N;
    if (t >= dom.count() * per_cell) return;
    const size_t c = t / per_cell;
    const int r = (int)(t - c * per_cell);
    const int k = r % N, d = k / P, e = k - P * d;
    const int i = dom.i0 + (int)(c / dom.nj), j = dom.j0 + (int)(c % dom.nj);
    const size_t cell = (size_t)i * (ny + 2) + j;
    const T *tc = xs + cell * N * N + k;
    const T *f = phi + cell * N;
    T v = T(0);
#pragma unroll
    for (int uv = 0; uv < N; ++uv) v += tc[(size_t)uv * N] * f[uv];
    v *= T((2 * d + 1) * (2 * e + 1)) / T(16.0 * M_PI);
    const T val = v * alpha;
    T *y = psi + cell * per_cell + r;
    *y = beta == T(0) ? val : beta * *y + val;
}

template <typename T>
void launch_smv(const Geom &g, const T *sigma_s, const T *phi, T *psi, T alpha, T beta, int all_cells,
                hipStream_t s, const T *xs) {
    const Domain dom = domain(g, all_cells);
    const size_t n = dom.count() * g.nq * g.N();
    if (xs) {
        switch (g.p) {
#define TSW_SMV_XS(PP) \
    case PP - 1: hipLaunchKernelGGL((k_smv_xs<T, PP>), dim3(grid_for(n)), dim3(kBlock), 0, s, xs, phi, psi, alpha, beta, dom, g.ny, g.nq); break;
            TSW_SMV_XS(1) TSW_SMV_XS(2) TSW_SMV_XS(3) TSW_SMV_XS(4)
#undef TSW_SMV_XS
            default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built (0..3)");
        }
    } else {
        hipLaunchKernelGGL(k_smv<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, sigma_s, phi, psi, alpha, beta, dom,
                           g.ny, g.nq, g.N());
    }
    TSW_HIP(hipGetLastError());
}

// ---- tensor of a degree-p cross section (CrossSection::ComputeTensor, CrossSection.cpp:165-189) ----
__global__ __launch_bounds__(kBlock) void k_xs_tensor(const double *__restrict__ sig, double *__restrict__ t,
                                                      size_t ncell, int P, Tpi tpi) {
    const int N = P * P;
    const size_t x = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= ncell * N * N) return;
    const size_t cell = x / ((size_t)N * N);
    const int r = (int)(x - cell * N * N), de = r / N, ab = r - de * N;
    const int d = de / P, e = de - d * P, a = ab / P, b = ab - a * P;
    const double *sc = sig + cell * N;
    double acc = 0.0;
    for (int u = 0; u < P; ++u)
        for (int v = 0; v < P; ++v)
            acc += sc[u * P + v] * tpi.v[a + P * (d + P * u)] * tpi.v[b + P * (e + P * v)];
    t[x] = acc;
}

void launch_xs_tensor(const Geom &g, const double *sig, const Tpi &tpi, double *t, hipStream_t s) {
    const size_t n = g.ncell_h() * g.N() * g.N();
    hipLaunchKernelGGL(k_xs_tensor, dim3(grid_for(n)), dim3(kBlock), 0, s, sig, t, g.ncell_h(), g.p + 1, tpi);
    TSW_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_xs_means(const double *__restrict__ sig, double *__restrict__ out,
                                                     size_t ncell, int N) {
    const size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < ncell) out[c] = sig[c * N];
}

void launch_xs_means(const Geom &g, const double *sig, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_xs_means, dim3(grid_for(g.ncell_h())), dim3(kBlock), 0, s, sig, out, g.ncell_h(), g.N());
    TSW_HIP(hipGetLastError());
}

// ---- Lmv: y = beta*y + alpha*(Omega.grad + sigma) x on interior cells (TransportOperator.cpp:1576-1658) ----
template <typename T, int P>
__global__ __launch_bounds__(kBlock) void k_lmv(const T *__restrict__ xi_, const T *__restrict__ eta_,
                                                const T *__restrict__ sigma, const T *__restrict__ x,
                                                T *__restrict__ y, T alpha, T beta, int nx, int ny, int nq, T dx,
                                                T dy, const T *__restrict__ xs) {
    constexpr int N = P * P;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nx * ny * nq) return;
    const int q = (int)(t % nq);
    const size_t c = t / nq;
    const int i = 1 + (int)(c / ny), j = 1 + (int)(c % ny);
    const size_t sj = (size_t)nq * N, si = (size_t)(ny + 2) * sj;
    const size_t base = (size_t)i * si + (size_t)j * sj + (size_t)q * N;
    const T *xc = x + base;
    const T xi = xi_[q], eta = eta_[q];
    T out[N];
    if (!xs) {
        const T sig = sigma[(size_t)i * (ny + 2) + j];
#pragma unroll
        for (int k = 0; k < N; ++k) out[k] = (beta == T(0) ? T(0) : y[base + k] * beta) + alpha * sig * xc[k];
    } else {  // degree-p sigma (:1587-1596): alpha (2d+1)(2e+1) sigma(i,j,d,e,u,v) x(u,v) / 4
        const T *tc = xs + ((size_t)i * (ny + 2) + j) * N * N;
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const int d = k / P, e = k - P * d;
            T acc = beta == T(0) ? T(0) : y[base + k] * beta;
#pragma unroll
            for (int uv = 0; uv < N; ++uv) acc += alpha * T((2 * d + 1) * (2 * e + 1)) * tc[uv * N + k] * xc[uv] / T(4);
            out[k] = acc;
        }
    }
    T coeff = alpha * xi / dx;
    if (xi < T(0)) {
        const T *xn = xc + si;
#pragma unroll
        for (int d = 0; d < P; ++d)
#pragma unroll
            for (int a = 0; a < P; ++a)
#pragma unroll
                for (int e = 0; e < P; ++e)
                    out[e + P * d] += coeff * T(2 * d + 1) *
                                      (T(ipow_m1(a)) * xn[e + P * a] + T(ipow_m1(d + a + 1)) * xc[e + P * a]);
    } else {
        const T *xn = xc - si;
#pragma unroll
        for (int d = 0; d < P; ++d)
#pragma unroll
            for (int a = 0; a < P; ++a)
#pragma unroll
                for (int e = 0; e < P; ++e)
                    out[e + P * d] += coeff * T(2 * d + 1) * (xc[e + P * a] + T(ipow_m1(d + 1)) * xn[e + P * a]);
    }
    coeff *= T(-2);
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int a = d - 1; a >= 0; a -= 2)
#pragma unroll
            for (int e = 0; e < P; ++e) out[e + P * d] += coeff * T(2 * d + 1) * xc[e + P * a];
    coeff = alpha * eta / dy;
    if (eta < T(0)) {
        const T *xn = xc + sj;
#pragma unroll
        for (int d = 0; d < P; ++d)
#pragma unroll
            for (int e = 0; e < P; ++e)
#pragma unroll
                for (int b = 0; b < P; ++b)
                    out[e + P * d] += coeff * T(2 * e + 1) *
                                      (T(ipow_m1(b)) * xn[b + P * d] + T(ipow_m1(e + b + 1)) * xc[b + P * d]);
    } else {
        const T *xn = xc - sj;
#pragma unroll
        for (int d = 0; d < P; ++d)
#pragma unroll
            for (int e = 0; e < P; ++e)
#pragma unroll
                for (int b = 0; b < P; ++b)
                    out[e + P * d] += coeff * T(2 * e + 1) * (xc[b + P * d] + T(ipow_m1(e + 1)) * xn[b + P * d]);
    }
    coeff *= T(-2);
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int e = 0; e < P; ++e)
#pragma unroll
            for (int b = e - 1; b >= 0; b -= 2) out[e + P * d] += coeff * T(2 * e + 1) * xc[b + P * d];
#pragma unroll
    for (int k = 0; k < N; ++k) y[base + k] = out[k];
}

template <typename T>
void launch_lmv(const Geom &g, const T *xi, const T *eta, const T *sigma_t, const T *x, T *y, T alpha, T beta,
                hipStream_t s, const T *xs) {
    const size_t n = (size_t)g.nx * g.ny * g.nq;
    const dim3 grid(grid_for(n)), block(kBlock);
    const T dx = T(g.dx), dy = T(g.dy);
    switch (g.p) {
        case 0: hipLaunchKernelGGL((k_lmv<T, 1>), grid, block, 0, s, xi, eta, sigma_t, x, y, alpha, beta, g.nx, g.ny, g.nq, dx, dy, xs); break;
        case 1: hipLaunchKernelGGL((k_lmv<T, 2>), grid, block, 0, s, xi, eta, sigma_t, x, y, alpha, beta, g.nx, g.ny, g.nq, dx, dy, xs); break;
        case 2: hipLaunchKernelGGL((k_lmv<T, 3>), grid, block, 0, s, xi, eta, sigma_t, x, y, alpha, beta, g.nx, g.ny, g.nq, dx, dy, xs); break;
        case 3: hipLaunchKernelGGL((k_lmv<T, 4>), grid, block, 0, s, xi, eta, sigma_t, x, y, alpha, beta, g.nx, g.ny, g.nq, dx, dy, xs); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built");
    }
    TSW_HIP(hipGetLastError());
}

template void launch_pmv<double, double>(const Geom &, const double *, double *, const double *, double, double, int, hipStream_t);
template void launch_pmv<float, float>(const Geom &, const float *, float *, const float *, float, float, int, hipStream_t);
template void launch_pmv<float, double>(const Geom &, const float *, double *, const double *, double, double, int, hipStream_t);
template void launch_smv<double>(const Geom &, const double *, const double *, double *, double, double, int, hipStream_t,
                                 const double *);
template void launch_smv<float>(const Geom &, const float *, const float *, float *, float, float, int, hipStream_t,
                                const float *);
template void launch_lmv<double>(const Geom &, const double *, const double *, const double *, const double *, double *, double, double,
                                 hipStream_t, const double *);
template void launch_lmv<float>(const Geom &, const float *, const float *, const float *, const float *, float *, float, float,
                                hipStream_t, const float *);

// ---- precision conversion ----
__global__ void k_d2f(const double *__restrict__ in, float *__restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (float)in[t];
}
__global__ void k_f2d(const float *__restrict__ in, double *__restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = (double)in[t];
}
void launch_convert_d2f(const double *in, float *out, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_d2f, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, n);
    TSW_HIP(hipGetLastError());
}
void launch_convert_f2d(const float *in, double *out, size_t n, hipStream_t s) {
    if (!n) return;
    hipLaunchKernelGGL(k_f2d, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, n);
    TSW_HIP(hipGetLastError());
}

// out = (double) in on the interior cells of a per-cell array (halo cells of out untouched)
__global__ void k_f2d_interior(const float *__restrict__ in, double *__restrict__ out, int nx, int ny,
                               size_t per_cell) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)(nx + 2) * (ny + 2) * per_cell) return;
    const size_t cell = t / per_cell;
    const int i = (int)(cell / (ny + 2)), j = (int)(cell % (ny + 2));
    if (i >= 1 && j >= 1 && i <= nx && j <= ny) out[t] = (double)in[t];
}
void launch_convert_f2d_interior(const Geom &g, size_t per_cell, const float *in, double *out, hipStream_t s) {
    const size_t n = g.ncell_h() * per_cell;
    hipLaunchKernelGGL(k_f2d_interior, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, g.nx, g.ny, per_cell);
    TSW_HIP(hipGetLastError());
}

// ---- region fill: halo cells (halo=1) or interior cells (halo=0) of a per-cell array get `value` ----
template <typename T>
__global__ void k_fill_region(T *__restrict__ a, int nx, int ny, size_t per_cell, int halo, T value) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)(nx + 2) * (ny + 2) * per_cell;
    if (t >= n) return;
    const size_t cell = t / per_cell;
    const int i = (int)(cell / (ny + 2)), j = (int)(cell % (ny + 2));
    const bool is_halo = i == 0 || j == 0 || i == nx + 1 || j == ny + 1;
    if (is_halo == (halo != 0)) a[t] = value;
}
template <typename T>
void launch_fill_region(const Geom &g, T *a, size_t per_cell, int halo, T value, hipStream_t s) {
    const size_t n = g.ncell_h() * per_cell;
    hipLaunchKernelGGL(k_fill_region<T>, dim3(grid_for(n)), dim3(kBlock), 0, s, a, g.nx, g.ny, per_cell, halo, value);
    TSW_HIP(hipGetLastError());
}
template void launch_fill_region<double>(const Geom &, double *, size_t, int, double, hipStream_t);
template void launch_fill_region<float>(const Geom &, float *, size_t, int, float, hipStream_t);

// out[k] += (double) in[k]
__global__ void k_acc_f2d(const float *__restrict__ in, double *__restrict__ out, size_t n) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] += (double)in[t];
}
void launch_accumulate_f2d(const float *in, double *out, size_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_acc_f2d, dim3(grid_for(n)), dim3(kBlock), 0, s, in, out, n);
    TSW_HIP(hipGetLastError());
}

// ---- input change detection: flag[0] = 1 when a and b differ bitwise in any of n doubles ----
__global__ void k_clear_flag(int *flag) {
    if (threadIdx.x == 0 && blockIdx.x == 0) flag[0] = 0;
}
__global__ void k_differs(size_t n, const unsigned long long *__restrict__ a, const unsigned long long *__restrict__ b,
                          int *flag) {
    bool diff = false;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        diff |= a[t] != b[t];
    if (diff) flag[0] = 1;  // every writer stores the same value
}
void launch_differs(size_t n, const double *a, const double *b, int *flag, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, s, flag);
    const size_t blocks = std::min<size_t>(grid_for(n), 1024);
    hipLaunchKernelGGL(k_differs, dim3((unsigned)blocks), dim3(kBlock), 0, s, n,
                       reinterpret_cast<const unsigned long long *>(a), reinterpret_cast<const unsigned long long *>(b),
                       flag);
    TSW_HIP(hipGetLastError());
}

// ---- FP64 vector kernels for the Krylov driver ----
__global__ void k_axpby(size_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) y[t] = b == 0.0 ? a * x[t] : a * x[t] + b * y[t];
}
void launch_axpby(size_t n, double a, const double *x, double b, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_axpby, dim3(grid_for(n)), dim3(kBlock), 0, s, n, a, x, b, y);
    TSW_HIP(hipGetLastError());
}

// y = sum_j a_j x_j over psi-layout arrays (per_cell values per cell of the (nx+2) x (ny+2) mesh) in one
// pass, accumulated in term order with fused multiply-adds; on halo cells only the terms with bit j of
// halo_terms set contribute (none: the halo is zero).  y may be one of the x_j.
template <int M>
__global__ void k_lincomb(size_t n, LinComb L, double *y, int nx, int ny, size_t per_cell) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const size_t cell = t / per_cell;
    const int i = (int)(cell / (size_t)(ny + 2)), j = (int)(cell % (size_t)(ny + 2));
    const bool halo = i == 0 || j == 0 || i == nx + 1 || j == ny + 1;
    const unsigned use = halo ? L.halo_terms : ~0u;
    double v[M];
#pragma unroll
    for (int k = 0; k < M; ++k) v[k] = ((use >> k) & 1u) ? L.x[k][t] : 0.0;
    double acc = L.a[0] * v[0];
#pragma unroll
    for (int k = 1; k < M; ++k) acc = __builtin_fma(L.a[k], v[k], acc);
    y[t] = use ? acc : 0.0;
}

void launch_lincomb(const Geom &g, size_t per_cell, const LinComb &L, double *y, hipStream_t s) {
    const size_t n = g.ncell_h() * per_cell;
    switch (L.m) {
#define TSW_LINCOMB_CASE(M) \
    case M: hipLaunchKernelGGL(k_lincomb<M>, dim3(grid_for(n)), dim3(kBlock), 0, s, n, L, y, g.nx, g.ny, per_cell); break;
        TSW_LINCOMB_CASE(1) TSW_LINCOMB_CASE(2) TSW_LINCOMB_CASE(3) TSW_LINCOMB_CASE(4) TSW_LINCOMB_CASE(5)
        TSW_LINCOMB_CASE(6) TSW_LINCOMB_CASE(7) TSW_LINCOMB_CASE(8) TSW_LINCOMB_CASE(9) TSW_LINCOMB_CASE(10)
        TSW_LINCOMB_CASE(11) TSW_LINCOMB_CASE(12) TSW_LINCOMB_CASE(13) TSW_LINCOMB_CASE(14) TSW_LINCOMB_CASE(15)
        TSW_LINCOMB_CASE(16)
#undef TSW_LINCOMB_CASE
        default: throw Error(TSW_E_ARG, "linear combination of 1..16 arrays");
    }
    TSW_HIP(hipGetLastError());
}

__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red[kBlock / 64];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) red[wid] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kBlock / 64; ++k) s += red[k];
    __syncthreads();
    return s;
}

// partials[i*nblocks + b] = partial dot of X_i with y over block b's grid-stride share
__global__ __launch_bounds__(kBlock) void k_multi_dot_partial(size_t n, const double *__restrict__ X, size_t ld,
                                                              const double *__restrict__ y,
                                                              double *__restrict__ partials) {
    const int i = blockIdx.y;
    const double *x = X + (size_t)i * ld;
    double acc = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        acc += x[t] * y[t];
    acc = block_sum(acc);
    if (threadIdx.x == 0) partials[(size_t)i * gridDim.x + blockIdx.x] = acc;
}
__global__ __launch_bounds__(kBlock) void k_finish_partials(const double *__restrict__ partials, int nblocks,
                                                            double *__restrict__ out) {
    const int i = blockIdx.x;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += partials[(size_t)i * nblocks + b];
    acc = block_sum(acc);
    if (threadIdx.x == 0) out[i] = acc;
}

int reduction_blocks() { return 512; }

size_t max_lds_per_block() {
    static thread_local int cached_dev = -1;
    static thread_local size_t cached = 0;
    int dev = 0;
    TSW_HIP(hipGetDevice(&dev));
    if (dev != cached_dev) {
        int v = 0;
        TSW_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        cached = (size_t)v;
        cached_dev = dev;
    }
    return cached;
}

void launch_multi_dot(size_t n, int m, const double *X, size_t ld, const double *y, double *out, double *partials,
                      int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_multi_dot_partial, dim3(nblocks, m), dim3(kBlock), 0, s, n, X, ld, y, partials);
    hipLaunchKernelGGL(k_finish_partials, dim3(m), dim3(kBlock), 0, s, partials, nblocks, out);
    TSW_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_multi_axpy(size_t n, int m, const double *__restrict__ X, size_t ld,
                                                       const double *__restrict__ h, double *__restrict__ y) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    double acc = y[t];
    for (int i = 0; i < m; ++i) acc -= h[i] * X[(size_t)i * ld + t];
    y[t] = acc;
}
void launch_multi_axpy(size_t n, int m, const double *X, size_t ld, const double *h, double *y, hipStream_t s) {
    hipLaunchKernelGGL(k_multi_axpy, dim3(grid_for(n)), dim3(kBlock), 0, s, n, m, X, ld, h, y);
    TSW_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_normalize(size_t n, const double *__restrict__ sq, double *__restrict__ v,
                                                      double *__restrict__ hn) {
    const double h = sqrt(*sq);
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t == 0) *hn = h;
    if (t < n && h > 0.0) v[t] = (1.0 / h) * v[t];
}
void launch_normalize(size_t n, const double *sq, double *v, double *hn, hipStream_t s) {
    hipLaunchKernelGGL(k_normalize, dim3(grid_for(n)), dim3(kBlock), 0, s, n, sq, v, hn);
    TSW_HIP(hipGetLastError());
}

__global__ __launch_bounds__(kBlock) void k_sqdist_partial(const double *__restrict__ a, const double *__restrict__ b,
                                                           int nx, int ny, int N, double *__restrict__ partials) {
    const size_t n = (size_t)nx * ny * N;
    double acc = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t c = t / N;
        const int k = (int)(t - c * N);
        const size_t cell = (size_t)(1 + c / ny) * (ny + 2) + 1 + c % ny;
        const double d = a[cell * N + k] - b[cell * N + k];
        acc += d * d;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}
// Source-iteration control for replayed graphs (SISolver.cpp:219-241 loop test, on the device): the
// partial sums of |cur - last|^2 over interior coefficients, then one block that finishes the norm,
// counts the iteration and sets ctl->done when the reference loop would stop (err <= tol or
// iters >= max_its).  Both do nothing once ctl->done is set, as does every kernel of the iteration.
__global__ __launch_bounds__(kBlock) void k_si_dist_partial(const double *__restrict__ a, const double *__restrict__ b,
                                                            int nx, int ny, int N, double *__restrict__ partials,
                                                            const int *__restrict__ skip) {
    if (*skip) return;
    const size_t n = (size_t)nx * ny * N;
    double acc = 0.0;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t c = t / N;
        const int k = (int)(t - c * N);
        const size_t cell = (size_t)(1 + c / ny) * (ny + 2) + 1 + c % ny;
        const double d = a[cell * N + k] - b[cell * N + k];
        acc += d * d;
    }
    acc = block_sum(acc);
    if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

__global__ __launch_bounds__(kBlock) void k_si_control(const double *__restrict__ partials, int nblocks, SiControl *ctl) {
    if (ctl->done) return;
    double acc = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) acc += partials[b];
    acc = block_sum(acc);
    if (threadIdx.x == 0) {
        const double err = sqrt(acc);
        ctl->err = err;
        ctl->iters += 1;
        if (err <= ctl->tol || ctl->iters >= ctl->max_its) ctl->done = 1;
    }
}

void launch_si_check(const Geom &g, const double *cur, const double *last, double *partials, int nblocks,
                     SiControl *ctl, hipStream_t s) {
    hipLaunchKernelGGL(k_si_dist_partial, dim3(nblocks), dim3(kBlock), 0, s, cur, last, g.nx, g.ny, g.N(), partials,
                       &ctl->done);
    hipLaunchKernelGGL(k_si_control, dim3(1), dim3(kBlock), 0, s, partials, nblocks, ctl);
    TSW_HIP(hipGetLastError());
}

void launch_sqdist_interior(const Geom &g, const double *a, const double *b, double *out, double *partials,
                            int nblocks, hipStream_t s) {
    hipLaunchKernelGGL(k_sqdist_partial, dim3(nblocks), dim3(kBlock), 0, s, a, b, g.nx, g.ny, g.N(), partials);
    hipLaunchKernelGGL(k_finish_partials, dim3(1), dim3(kBlock), 0, s, partials, nblocks, out);
    TSW_HIP(hipGetLastError());
}


// ---------------------------------------------------------------------------------------------------
// Reflecting boundaries: OrdinateFlux::ReflectBoundaries (RKDG/OrdinateFlux.cpp:1181-1420).  The ghost
// cells of edge `edge` (0 x-min, 1 x-max, 2 y-min, 3 y-max) take the mirrored ordinate's values of the
// adjacent interior cell, the coefficients odd in the reflected direction negated; x-min covers
// quadrants 0 and 1, x-max 2 and 3, y-min 0 and 2, y-max all four (as the reference does).
// ---------------------------------------------------------------------------------------------------
namespace {
__global__ void k_reflect_edge(double *__restrict__ psi, int nx, int ny, int nq, int P, int qo1, int qo2, int qo3,
                               const int *__restrict__ map, int edge) {
    const int N = P * P, dim = edge / 2, len = dim == 0 ? ny : nx;
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < (long long)nq * len * N;
         t += (long long)gridDim.x * blockDim.x) {
        const int k = (int)(t % N), r = (int)((t / N) % len), q = (int)(t / ((long long)N * len));
        const int Q = q >= qo3 ? 3 : (q >= qo2 ? 2 : (q >= qo1 ? 1 : 0));
        const bool on = edge == 0 ? Q <= 1 : (edge == 1 ? Q >= 2 : (edge == 2 ? (Q & 1) == 0 : true));
        if (!on) continue;
        const int d = k / P, e = k - d * P;
        int gi, gj, si, sj;
        if (dim == 0) {
            gi = edge == 0 ? 0 : nx + 1;
            si = edge == 0 ? 1 : nx;
            gj = sj = r + 1;
        } else {
            gj = edge == 2 ? 0 : ny + 1;
            sj = edge == 2 ? 1 : ny;
            gi = si = r + 1;
        }
        const int m = map[dim * nq + q];
        const double v = psi[(((size_t)si * (ny + 2) + sj) * nq + m) * N + k];
        psi[(((size_t)gi * (ny + 2) + gj) * nq + q) * N + k] = ((dim == 0 ? d : e) & 1) ? -v : v;
    }
}
}  // namespace

namespace {
__global__ void k_copy_outside_reflect(const double *__restrict__ src, double *__restrict__ dst, int nx, int ny,
                                       size_t per, int mask) {
    const size_t n = (size_t)(nx + 2) * (ny + 2) * per;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x) {
        const size_t cell = t / per;
        const int i = (int)(cell / (ny + 2)), j = (int)(cell % (ny + 2));
        const bool skip = (i == 0 && (mask & 1)) || (i == nx + 1 && (mask & 2)) || (j == 0 && (mask & 4)) ||
                          (j == ny + 1 && (mask & 8));
        if (!skip) dst[t] = src[t];
    }
}
}  // namespace

// dst <- src on the interior and the ghost cells of the edges that do not reflect (AngularFlux::Copy over
// OpDomain::Interior | OpDomain::Boundary, whose edge rule is DomainDecomposition.cpp:285-336).
void launch_copy_outside_reflect(const Geom &g, const double *src, double *dst, int mask, hipStream_t s) {
    const size_t n = g.npsi();
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_copy_outside_reflect, dim3(blocks), dim3(256), 0, s, src, dst, g.nx, g.ny,
                       (size_t)g.nq * g.N(), mask);
    TSW_HIP(hipGetLastError());
}

namespace {
// DensityFunction::SynchronizeHalos across periodic boundaries (Abstract/DensityFunction.cpp:927-956):
// phase 0 the y-halos of columns 1..nx, phase 1 the x-halos of rows 1..ny, every value of the cell.
__global__ void k_periodic_halos(double *__restrict__ psi, int nx, int ny, size_t per, int phase) {
    const int len = phase == 0 ? nx : ny;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < (size_t)len * per;
         t += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(t / per) + 1;
        const size_t l = t % per;
        if (phase == 0) {
            psi[((size_t)k * (ny + 2) + 0) * per + l] = psi[((size_t)k * (ny + 2) + ny) * per + l];
            psi[((size_t)k * (ny + 2) + ny + 1) * per + l] = psi[((size_t)k * (ny + 2) + 1) * per + l];
        } else {
            psi[((size_t)0 * (ny + 2) + k) * per + l] = psi[((size_t)nx * (ny + 2) + k) * per + l];
            psi[((size_t)(nx + 1) * (ny + 2) + k) * per + l] = psi[((size_t)1 * (ny + 2) + k) * per + l];
        }
    }
}
}  // namespace

void launch_periodic_halos(const Geom &g, double *psi, hipStream_t s) {
    const size_t per = (size_t)g.nq * g.N();
    for (int phase = 0; phase < 2; ++phase) {
        const size_t n = (size_t)(phase == 0 ? g.nx : g.ny) * per;
        const int blocks = (int)std::min<size_t>((n + 255) / 256, 2048);
        hipLaunchKernelGGL(k_periodic_halos, dim3(blocks), dim3(256), 0, s, psi, g.nx, g.ny, per, phase);
        TSW_HIP(hipGetLastError());
    }
}

void launch_reflect_boundaries(const Geom &g, double *psi, const int *map, int mask, hipStream_t s) {
    for (int edge = 0; edge < 4; ++edge) {
        if (!(mask & (1 << edge))) continue;
        const long long n = (long long)g.nq * (edge < 2 ? g.ny : g.nx) * g.N();
        const int blocks = (int)std::min<long long>((n + 255) / 256, 2048);
        hipLaunchKernelGGL(k_reflect_edge, dim3(blocks), dim3(256), 0, s, psi, g.nx, g.ny, g.nq, g.p + 1,
                           g.quad_off[1], g.quad_off[2], g.quad_off[3], map, edge);
        TSW_HIP(hipGetLastError());
    }
}

}  // namespace tsw
