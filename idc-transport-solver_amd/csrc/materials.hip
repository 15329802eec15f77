// Element-matrix inverses per (material, ordinate).
//
// With piecewise-constant cross sections the element matrix A = Aq + (sigma_t + 1/dt) dx dy I takes
// one of M values per ordinate (SweepLinearSolveSpec.cpp:284-298), M = number of distinct sigma_t
// cell values.  When M is small, the sweep multiplies by a precomputed A^{-1} instead of eliminating
// per cell.  The inverses are formed here in FP64 by the same elimination applied to the identity
// columns, scaled by dx dy (the sweep then applies them to B / (dx dy), whose source term is the
// unscaled source), and stored row-major as [M][nq][N][N] in the sweep's precision, rounded once.
#include "cell_solve.hpp"
#include "tsw_internal.hpp"

namespace tsw {

template <typename T, int P>
__global__ __launch_bounds__(64) void k_material_inverses(const double *__restrict__ xi, const double *__restrict__ eta,
                                                          const double *__restrict__ sig_m, int M, int nq, double dx,
                                                          double dy, double inv_dt, T *__restrict__ out) {
    constexpr int N = P * P;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= M * nq) return;
    const int m = t / nq, q = t - m * nq;
    T *dst = out + (size_t)t * N * N;
    for (int c = 0; c < N; ++c) {
        double A[N * N];
        assemble_matrix<double, P>(A, xi[q], eta[q], dx, dy, (sig_m[m] + inv_dt) * dx * dy);
        double e[N];
#pragma unroll
        for (int k = 0; k < N; ++k) e[k] = (k == c) ? 1.0 : 0.0;
        ge_solve<double, N>(A, e);
#pragma unroll
        for (int r = 0; r < N; ++r) dst[r * N + c] = T(e[r] * (dx * dy));
    }
}

template <typename T>
void launch_material_inverses(const Geom &g, const double *xi, const double *eta, const double *sig_m, int M,
                              double inv_dt, T *out, hipStream_t s) {
    const int n = M * g.nq;
    const dim3 grid((n + 63) / 64), block(64);
    switch (g.p) {
        case 0: hipLaunchKernelGGL((k_material_inverses<T, 1>), grid, block, 0, s, xi, eta, sig_m, M, g.nq, g.dx, g.dy, inv_dt, out); break;
        case 1: hipLaunchKernelGGL((k_material_inverses<T, 2>), grid, block, 0, s, xi, eta, sig_m, M, g.nq, g.dx, g.dy, inv_dt, out); break;
        case 2: hipLaunchKernelGGL((k_material_inverses<T, 3>), grid, block, 0, s, xi, eta, sig_m, M, g.nq, g.dx, g.dy, inv_dt, out); break;
        case 3: hipLaunchKernelGGL((k_material_inverses<T, 4>), grid, block, 0, s, xi, eta, sig_m, M, g.nq, g.dx, g.dy, inv_dt, out); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built");
    }
    TSW_HIP(hipGetLastError());
}

template void launch_material_inverses<double>(const Geom &, const double *, const double *, const double *, int, double, double *, hipStream_t);
template void launch_material_inverses<float>(const Geom &, const double *, const double *, const double *, int, double, float *, hipStream_t);

}  // namespace tsw
