// Per-cell, per-ordinate upwind-DG element solve, evaluated entirely in registers.
//
// For ordinate (xi, eta) on cell (i,j) with sizes dx, dy, Legendre degree p per axis and
// N = (p+1)^2 unknowns ordered r = e + (p+1)*d (d: x-degree, e: y-degree):
//   A = Aq + (sigma_t + 1/dt) dx dy I
//   Aq[(d,e),(a,e)] = cx (2d+1) [ s_x(d,a) - 2*[a<d, d-a odd] ],  cx = dy*xi,
//       s_x(d,a) = (-1)^(d+a+1) if xi < 0 else 1                       (SweepLinearSolveSpec.cpp:120-138)
//   Aq[(d,e),(d,b)] += the same in y with cy = dx*eta, (2e+1)         (:141-155)
//   B(d,e) = dx dy src(d,e) + (2d+1) sum_a c_x(d,a) cx psi_up_x(a,e) + (2e+1) sum_b c_y(e,b) cy psi_up_y(d,b)
//       c_x(d,a) = (-1)^(a+1) if xi <= 0 else (-1)^d                   (:473-598)
// then unpivoted Gaussian elimination and back substitution (SIMD_LinearAlgebra.hpp:78-115) with
// the reference's operation order, and the solution stored back over the source.
#pragma once

#include <hip/hip_runtime.h>

namespace tsw {

__host__ __device__ constexpr int ipow_m1(int k) { return (k & 1) ? -1 : 1; }

// c + a * b with one rounding.  Sums of two products (a0 b0 + a1 b1) may be contracted with either product
// rounded first, so the element matvecs accumulate through this to give every kernel variant the same bits.
template <typename T>
__device__ __forceinline__ T fmadd(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

// Builds B (right-hand side) for one lane from the source and the two upwind neighbours.
template <typename T, int P>
__device__ __forceinline__ void assemble_rhs(T (&B)[P * P], const T *__restrict__ src, const T *__restrict__ upx,
                                             const T *__restrict__ upy, T xi, T eta, T dx, T dy) {
    constexpr int N = P * P;
    const bool xneg = xi <= T(0), yneg = eta <= T(0);
    T W0[N], W1[N];
#pragma unroll
    for (int k = 0; k < N; ++k) {
        W0[k] = upx[k] * (dy * xi);
        W1[k] = upy[k] * (dx * eta);
    }
    const T area = dx * dy;
#pragma unroll
    for (int k = 0; k < N; ++k) B[k] = area * src[k];
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int e = 0; e < P; ++e)
#pragma unroll
            for (int a = 0; a < P; ++a) {
                const T c = xneg ? T(ipow_m1(a + 1) * (2 * d + 1)) : T(ipow_m1(d) * (2 * d + 1));
                B[e + P * d] += W0[e + P * a] * c;
            }
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int e = 0; e < P; ++e)
#pragma unroll
            for (int b = 0; b < P; ++b) {
                const T c = yneg ? T(ipow_m1(b + 1) * (2 * e + 1)) : T(ipow_m1(e) * (2 * e + 1));
                B[e + P * d] += W1[b + P * d] * c;
            }
}

// Builds A = Aq + diag_add * I (column-major, A[row + N*col]).
template <typename T, int P>
__device__ __forceinline__ void assemble_matrix(T (&A)[P * P * P * P], T xi, T eta, T dx, T dy, T diag_add) {
    constexpr int N = P * P;
    const bool xneg = xi < T(0), yneg = eta < T(0);
    const T cx = dy * xi, cy = dx * eta;
#pragma unroll
    for (int k = 0; k < N * N; ++k) A[k] = T(0);
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int e = 0; e < P; ++e) {
            const int row = e + P * d;
            const T tx = cx * T(2 * d + 1);
            const T ty = cy * T(2 * e + 1);
#pragma unroll
            for (int a = d - 1; a >= 0; a -= 2) A[row + N * (e + P * a)] -= T(2) * tx;
#pragma unroll
            for (int a = 0; a < P; ++a) A[row + N * (e + P * a)] += xneg ? T(ipow_m1(d + a + 1)) * tx : tx;
#pragma unroll
            for (int b = e - 1; b >= 0; b -= 2) A[row + N * (b + P * d)] -= T(2) * ty;
#pragma unroll
            for (int b = 0; b < P; ++b) A[row + N * (b + P * d)] += yneg ? T(ipow_m1(e + b + 1)) * ty : ty;
        }
#pragma unroll
    for (int r = 0; r < N; ++r) A[r + N * r] += diag_add;
}

// y = A x for the matrix assemble_matrix builds, without forming it: A is Kx (x) I + I (x) Ky plus
// diag_add on the diagonal, Kx(d,a) = cx (2d+1) (s(d,a) - 2 [a < d, d - a odd]) with s = 1, or
// (-1)^(d+a+1) for xi < 0 (and Ky likewise in e, b with cy, eta), so the product costs 2 P^3 + P^2
// multiply-adds and P^2 coefficients instead of P^4 of each.
template <typename T, int P>
__device__ __forceinline__ void apply_element(const T (&x)[P * P], T xi, T eta, T dx, T dy, T diag_add,
                                              T (&y)[P * P]) {
    const bool xneg = xi < T(0), yneg = eta < T(0);
    const T cx = dy * xi, cy = dx * eta;
    T kx[P][P], ky[P][P];
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int a = 0; a < P; ++a) {
            const T sx = xneg ? T(ipow_m1(d + a + 1)) : T(1), sy = yneg ? T(ipow_m1(d + a + 1)) : T(1);
            const T low = (a < d && ((d - a) & 1)) ? T(2) : T(0);
            kx[d][a] = cx * T(2 * d + 1) * (sx - low);
            ky[d][a] = cy * T(2 * d + 1) * (sy - low);
        }
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int e = 0; e < P; ++e) {
            T acc = diag_add * x[e + P * d];
#pragma unroll
            for (int a = 0; a < P; ++a) acc += kx[d][a] * x[e + P * a];
#pragma unroll
            for (int b = 0; b < P; ++b) acc += ky[e][b] * x[b + P * d];
            y[e + P * d] = acc;
        }
}

// Unpivoted Gaussian elimination, same operation order as the reference's SIMD kernel.
template <typename T, int N>
__device__ __forceinline__ void ge_solve(T (&A)[N * N], T (&b)[N]) {
#pragma unroll
    for (int j = 0; j < N; ++j) {
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            A[i + N * j] /= -A[j + N * j];
            b[i] += A[i + N * j] * b[j];
        }
#pragma unroll
        for (int k = j + 1; k < N; ++k)
#pragma unroll
            for (int i = j + 1; i < N; ++i) A[i + N * k] += A[i + N * j] * A[j + N * k];
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        b[i] /= A[i + N * i];
#pragma unroll
        for (int j = 0; j < i; ++j) b[j] -= A[j + N * i] * b[i];
    }
}

}  // namespace tsw

namespace tsw {

// ---------------------------------------------------------------------------------------------
// Face traces.  The upwind contribution to B only needs the upwind cell's polynomial restricted to
// the shared face: for xi > 0 the right-face trace Tx(e) = sum_a psi(a,e), for xi <= 0 the
// left-face trace Tx(e) = sum_a (-1)^a psi(a,e) (P_a(1) = 1, P_a(-1) = (-1)^a); same in y.
// In terms of traces the reference's CalcB (SweepLinearSolveSpec.cpp:545-598) reads
//   B(d,e) = dx dy src(d,e) + (2d+1) cx kx(d) Tx(e) + (2e+1) cy ky(e) Ty(d),
//   kx(d) = (-1)^d if xi > 0 else -1,   ky(e) = (-1)^e if eta > 0 else -1.
// ---------------------------------------------------------------------------------------------
template <typename T, int P>
__device__ __forceinline__ void trace_x(const T *__restrict__ psi, bool xneg, T (&tx)[P]) {
#pragma unroll
    for (int e = 0; e < P; ++e) {
        T s = T(0);
#pragma unroll
        for (int a = 0; a < P; ++a) s += (xneg && (a & 1)) ? -psi[e + P * a] : psi[e + P * a];
        tx[e] = s;
    }
}

template <typename T, int P>
__device__ __forceinline__ void trace_y(const T *__restrict__ psi, bool yneg, T (&ty)[P]) {
#pragma unroll
    for (int d = 0; d < P; ++d) {
        T s = T(0);
#pragma unroll
        for (int b = 0; b < P; ++b) s += (yneg && (b & 1)) ? -psi[b + P * d] : psi[b + P * d];
        ty[d] = s;
    }
}

template <typename T, int P>
__device__ __forceinline__ void rhs_from_traces(T (&B)[P * P], const T (&src)[P * P], const T (&tx)[P],
                                                const T (&ty)[P], T xi, T eta, T dx, T dy) {
    const bool xneg = xi <= T(0), yneg = eta <= T(0);
    const T cx = dy * xi, cy = dx * eta, area = dx * dy;
#pragma unroll
    for (int d = 0; d < P; ++d)
#pragma unroll
        for (int e = 0; e < P; ++e) {
            const T kx = xneg ? T(-(2 * d + 1)) : T(ipow_m1(d) * (2 * d + 1));
            const T ky = yneg ? T(-(2 * e + 1)) : T(ipow_m1(e) * (2 * e + 1));
            B[e + P * d] = area * src[e + P * d] + (cx * kx) * tx[e] + (cy * ky) * ty[d];
        }
}

}  // namespace tsw

namespace tsw {

// Unpivoted Gaussian elimination in the reference's order, with each pivot inverted once and the
// divisions of SIMD_LinearAlgebra.hpp:91,105 replaced by multiplications with that reciprocal
// (N reciprocals instead of N(N+1)/2 divisions; the results differ from the dividing form by
// rounding only).
template <typename T, int N>
__device__ __forceinline__ void ge_solve_rcp(T (&A)[N * N], T (&b)[N]) {
    T inv[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        inv[j] = T(1) / A[j + N * j];
        const T ninv = -inv[j];
#pragma unroll
        for (int i = j + 1; i < N; ++i) {
            A[i + N * j] *= ninv;
            b[i] += A[i + N * j] * b[j];
        }
#pragma unroll
        for (int k = j + 1; k < N; ++k)
#pragma unroll
            for (int i = j + 1; i < N; ++i) A[i + N * k] += A[i + N * j] * A[j + N * k];
    }
#pragma unroll
    for (int i = N - 1; i >= 0; --i) {
        b[i] *= inv[i];
#pragma unroll
        for (int j = 0; j < i; ++j) b[j] -= A[j + N * i] * b[i];
    }
}

// Workgroup barrier that orders LDS only: outstanding global loads (prefetches) stay in flight.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

}  // namespace tsw

namespace tsw {

// Dense matrix-vector product psi = M b with M row-major in LDS (M[r*N + c]).
template <typename T, int N>
__device__ __forceinline__ void matvec_lds(const T *__restrict__ M, const T (&b)[N], T (&psi)[N]) {
#pragma unroll
    for (int r = 0; r < N; ++r) {
        T acc = M[r * N] * b[0];
#pragma unroll
        for (int c = 1; c < N; ++c) acc = fmadd(M[r * N + c], b[c], acc);
        psi[r] = acc;
    }
}

}  // namespace tsw
