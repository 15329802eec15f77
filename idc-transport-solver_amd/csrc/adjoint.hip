// Kernels of the adjoint (vector-Jacobian) path.
//
// The upwind-DG sweep solves M(Omega) psi = dx dy src, where M = W Mt(Omega) with W = diag((2d+1)(2e+1))
// the inverse Legendre mass scaling and Mt the weak-form operator.  Integration by parts gives
// Mt(Omega)^T = Mt(-Omega), so M(Omega)^{-T} = W^{-1} M(-Omega)^{-1} W: the transposed sweep of an ordinate
// is the ordinary sweep of the opposite ordinate applied to W-scaled data and unscaled afterwards.  The
// library runs it on a second problem whose ordinate set is the negated set in reverse order (reversal
// keeps the quadrant blocks ordered 0..3).  These kernels move data between the two orderings and form
// the gradients from the forward and adjoint fluxes.
#include "tsw_internal.hpp"

namespace tsw {

// out[cell][q][k] = in[cell][nq-1-q][k] * kscale(k) * qscale, interior cells only (halos of out are not
// touched).  kmode 1 multiplies by (2d+1)(2e+1), 2 divides by it; qmode 1 multiplies by w, 2 divides,
// with w taken at the output ordinate (w_of_out) or at the input ordinate.
__global__ void k_reverse_ordinates(const double *__restrict__ in, double *__restrict__ out, int nx, int ny, int nq,
                                    int P, int kmode, const double *__restrict__ w, int qmode, int w_of_out) {
    const int N = P * P;
    const size_t per_cell = (size_t)nq * N;
    const size_t total = (size_t)nx * ny * per_cell;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < total; t += (size_t)gridDim.x * blockDim.x) {
        const size_t cell = t / per_cell;
        const int m = (int)(t - cell * per_cell);
        const int q = m / N, k = m - q * N;
        const int ci = (int)(cell / ny) + 1, cj = (int)(cell % ny) + 1;
        const size_t base = ((size_t)ci * (ny + 2) + cj) * per_cell;
        double v = in[base + (size_t)(nq - 1 - q) * N + k];
        const int d = k / P, e = k - d * P;
        const double wk = (double)((2 * d + 1) * (2 * e + 1));
        if (kmode == 1) v *= wk;
        else if (kmode == 2) v /= wk;
        if (qmode) {
            const double wq = w[w_of_out ? q : nq - 1 - q];
            v = qmode == 1 ? v * wq : v / wq;
        }
        out[base + (size_t)q * N + k] = v;
    }
}

void launch_reverse_ordinates(const Geom &g, const double *in, double *out, int kmode, const double *w, int qmode,
                              bool w_of_out, hipStream_t s) {
    if (in == out) throw Error(TSW_E_ARG, "ordinate reversal needs distinct buffers");
    const size_t total = (size_t)g.nx * g.ny * g.nq * g.N();
    const int threads = 256;
    const int blocks = (int)std::min<size_t>((total + threads - 1) / threads, 65536);
    hipLaunchKernelGGL(k_reverse_ordinates, dim3(blocks), dim3(threads), 0, s, in, out, g.nx, g.ny, g.nq, g.p + 1,
                       kmode, w, qmode, (int)w_of_out);
    TSW_HIP(hipGetLastError());
}

// Per interior cell c (one wavefront per cell), written or (accumulate) added:
//   gst[c] (+)= -scale sum_{q,k} zeta psi               (d/dsigma_t of <seed, psi>)
//   gss[c] (+)=  scale sum_{q,k} zeta_qk phi_k / (4 pi)  (d/dsigma_s, when phi is given)
// psi, zeta: reference layout; phi: phi layout; gst/gss: cell arrays (halo entries not written).
__global__ __launch_bounds__(256) void k_cell_grads(const double *__restrict__ psi, const double *__restrict__ zeta,
                                                    const double *__restrict__ phi, double *__restrict__ gst,
                                                    double *__restrict__ gss, int nx, int ny, int nq, int N,
                                                    double scale, int accumulate) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const long long cell = (long long)blockIdx.x * 4 + wave;
    if (cell >= (long long)nx * ny) return;  // whole wavefront leaves; no block barrier follows
    const int ci = (int)(cell / ny) + 1, cj = (int)(cell % ny) + 1;
    const size_t c = (size_t)ci * (ny + 2) + cj;
    const int per = nq * N;
    const size_t base = c * per;
    double st = 0.0, ss = 0.0;
    for (int m = lane; m < per; m += 64) {
        const double z = zeta[base + m];
        if (gst) st += z * psi[base + m];
        if (gss) ss += z * phi[c * N + m % N];
    }
    for (int off = 32; off > 0; off >>= 1) {
        st += __shfl_down(st, off);
        ss += __shfl_down(ss, off);
    }
    if (lane == 0) {
        if (gst) gst[c] = (accumulate ? gst[c] : 0.0) - scale * st;
        if (gss) gss[c] = (accumulate ? gss[c] : 0.0) + scale * ss * (1.0 / (4.0 * 3.14159265358979323846));
    }
}

void launch_cell_grads(const Geom &g, const double *psi, const double *zeta, const double *phi, double *gst,
                       double *gss, hipStream_t s, double scale, bool accumulate) {
    const long long cells = (long long)g.nx * g.ny;
    const int blocks = (int)((cells + 3) / 4);
    hipLaunchKernelGGL(k_cell_grads, dim3(blocks), dim3(256), 0, s, psi, zeta, phi, gst, gss, g.nx, g.ny, g.nq, g.N(),
                       scale, (int)accumulate);
    TSW_HIP(hipGetLastError());
}

// Transpose of the scattering source S P on the interior: out_q += alpha w_q sigma_s sum_q' y_q' / (4 pi).
// One thread per (interior cell, coefficient).
__global__ void k_kt_apply(const double *__restrict__ y, double *__restrict__ out, const double *__restrict__ ss,
                           const double *__restrict__ w, double alpha, int nx, int ny, int nq, int N) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)nx * ny * N) return;
    const size_t c = t / N;
    const int k = (int)(t - c * N);
    const size_t cell = (size_t)(1 + c / ny) * (ny + 2) + 1 + c % ny;
    const size_t base = cell * nq * N + k;
    double sum = 0.0;
    for (int q = 0; q < nq; ++q) sum += y[base + (size_t)q * N];
    const double f = alpha * ss[cell] * sum * (1.0 / (4.0 * 3.14159265358979323846));
    for (int q = 0; q < nq; ++q) out[base + (size_t)q * N] += f * w[q];
}

void launch_kt_apply(const Geom &g, const double *y, double *out, const double *ss, const double *w, double alpha,
                     hipStream_t s) {
    const size_t n = (size_t)g.nx * g.ny * g.N();
    hipLaunchKernelGGL(k_kt_apply, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, y, out, ss, w, alpha, g.nx, g.ny,
                       g.nq, g.N());
    TSW_HIP(hipGetLastError());
}

// Gradient with respect to the inflow halo values, added into g's halo cells: for ordinate q the halo
// column beside its x-inflow face receives Cx^T zeta_b / (dx dy) and the halo row beside its y-inflow face
// Cy^T zeta_b / (dx dy), zeta_b the adjacent boundary cell's adjoint flux and
//   Cx[(d,e),(a,e)] = dy xi (2d+1) kx(d,a),  kx = (-1)^(a+1) if xi <= 0 else (-1)^d   (y alike),
// the coupling through which the forward sweep reads the halo (SweepLinearSolveSpec.cpp:545-598).
// One thread per (ordinate, face cell, output coefficient); corners are not inflow for any face.
__global__ void k_inflow_grad(const double *__restrict__ zeta, double *__restrict__ g, const double *__restrict__ xi,
                              const double *__restrict__ eta, int nx, int ny, int nq, int P, double dx, double dy) {
    const int N = P * P;
    const long long per_q = (long long)(nx + ny) * N;
    const long long total = per_q * nq;
    const double inv_area = 1.0 / (dx * dy);
    for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (long long)gridDim.x * blockDim.x) {
        const int q = (int)(t / per_q);
        const int r = (int)(t - (long long)q * per_q);
        const int f = r / N, k = r - f * N;
        const int hi = k / P, lo = k - hi * P;  // output coefficient (d,e) = (hi,lo)
        auto at = [&](int i, int j) { return (((size_t)i * (ny + 2) + j) * nq + q) * N; };
        if (f < ny) {
            const int j = f + 1;
            const bool xneg = xi[q] <= 0.0;
            const double cx = dy * xi[q];
            const int ih = xneg ? nx + 1 : 0, ib = xneg ? nx : 1;
            const double *zb = zeta + at(ib, j);
            double v = 0.0;
            for (int d = 0; d < P; ++d) {
                const double kx = xneg ? ((hi + 1) & 1 ? -1.0 : 1.0) : (d & 1 ? -1.0 : 1.0);
                v += cx * (2 * d + 1) * kx * zb[d * P + lo];
            }
            g[at(ih, j) + k] += v * inv_area;
        } else {
            const int i = f - ny + 1;
            const bool yneg = eta[q] <= 0.0;
            const double cy = dx * eta[q];
            const int jh = yneg ? ny + 1 : 0, jb = yneg ? ny : 1;
            const double *zb = zeta + at(i, jb);
            double v = 0.0;
            for (int e = 0; e < P; ++e) {
                const double ky = yneg ? ((lo + 1) & 1 ? -1.0 : 1.0) : (e & 1 ? -1.0 : 1.0);
                v += cy * (2 * e + 1) * ky * zb[hi * P + e];
            }
            g[at(i, jh) + k] += v * inv_area;
        }
    }
}

void launch_inflow_grad(const Geom &g, const double *zeta, double *out, const double *xi, const double *eta,
                        hipStream_t s) {
    const long long total = (long long)g.nq * (g.nx + g.ny) * g.N();
    const int threads = 256;
    const int blocks = (int)std::min<long long>((total + threads - 1) / threads, 65536);
    hipLaunchKernelGGL(k_inflow_grad, dim3(blocks), dim3(threads), 0, s, zeta, out, xi, eta, g.nx, g.ny, g.nq, g.p + 1,
                       g.dx, g.dy);
    TSW_HIP(hipGetLastError());
}

}  // namespace tsw
