// Staggered-sweep dispatch and the companion operators on the diagonal-major layout (diag_layout.hpp):
// layout conversions, inflow traces, the moment contraction and the FP64 residual.  The sweep kernel
// itself is in sweep_kernel.hpp.
#include <type_traits>

#include "sweep_kernel.hpp"

namespace tsw {

TSW_SWEEP_EXTERN(double, 1)
TSW_SWEEP_EXTERN(double, 2)
TSW_SWEEP_EXTERN(double, 3)
TSW_SWEEP_EXTERN(double, 4)
TSW_SWEEP_EXTERN(float, 1)
TSW_SWEEP_EXTERN(float, 2)
TSW_SWEEP_EXTERN(float, 3)
TSW_SWEEP_EXTERN(float, 4)

template <typename T>
size_t diag_sweep_lds_bytes(const DiagSweepParams<T> &sp, int p) {
    const int BR = band_rows_of(sp);
    if (BR > 448) return 0;
    // the kernel addresses an ordinate's coefficient planes with 32-bit byte offsets
    if ((long long)sp.nx * sp.ny * (p + 1) * (p + 1) * (long long)sizeof(T) >= (1ll << 32)) return 0;
    if (sp.wave1 && BR > 64) return 0;
    const int cthreads = ((BR + 63) / 64) * 64;
    const int M = sp.ainv ? sp.M : 0;
    const bool direct = (sp.ainv && sp.direct) || sp.wave1;
    if (sp.ainv && sp.direct && sp.nbands > 1) return 0;
    size_t total;
    switch (p) {
        case 0: total = DiagLds<T, 1>(cthreads, sp.nx, M, direct).total; break;
        case 1: total = DiagLds<T, 2>(cthreads, sp.nx, M, direct).total; break;
        case 2: total = DiagLds<T, 3>(cthreads, sp.nx, M, direct).total; break;
        case 3: total = DiagLds<T, 4>(cthreads, sp.nx, M, direct).total; break;
        default: return 0;
    }
    const size_t bytes = total * sizeof(T);
    return bytes <= max_lds_per_block() ? bytes : 0;
}

template <typename T>
void launch_diag_sweep(const DiagSweepParams<T> &sp, int p, hipStream_t s) {
    const size_t lds = diag_sweep_lds_bytes<T>(sp, p);
    if (lds == 0) throw Error(TSW_E_UNSUPPORTED, "hip-staggered needs <= 448 rows per band and its tables within one CU's LDS");
    if (sp.nbands > 1) {
        if (!sp.xtr || !sp.abort_flag || sp.band_rows < 1 || (long long)sp.band_rows * (sp.nbands - 1) >= sp.ny ||
            (long long)sp.band_rows * sp.nbands < sp.ny)
            throw Error(TSW_E_ARG, "inconsistent row-band configuration");
    }
    switch (p) {
        case 0: launch_diag_sweep_p<T, 1>(sp, lds, s); break;
        case 1: launch_diag_sweep_p<T, 2>(sp, lds, s); break;
        case 2: launch_diag_sweep_p<T, 3>(sp, lds, s); break;
        case 3: launch_diag_sweep_p<T, 4>(sp, lds, s); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree " + std::to_string(p) + " not built (0..3)");
    }
}

template <typename U>
__global__ void k_fill_bits(U *p, size_t n, U v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

template <typename T>
void launch_exchange_fill(T *xtr, size_t n, hipStream_t s) {
    using U = typename ExchangeSentinel<T>::U;
    if (n == 0) return;
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_fill_bits<U>, dim3(blocks), dim3(256), 0, s, reinterpret_cast<U *>(xtr), n,
                       ExchangeSentinel<T>::bits);
    TSW_HIP(hipGetLastError());
}
template void launch_exchange_fill<double>(double *, size_t, hipStream_t);
template void launch_exchange_fill<float>(float *, size_t, hipStream_t);

// Batched sweep: one launch over work items {problem, ordinate} that share the DG degree, the mode
// (which sources are present) and the material/elimination choice; `probs` are device copies.
template <typename T>
void launch_diag_sweep_batch(const DiagSweepParams<T> &rep, const DiagSweepParams<T> *probs, const int2 *items,
                             int nitems, int max_ny, size_t lds, int p, hipStream_t s) {
    switch (p) {
        case 0: launch_diag_sweep_batch_p<T, 1>(rep, probs, items, nitems, max_ny, lds, s); break;
        case 1: launch_diag_sweep_batch_p<T, 2>(rep, probs, items, nitems, max_ny, lds, s); break;
        case 2: launch_diag_sweep_batch_p<T, 3>(rep, probs, items, nitems, max_ny, lds, s); break;
        case 3: launch_diag_sweep_batch_p<T, 4>(rep, probs, items, nitems, max_ny, lds, s); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built");
    }
}
template void launch_diag_sweep_batch<double>(const DiagSweepParams<double> &, const DiagSweepParams<double> *,
                                              const int2 *, int, int, size_t, int, hipStream_t);
template void launch_diag_sweep_batch<float>(const DiagSweepParams<float> &, const DiagSweepParams<float> *,
                                             const int2 *, int, int, size_t, int, hipStream_t);

template size_t diag_sweep_lds_bytes<double>(const DiagSweepParams<double> &, int);
template size_t diag_sweep_lds_bytes<float>(const DiagSweepParams<float> &, int);
template void launch_diag_sweep<double>(const DiagSweepParams<double> &, int, hipStream_t);
template void launch_diag_sweep<float>(const DiagSweepParams<float> &, int, hipStream_t);

// =====================================================================================================
// Layout conversions.  Grids are (diagonal s, ordinate or quadrant) with one thread per row r, so the
// diagonal-layout side of every access is contiguous.
// =====================================================================================================
namespace {
inline int rows_block(int ny) { return ny <= 64 ? 64 : (ny <= 128 ? 128 : (ny <= 256 ? 256 : 512)); }
}

// reference psi [i][j][q][k] (interior) <-> diagonal psi [q][k][pos]; dir 0: ref -> diag, 1: diag -> ref.
// A block takes R consecutive cells of one diagonal for the ordinates of one quadrant and transposes
// them through LDS (held in the diagonal precision): on the reference side each cell's ordinates of
// the quadrant are one contiguous run, on the diagonal side each (q, k) plane receives R consecutive
// positions.  Both sides are walked with all threads over the flattened tile.
// SUM (dir 1 only): ref = (TR) diag + (TR) diag2, the tile held in TR.
template <typename TR, typename TD, bool SUM = false>
__global__ __launch_bounds__(256) void k_psi_convert(TR *ref, TD *diag, DiagMap dm, int nq, int N, int qo1, int qo2,
                                                     int qo3, int R, int dir, const TD *diag2 = nullptr) {
    using TT = std::conditional_t<SUM, TR, TD>;
    extern __shared__ unsigned char smem_raw[];
    TT *tile = reinterpret_cast<TT *>(smem_raw);  // [R][nqQ*N]
    const int s = blockIdx.x, Q = blockIdx.z;
    const int r0 = dm.rlo(s) + (int)blockIdx.y * R;
    const int rhi = s < dm.ny - 1 ? s : dm.ny - 1;
    if (r0 > rhi) return;
    const int nr = (rhi - r0 + 1) < R ? (rhi - r0 + 1) : R;
    const int q0 = Q == 0 ? 0 : (Q == 1 ? qo1 : (Q == 2 ? qo2 : qo3));
    const int q1 = Q == 0 ? qo1 : (Q == 1 ? qo2 : (Q == 2 ? qo3 : nq));
    const int run = (q1 - q0) * N;  // values per cell on the reference side
    if (run == 0) return;
    const size_t ncell = (size_t)dm.cells();
    const long long p0 = dm.pos(s, r0);
    const bool xn = quad_xneg(Q), yn = quad_yneg(Q);
    // cells of one diagonal step by (-1, +1) in (c, r): the mesh index moves by a constant stride
    int i0, j0, i1, j1;
    dm.mesh(s - r0, r0, xn, yn, i0, j0);
    dm.mesh(s - r0 - 1, r0 + 1, xn, yn, i1, j1);
    const long long cstride = ((long long)(i1 - i0) * (dm.ny + 2) + (j1 - j0)) * nq * N;
    const gptr<TR> base = (gptr<TR>)(ref + (((size_t)i0 * (dm.ny + 2) + j0) * nq + q0) * N);
    const gptr<TD> dbase = (gptr<TD>)(diag + (size_t)q0 * N * ncell + (size_t)p0);
    const int total = run * nr;
    if (!SUM && dir == 0) {
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e / run, v = e - t * run;
            tile[e] = TT(base[t * cstride + v]);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e % nr, qk = e / nr;  // consecutive threads: consecutive positions
            dbase[(size_t)qk * ncell + t] = TD(tile[t * run + qk]);
        }
    } else {
        const gptr<const TD> dbase2 = SUM ? (gptr<const TD>)(diag2 + (size_t)q0 * N * ncell + (size_t)p0) : nullptr;
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e % nr, qk = e / nr;
            if (SUM) tile[t * run + qk] = TT(dbase[(size_t)qk * ncell + t]) + TT(dbase2[(size_t)qk * ncell + t]);
            else tile[t * run + qk] = TT(dbase[(size_t)qk * ncell + t]);
        }
        __syncthreads();
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e / run, v = e - t * run;
            base[t * cstride + v] = TR(tile[e]);
        }
    }
}

template <typename TR, typename TD>
void launch_psi_convert(const Geom &g, TR *ref, TD *diag, int dir, hipStream_t s) {
    const DiagMap dm{g.nx, g.ny};
    int max_run = 0;
    for (int Q = 0; Q < 4; ++Q) max_run = std::max(max_run, (g.quad_off[Q + 1] - g.quad_off[Q]) * g.N());
    if (max_run == 0) return;
    int R = 64;
    while (R > 1 && (size_t)R * max_run * sizeof(TD) > 48 * 1024) R /= 2;
    const int maxlen = g.nx < g.ny ? g.nx : g.ny;
    const dim3 grid(g.nx + g.ny - 1, (maxlen + R - 1) / R, 4);
    hipLaunchKernelGGL((k_psi_convert<TR, TD>), grid, dim3(256), (size_t)R * max_run * sizeof(TD), s, ref, diag,
                       dm, g.nq, g.N(), g.quad_off[1], g.quad_off[2], g.quad_off[3], R, dir);
    TSW_HIP(hipGetLastError());
}
template void launch_psi_convert<double, double>(const Geom &, double *, double *, int, hipStream_t);
template void launch_psi_convert<double, float>(const Geom &, double *, float *, int, hipStream_t);

void launch_psi_store_sum(const Geom &g, double *ref, const float *a, const float *b, hipStream_t s) {
    const DiagMap dm{g.nx, g.ny};
    int max_run = 0;
    for (int Q = 0; Q < 4; ++Q) max_run = std::max(max_run, (g.quad_off[Q + 1] - g.quad_off[Q]) * g.N());
    if (max_run == 0) return;
    int R = 64;
    while (R > 1 && (size_t)R * max_run * sizeof(double) > 48 * 1024) R /= 2;
    const int maxlen = g.nx < g.ny ? g.nx : g.ny;
    const dim3 grid(g.nx + g.ny - 1, (maxlen + R - 1) / R, 4);
    hipLaunchKernelGGL((k_psi_convert<double, float, true>), grid, dim3(256), (size_t)R * max_run * sizeof(double), s,
                       ref, const_cast<float *>(a), dm, g.nq, g.N(), g.quad_off[1], g.quad_off[2], g.quad_off[3], R, 1,
                       b);
    TSW_HIP(hipGetLastError());
}

// per-cell array [i][j][k] (k < K, halo layout) -> per-quadrant diagonal [Q][k][pos]
template <typename TI, typename TO>
__global__ void k_cell_to_diag(const TI *cell, TO *out, DiagMap dm, int K, const int *skip) {
    if (skip && *skip) return;
    const int s = blockIdx.x, Q = blockIdx.y;
    const int r = dm.rlo(s) + (int)threadIdx.x;
    const int r1 = s < dm.ny - 1 ? s : dm.ny - 1;
    if (r > r1) return;
    int i, j;
    dm.mesh(s - r, r, quad_xneg(Q), quad_yneg(Q), i, j);
    const size_t ncell = (size_t)dm.cells();
    const TI *pc = cell + ((size_t)i * (dm.ny + 2) + j) * K;
    TO *po = out + (size_t)Q * K * ncell + (size_t)dm.pos(s, r);
    for (int k = 0; k < K; ++k) po[(size_t)k * ncell] = TO(pc[k]);
}

template <typename TI, typename TO>
void launch_cell_to_diag(const Geom &g, const TI *cell, TO *out, int K, hipStream_t s, const int *skip) {
    const DiagMap dm{g.nx, g.ny};
    hipLaunchKernelGGL((k_cell_to_diag<TI, TO>), dim3(g.nx + g.ny - 1, 4), dim3(rows_block(g.ny)), 0, s, cell, out, dm,
                       K, skip);
    TSW_HIP(hipGetLastError());
}
template void launch_cell_to_diag<double, double>(const Geom &, const double *, double *, int, hipStream_t, const int *);
template void launch_cell_to_diag<double, float>(const Geom &, const double *, float *, int, hipStream_t, const int *);
template void launch_cell_to_diag<int32_t, int32_t>(const Geom &, const int32_t *, int32_t *, int, hipStream_t,
                                                    const int *);

// Inflow face traces of the halo cells of a reference-layout psi: tin_x[q][r][P] (x-inflow of row r)
// and tin_y[q][c][P] (y-inflow of column c), as the sweep consumes them.
template <typename TO, int P>
__global__ void k_inflow(const double *ref, TO *tin_x, TO *tin_y, int nx, int ny, int nq, int qo1, int qo2, int qo3) {
    constexpr int N = P * P;
    const int q = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int Q = q >= qo3 ? 3 : (q >= qo2 ? 2 : (q >= qo1 ? 1 : 0));
    const bool xneg = quad_xneg(Q), yneg = quad_yneg(Q);
    if (t < ny) {  // row r = t
        const int j = yneg ? ny - t : t + 1, i = xneg ? nx + 1 : 0;
        double tr[P];
        trace_x<double, P>(ref + (((size_t)i * (ny + 2) + j) * nq + q) * N, xneg, tr);
#pragma unroll
        for (int d = 0; d < P; ++d) tin_x[((size_t)q * ny + t) * P + d] = TO(tr[d]);
    }
    if (t < nx) {  // column c = t
        const int i = xneg ? nx - t : t + 1, j = yneg ? ny + 1 : 0;
        double tr[P];
        trace_y<double, P>(ref + (((size_t)i * (ny + 2) + j) * nq + q) * N, yneg, tr);
#pragma unroll
        for (int d = 0; d < P; ++d) tin_y[((size_t)q * nx + t) * P + d] = TO(tr[d]);
    }
}

template <typename TO>
void launch_inflow(const Geom &g, const double *ref, TO *tin_x, TO *tin_y, hipStream_t s) {
    const int n = g.nx > g.ny ? g.nx : g.ny;
    const dim3 grid((n + 127) / 128, g.nq), block(128);
    const int a = g.quad_off[1], b = g.quad_off[2], c = g.quad_off[3];
    switch (g.p) {
        case 0: hipLaunchKernelGGL((k_inflow<TO, 1>), grid, block, 0, s, ref, tin_x, tin_y, g.nx, g.ny, g.nq, a, b, c); break;
        case 1: hipLaunchKernelGGL((k_inflow<TO, 2>), grid, block, 0, s, ref, tin_x, tin_y, g.nx, g.ny, g.nq, a, b, c); break;
        case 2: hipLaunchKernelGGL((k_inflow<TO, 3>), grid, block, 0, s, ref, tin_x, tin_y, g.nx, g.ny, g.nq, a, b, c); break;
        case 3: hipLaunchKernelGGL((k_inflow<TO, 4>), grid, block, 0, s, ref, tin_x, tin_y, g.nx, g.ny, g.nq, a, b, c); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built");
    }
    TSW_HIP(hipGetLastError());
}
template void launch_inflow<double>(const Geom &, const double *, double *, double *, hipStream_t);
template void launch_inflow<float>(const Geom &, const double *, float *, float *, hipStream_t);

// =====================================================================================================
// Moment contraction on the diagonal layout: per quadrant, phiQ[Q][k][pos] = sum_{q in Q} w_q psi[q][k][pos]
// (contiguous over pos), then phi[i][j][k] = beta phi + alpha sum_Q phiQ[Q][k][posQ(i,j)]
// (TransportOperator.cpp:140-158).
// =====================================================================================================
template <typename TI>
__global__ __launch_bounds__(256) void k_pmv_quadrant(const TI *__restrict__ psi, const double *__restrict__ w,
                                                      double *__restrict__ phiQ, size_t ncell, int N, int qo1,
                                                      int qo2, int qo3, int nq, const int *skip) {
    if (skip && *skip) return;
    const int Q = blockIdx.y;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * N) return;
    const int q0 = Q == 0 ? 0 : (Q == 1 ? qo1 : (Q == 2 ? qo2 : qo3));
    const int q1 = Q == 0 ? qo1 : (Q == 1 ? qo2 : (Q == 2 ? qo3 : nq));
    double acc = 0.0;
    for (int q = q0; q < q1; ++q) acc += w[q] * (double)psi[(size_t)q * N * ncell + t];
    phiQ[(size_t)Q * N * ncell + t] = acc;
}

// phi = beta xb + alpha sum_Q phiQ on the interior (xb null: xb = phi); halos of phi are not written
__global__ void k_phi_from_quadrants(const double *__restrict__ phiQ, double *phi, const double *xb, DiagMap dm, int N,
                                     double alpha, double beta, const int *skip) {
    if (skip && *skip) return;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ncell = (size_t)dm.cells();
    if (t >= ncell * N) return;
    const int k = (int)(t % N);
    const size_t cidx = t / N;
    const int i = 1 + (int)(cidx / dm.ny), j = 1 + (int)(cidx % dm.ny);
    double acc = 0.0;
    for (int Q = 0; Q < 4; ++Q)
        acc += phiQ[(size_t)Q * N * ncell + (size_t)k * ncell + (size_t)dm.pos_of_cell(i, j, quad_xneg(Q), quad_yneg(Q))];
    const size_t at = ((size_t)i * (dm.ny + 2) + j) * N + k;
    phi[at] = (beta == 0.0 ? 0.0 : beta * (xb ? xb[at] : phi[at])) + alpha * acc;
}

template <typename TI>
void launch_pmv_diag(const Geom &g, const TI *psi, const double *w, double *phiQ, double *phi, double alpha, double beta,
                     hipStream_t s, const int *skip, const double *xb) {
    const DiagMap dm{g.nx, g.ny};
    const size_t n = (size_t)dm.cells() * g.N();
    hipLaunchKernelGGL((k_pmv_quadrant<TI>), dim3((unsigned)((n + 255) / 256), 4), dim3(256), 0, s, psi, w, phiQ,
                       (size_t)dm.cells(), g.N(), g.quad_off[1], g.quad_off[2], g.quad_off[3], g.nq, skip);
    hipLaunchKernelGGL(k_phi_from_quadrants, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, phiQ, phi, xb, dm,
                       g.N(), alpha, beta, skip);
    TSW_HIP(hipGetLastError());
}
template void launch_pmv_diag<double>(const Geom &, const double *, const double *, double *, double *, double, double,
                                      hipStream_t, const int *, const double *);
template void launch_pmv_diag<float>(const Geom &, const float *, const double *, double *, double *, double, double,
                                     hipStream_t, const int *, const double *);

// =====================================================================================================
// FP64 angular residual on the diagonal layout, the refinement's r = Q + sigma_s phi/(4 pi) - L psi
// (Lmv, TransportOperator.cpp:1576-1658, written with face traces), stored in TO; phi_d null: no
// scattering term.
// psi's inflow comes from tin_x/tin_y (traces of the halo cells).  Grid (diagonal, ordinate).
// =====================================================================================================
template <typename TO, typename TP, int P>
__global__ void k_residual_diag(const double *__restrict__ qsrc, const double *__restrict__ phi_d,
                                const double *__restrict__ sigs_d, const double *__restrict__ sigt_d,
                                const TP *__restrict__ psi, const double *__restrict__ tin_x,
                                const double *__restrict__ tin_y, const double *__restrict__ xi_,
                                const double *__restrict__ eta_, TO *__restrict__ res, DiagMap dm, int nq, int qo1,
                                int qo2, int qo3, double dx, double dy, double inv_dt) {
    constexpr int N = P * P;
    const int s = blockIdx.x, q = blockIdx.y;
    const int r = dm.rlo(s) + (int)threadIdx.x;
    const int r1 = s < dm.ny - 1 ? s : dm.ny - 1;
    if (r > r1) return;
    const int c = s - r;
    const int Q = q >= qo3 ? 3 : (q >= qo2 ? 2 : (q >= qo1 ? 1 : 0));
    const size_t ncell = (size_t)dm.cells();
    const double xi = xi_[q], eta = eta_[q];
    const bool xneg = xi <= 0.0, yneg = eta <= 0.0;
    const size_t p = (size_t)dm.pos(s, r);
    const TP *pq = psi + (size_t)q * N * ncell;
    double x[N];
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = (double)pq[(size_t)k * ncell + p];
    double tx[P], ty[P];
    if (c == 0) {
#pragma unroll
        for (int d = 0; d < P; ++d) tx[d] = tin_x[((size_t)q * dm.ny + r) * P + d];
    } else {
        const size_t pu = (size_t)dm.pos(s - 1, r);
        double u[N];
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = (double)pq[(size_t)k * ncell + pu];
        trace_x<double, P>(u, xneg, tx);
    }
    if (r == 0) {
#pragma unroll
        for (int d = 0; d < P; ++d) ty[d] = tin_y[((size_t)q * dm.nx + c) * P + d];
    } else {
        const size_t pu = (size_t)dm.pos(s - 1, r - 1);
        double u[N];
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = (double)pq[(size_t)k * ncell + pu];
        trace_y<double, P>(u, yneg, ty);
    }
    // (L + 1/dt) x = (Aq x + (sigma + 1/dt) dx dy x - Kx tx - Ky ty) / (dx dy): the sweep's element
    // system with the inflow terms moved back to the left
    double A[N * N];
    const size_t pQ = (size_t)Q * ncell + p;
    assemble_matrix<double, P>(A, xi, eta, dx, dy, (sigt_d[pQ] + inv_dt) * dx * dy);
    double zero[N];
#pragma unroll
    for (int k = 0; k < N; ++k) zero[k] = 0.0;
    double inflow[N];
    rhs_from_traces<double, P>(inflow, zero, tx, ty, xi, eta, dx, dy);
    const double inv_area = 1.0 / (dx * dy);
    TO *pr = res + (size_t)q * N * ncell + p;
#pragma unroll
    for (int row = 0; row < N; ++row) {
        double lx = 0.0;
#pragma unroll
        for (int col = 0; col < N; ++col) lx += A[row + N * col] * x[col];
        lx = (lx - inflow[row]) * inv_area;
        double v = qsrc[(size_t)q * N * ncell + (size_t)row * ncell + p];
        if (phi_d) v += sigs_d[pQ] * phi_d[(size_t)Q * N * ncell + (size_t)row * ncell + p] * kInv4Pi;
        pr[(size_t)row * ncell] = TO(v - lx);
    }
}

template <typename TO, typename TP>
void launch_residual_diag(const Geom &g, const double *qsrc, const double *phi_d, const double *sigs_d,
                          const double *sigt_d, const TP *psi, const double *tin_x, const double *tin_y,
                          const double *xi, const double *eta, double inv_dt, TO *res, hipStream_t s) {
    const DiagMap dm{g.nx, g.ny};
    const dim3 grid(g.nx + g.ny - 1, g.nq), block(rows_block(g.ny));
    const int a = g.quad_off[1], b = g.quad_off[2], c = g.quad_off[3];
#define TSW_RES(PP)                                                                                                   \
    hipLaunchKernelGGL((k_residual_diag<TO, TP, PP>), grid, block, 0, s, qsrc, phi_d, sigs_d, sigt_d, psi, tin_x, tin_y, \
                       xi, eta, res, dm, g.nq, a, b, c, g.dx, g.dy, inv_dt)
    switch (g.p) {
        case 0: TSW_RES(1); break;
        case 1: TSW_RES(2); break;
        case 2: TSW_RES(3); break;
        case 3: TSW_RES(4); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built");
    }
#undef TSW_RES
    TSW_HIP(hipGetLastError());
}
template void launch_residual_diag<float, double>(const Geom &, const double *, const double *, const double *,
                                                  const double *, const double *, const double *, const double *,
                                                  const double *, const double *, double, float *, hipStream_t);
template void launch_residual_diag<double, double>(const Geom &, const double *, const double *, const double *,
                                                   const double *, const double *, const double *, const double *,
                                                   const double *, const double *, double, double *, hipStream_t);
template void launch_residual_diag<float, float>(const Geom &, const double *, const double *, const double *,
                                                 const double *, const float *, const double *, const double *,
                                                 const double *, const double *, double, float *, hipStream_t);

}  // namespace tsw
