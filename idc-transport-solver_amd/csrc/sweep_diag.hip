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

// LDS of the lane-split sweep for bands of BR rows (0: not built for (T, P) or too many rows)
template <typename T, int P>
size_t split_lds_total(int BR, int nx, int M) {
    if constexpr (!split_supported<T, P>()) {
        return 0;
    } else {
        constexpr int LG = split_lanes<P>(), rpw = 64 / LG;
        const int slots = (BR + rpw - 1) / rpw * rpw;
        if (slots * LG > 512) return 0;
        return SplitLds<T, P>(slots, nx, M).total;
    }
}

template <typename T>
size_t split_sweep_lds_bytes(int p, int rows, int nx, int M) {
    size_t total = 0;
    switch (p) {
        case 1: total = split_lds_total<T, 2>(rows, nx, M); break;
        case 2: total = split_lds_total<T, 3>(rows, nx, M); break;
        case 3: total = split_lds_total<T, 4>(rows, nx, M); break;
        default: break;
    }
    const size_t bytes = total * sizeof(T);
    return bytes <= max_lds_per_block() ? bytes : 0;
}
template size_t split_sweep_lds_bytes<double>(int, int, int, int);
template size_t split_sweep_lds_bytes<float>(int, int, int, int);

template <typename T>
size_t diag_sweep_lds_bytes(const DiagSweepParams<T> &sp, int p) {
    const int BR = band_rows_of(sp);
    if (BR > 448) return 0;
    // the kernel addresses an ordinate's coefficient planes with 32-bit byte offsets
    if ((long long)sp.nx * sp.ny * (p + 1) * (p + 1) * (long long)sizeof(T) >= (1ll << 32)) return 0;
    if (sp.wave1 && BR > 64) return 0;
    if (sp.split) {
        if (!sp.ainv || sp.wave1) return 0;
        size_t total = 0;
        switch (p) {
            case 1: total = split_lds_total<T, 2>(BR, sp.nx, sp.M); break;
            case 2: total = split_lds_total<T, 3>(BR, sp.nx, sp.M); break;
            case 3: total = split_lds_total<T, 4>(BR, sp.nx, sp.M); break;
            default: break;
        }
        const size_t bytes = total * sizeof(T);
        return bytes <= max_lds_per_block() ? bytes : 0;
    }
    const int cthreads = ((BR + 63) / 64) * 64;
    const int M = sp.ainv ? sp.M : 0;
    const bool direct = (sp.ainv && sp.direct) || sp.wave1;
    const int stage_w = sp.qref && sp.ref_stage ? cthreads / 64 : 0;  // compute waves with a source stage
    if (sp.ainv && sp.direct && sp.nbands > 1) return 0;
    size_t total;
    switch (p) {
        case 0: total = DiagLds<T, 1>(cthreads, sp.nx, M, direct, stage_w).total; break;
        case 1: total = DiagLds<T, 2>(cthreads, sp.nx, M, direct, stage_w).total; break;
        case 2: total = DiagLds<T, 3>(cthreads, sp.nx, M, direct, stage_w).total; break;
        case 3: total = DiagLds<T, 4>(cthreads, sp.nx, M, direct, stage_w).total; break;
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
template <typename T>
void launch_diag_inflow_batch(const DiagSweepParams<T> *probs, int nprob, int p, hipStream_t s) {
    switch (p) {
        case 0: launch_diag_inflow_batch_p<T, 1>(probs, nprob, s); break;
        case 1: launch_diag_inflow_batch_p<T, 2>(probs, nprob, s); break;
        case 2: launch_diag_inflow_batch_p<T, 3>(probs, nprob, s); break;
        case 3: launch_diag_inflow_batch_p<T, 4>(probs, nprob, s); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built");
    }
}
template void launch_diag_inflow_batch<double>(const DiagSweepParams<double> *, int, int, hipStream_t);
template void launch_diag_inflow_batch<float>(const DiagSweepParams<float> *, int, int, hipStream_t);
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
// them through LDS: on the reference side each cell's ordinates of the quadrant are one contiguous run,
// on the diagonal side each (q, k) plane receives R consecutive positions.  Both sides are walked with
// all threads over the flattened tile, four independent accesses in flight per thread.
// dir 0: diag = (TD) ref, and d32 (if not null) = (float) ref as well.  dir 1: ref = (TR) diag, or with
// diag2 ref = (TR) diag + (TR) diag2; the tile is held in TT (double when two outputs or a sum need it).
struct ConvJob {
    double *ref;
    void *diag;         // diagonal array, TD
    const void *diag2;  // second summand of a store, TD, or null
    float *d32;         // second FP32 output of a load, or null
    int nx, ny, nq, N, qo1, qo2, qo3;
    int R;  // cells per block
};

template <typename TR, typename TD, typename TT>
__device__ __forceinline__ void convert_tile(const ConvJob &J, int s, int chunk, int Q, int dir, TT *tile) {
    const DiagMap dm{J.nx, J.ny};
    const int R = J.R;
    const int r0 = dm.rlo(s) + chunk * R;
    const int rhi = s < dm.ny - 1 ? s : dm.ny - 1;
    if (r0 > rhi) return;
    const int nr = (rhi - r0 + 1) < R ? (rhi - r0 + 1) : R;
    const int q0 = Q == 0 ? 0 : (Q == 1 ? J.qo1 : (Q == 2 ? J.qo2 : J.qo3));
    const int q1 = Q == 0 ? J.qo1 : (Q == 1 ? J.qo2 : (Q == 2 ? J.qo3 : J.nq));
    const int N = J.N, nq = J.nq;
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
    const gptr<TR> base = (gptr<TR>)(J.ref + (((size_t)i0 * (dm.ny + 2) + j0) * nq + q0) * N);
    const size_t doff = (size_t)q0 * N * ncell + (size_t)p0;
    const gptr<TD> dbase = (gptr<TD>)((TD *)J.diag + doff);
    const int total = run * nr;
    if (dir == 0) {
#pragma unroll 4
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e / run, v = e - t * run;
            tile[e] = TT(base[t * cstride + v]);
        }
        __syncthreads();
        const gptr<float> d32 = J.d32 ? (gptr<float>)(J.d32 + doff) : nullptr;
#pragma unroll 4
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e % nr, qk = e / nr;  // consecutive threads: consecutive positions
            const TT v = tile[t * run + qk];
            dbase[(size_t)qk * ncell + t] = TD(v);
            if (d32) d32[(size_t)qk * ncell + t] = float(v);
        }
    } else {
        const gptr<const TD> dbase2 = J.diag2 ? (gptr<const TD>)((const TD *)J.diag2 + doff) : nullptr;
#pragma unroll 4
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e % nr, qk = e / nr;
            TT v = TT(dbase[(size_t)qk * ncell + t]);
            if (dbase2) v += TT(dbase2[(size_t)qk * ncell + t]);
            tile[t * run + qk] = v;
        }
        __syncthreads();
#pragma unroll 4
        for (int e = threadIdx.x; e < total; e += blockDim.x) {
            const int t = e / run, v = e - t * run;
            base[t * cstride + v] = TR(tile[e]);
        }
    }
}

template <typename TR, typename TD, typename TT>
__global__ __launch_bounds__(1024) void k_psi_convert(ConvJob J, int dir) {
    extern __shared__ unsigned char smem_raw[];
    convert_tile<TR, TD, TT>(J, blockIdx.x, blockIdx.y, blockIdx.z, dir, reinterpret_cast<TT *>(smem_raw));
}

namespace {
// values per cell of the largest quadrant
int max_run_of(const Geom &g) {
    int m = 0;
    for (int Q = 0; Q < 4; ++Q) m = std::max(m, (g.quad_off[Q + 1] - g.quad_off[Q]) * g.N());
    return m;
}
// Cells per block: the largest power of two (<= 64) whose tile fits the direction's LDS budget.  Long
// runs of consecutive positions on the diagonal side pay off most when loading (measured on the
// 280x280 S12 Q2 lattice, FP64: 280 us with 64-cell tiles and 1024 threads against 353 us with 16-cell
// tiles; the store runs at 244 us with 32-cell tiles and 512 threads).
int conv_rows(int run, size_t tbytes, int dir) {
    const size_t budget = (dir == 0 ? 96 : 48) * 1024;
    int R = 64;
    while (R > 1 && (size_t)R * run * tbytes > budget) R /= 2;
    return R;
}
ConvJob conv_job(const Geom &g, double *ref, void *diag, const void *diag2, float *d32, int R) {
    ConvJob J{};
    J.ref = ref;
    J.diag = diag;
    J.diag2 = diag2;
    J.d32 = d32;
    J.nx = g.nx;
    J.ny = g.ny;
    J.nq = g.nq;
    J.N = g.N();
    J.qo1 = g.quad_off[1];
    J.qo2 = g.quad_off[2];
    J.qo3 = g.quad_off[3];
    J.R = R;
    return J;
}
template <typename TR, typename TD, typename TT>
void launch_convert_one(const Geom &g, double *ref, void *diag, const void *diag2, float *d32, int dir, hipStream_t s) {
    const int max_run = max_run_of(g);
    if (max_run == 0) return;
    const int R = conv_rows(max_run, sizeof(TT), dir);
    const ConvJob J = conv_job(g, ref, diag, diag2, d32, R);
    const int maxlen = g.nx < g.ny ? g.nx : g.ny;
    const dim3 grid(g.nx + g.ny - 1, (maxlen + R - 1) / R, 4);
    const size_t lds = (size_t)R * max_run * sizeof(TT);
    auto kern = k_psi_convert<TR, TD, TT>;
    if (lds > 64 * 1024)
        TSW_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, grid, dim3(dir == 0 ? 1024 : 512), lds, s, J, dir);
    TSW_HIP(hipGetLastError());
}
}  // namespace

template <typename TR, typename TD>
void launch_psi_convert(const Geom &g, TR *ref, TD *diag, int dir, hipStream_t s) {
    launch_convert_one<TR, TD, TD>(g, ref, diag, nullptr, nullptr, dir, s);
}
template void launch_psi_convert<double, double>(const Geom &, double *, double *, int, hipStream_t);
template void launch_psi_convert<double, float>(const Geom &, double *, float *, int, hipStream_t);

void launch_psi_load_dual(const Geom &g, const double *ref, double *d64, float *d32, hipStream_t s) {
    launch_convert_one<double, double, double>(g, const_cast<double *>(ref), d64, nullptr, d32, 0, s);
}

void launch_psi_store_sum(const Geom &g, double *ref, const float *a, const float *b, hipStream_t s) {
    launch_convert_one<double, float, double>(g, ref, const_cast<float *>(a), b, nullptr, 1, s);
}


// Square tiles of kT x kT cells for the transfers between the cell layout [i][j][k] and the four
// quadrant diagonal layouts [Q][k][pos].  In each quadrant's sweep coordinates (c, r) a tile is again a
// kT x kT square, whose cells lie on 2 kT - 1 diagonals as runs of consecutive positions; element e of
// the tile is placed so that consecutive e walk those runs (tile_cell), and the cell-layout side is read
// or written one tile row (kT cells x K values, contiguous) at a time.
constexpr int kT = 16;

__device__ __forceinline__ void tile_cell(int e, int &a, int &b) {
    constexpr int H = kT * (kT + 1) / 2;  // cells on the tile diagonals 0 .. kT-1
    const bool hi = e >= H;
    const int f = hi ? kT * kT - 1 - e : e;
    int d = (int)((sqrtf(8.0f * (float)f + 1.0f) - 1.0f) * 0.5f);
    if ((d + 1) * (d + 2) / 2 <= f) ++d;
    if (d * (d + 1) / 2 > f) --d;
    const int k = f - d * (d + 1) / 2;
    a = d - k;  // offset in c
    b = k;      // offset in r
    if (hi) {   // the upper diagonals, as the mirror image of the lower ones
        a = kT - 1 - a;
        b = kT - 1 - b;
    }
}

// phi = beta xb + alpha sum_Q phiQ on the interior, one tile per block (xb null: xb = phi)
__global__ __launch_bounds__(kT *kT) void k_phi_from_quadrants_tiled(const double *__restrict__ phiQ, double *phi,
                                                                     const double *xb, DiagMap dm, int N,
                                                                     double alpha, double beta, const int *skip) {
    if (skip && *skip) return;
    extern __shared__ double acc[];  // [kT][kT][N], tile rows in i
    const int i0 = 1 + (int)blockIdx.x * kT, j0 = 1 + (int)blockIdx.y * kT;
    const int e = threadIdx.x;
    for (int v = e; v < kT * kT * N; v += kT * kT) acc[v] = 0.0;
    __syncthreads();
    const size_t ncell = (size_t)dm.cells();
    int a, b;
    tile_cell(e, a, b);
    for (int Q = 0; Q < 4; ++Q) {
        const bool xn = quad_xneg(Q), yn = quad_yneg(Q);
        int c0, r0;  // sweep coordinates of the tile's lowest corner in this quadrant
        dm.local(xn ? i0 + kT - 1 : i0, yn ? j0 + kT - 1 : j0, xn, yn, c0, r0);
        const int c = c0 + a, r = r0 + b;
        int i, j;
        dm.mesh(c, r, xn, yn, i, j);
        if (c >= 0 && r >= 0 && c < dm.nx && r < dm.ny && i >= i0 && i < i0 + kT && j >= j0 && j < j0 + kT) {
            const double *src = phiQ + (size_t)Q * N * ncell + (size_t)dm.pos(c + r, r);
            double *dst = acc + ((size_t)(i - i0) * kT + (j - j0)) * N;
            for (int k = 0; k < N; ++k) dst[k] += src[(size_t)k * ncell];
        }
        __syncthreads();
    }
    const int ilim = dm.nx - i0 + 1 < kT ? dm.nx - i0 + 1 : kT, jlim = dm.ny - j0 + 1 < kT ? dm.ny - j0 + 1 : kT;
    const int row = jlim * N;  // values of one tile row, contiguous in phi
    for (int v = e; v < ilim * row; v += kT * kT) {
        const int ti = v / row, w = v - ti * row;
        const size_t at = ((size_t)(i0 + ti) * (dm.ny + 2) + j0) * N + w;
        const double s = acc[(size_t)ti * kT * N + w];
        phi[at] = (beta == 0.0 ? 0.0 : beta * (xb ? xb[at] : phi[at])) + alpha * s;
    }
}

// per-cell array [i][j][k] (halo layout) -> per-quadrant diagonal [Q][k][pos], one tile per block
template <typename TI, typename TO>
__global__ __launch_bounds__(kT *kT) void k_cell_to_diag_tiled(const TI *__restrict__ cell, TO *__restrict__ out,
                                                               DiagMap dm, int K, const int *skip) {
    if (skip && *skip) return;
    extern __shared__ double buf[];  // [kT][kT][K] as TI
    TI *tile = reinterpret_cast<TI *>(buf);
    const int i0 = 1 + (int)blockIdx.x * kT, j0 = 1 + (int)blockIdx.y * kT;
    const int e = threadIdx.x;
    const int ilim = dm.nx - i0 + 1 < kT ? dm.nx - i0 + 1 : kT, jlim = dm.ny - j0 + 1 < kT ? dm.ny - j0 + 1 : kT;
    const int row = jlim * K;
    for (int v = e; v < ilim * row; v += kT * kT) {
        const int ti = v / row, w = v - ti * row;
        tile[(size_t)ti * kT * K + w] = cell[((size_t)(i0 + ti) * (dm.ny + 2) + j0) * K + w];
    }
    __syncthreads();
    const size_t ncell = (size_t)dm.cells();
    int a, b;
    tile_cell(e, a, b);
    for (int Q = 0; Q < 4; ++Q) {
        const bool xn = quad_xneg(Q), yn = quad_yneg(Q);
        int c0, r0;
        dm.local(xn ? i0 + kT - 1 : i0, yn ? j0 + kT - 1 : j0, xn, yn, c0, r0);
        const int c = c0 + a, r = r0 + b;
        int i, j;
        dm.mesh(c, r, xn, yn, i, j);
        if (c >= 0 && r >= 0 && c < dm.nx && r < dm.ny && i >= i0 && i < i0 + kT && j >= j0 && j < j0 + kT) {
            TO *dst = out + (size_t)Q * K * ncell + (size_t)dm.pos(c + r, r);
            const TI *src = tile + ((size_t)(i - i0) * kT + (j - j0)) * K;
            for (int k = 0; k < K; ++k) dst[(size_t)k * ncell] = TO(src[k]);
        }
    }
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
    const size_t lds = (size_t)kT * kT * K * sizeof(TI);
    if (lds <= 32 * 1024) {
        const dim3 grid((unsigned)((g.nx + kT - 1) / kT), (unsigned)((g.ny + kT - 1) / kT));
        hipLaunchKernelGGL((k_cell_to_diag_tiled<TI, TO>), grid, dim3(kT * kT), lds, s, cell, out, dm, K, skip);
    } else {
        hipLaunchKernelGGL((k_cell_to_diag<TI, TO>), dim3(g.nx + g.ny - 1, 4), dim3(rows_block(g.ny)), 0, s, cell, out,
                           dm, K, skip);
    }
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
    const size_t lds = (size_t)kT * kT * g.N() * sizeof(double);
    if (lds <= 32 * 1024) {
        const dim3 grid((unsigned)((g.nx + kT - 1) / kT), (unsigned)((g.ny + kT - 1) / kT));
        hipLaunchKernelGGL(k_phi_from_quadrants_tiled, grid, dim3(kT * kT), lds, s, phiQ, phi, xb, dm, g.N(), alpha,
                           beta, skip);
    } else {
        hipLaunchKernelGGL(k_phi_from_quadrants, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, phiQ, phi, xb, dm,
                           g.N(), alpha, beta, skip);
    }
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
    const size_t pQ = (size_t)Q * ncell + p;
    double ax[N];
    apply_element<double, P>(x, xi, eta, dx, dy, (sigt_d[pQ] + inv_dt) * dx * dy, ax);
    double zero[N];
#pragma unroll
    for (int k = 0; k < N; ++k) zero[k] = 0.0;
    double inflow[N];
    rhs_from_traces<double, P>(inflow, zero, tx, ty, xi, eta, dx, dy);
    const double inv_area = 1.0 / (dx * dy);
    TO *pr = res + (size_t)q * N * ncell + p;
#pragma unroll
    for (int row = 0; row < N; ++row) {
        const double lx = (ax[row] - inflow[row]) * inv_area;
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
