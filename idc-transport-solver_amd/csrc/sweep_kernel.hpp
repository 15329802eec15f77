// Staggered sweep kernel on the diagonal-major layout (diag_layout.hpp), included by the translation
// units that instantiate it (one per precision and DG degree, so they compile in parallel).
//
// Sweep (k_diag_sweep): one workgroup per ordinate, one compute thread per mesh row r plus a storer
// wave.  At step s thread r solves cell (c = s - r, r), so a step is one diagonal and the workgroup
// barrier ending it is the diagonal barrier of the reference's wavefront
// (SweepPatternUniCartWavefront.cpp:253-296).  The x-upwind face trace stays in registers, the
// y-upwind trace passes through LDS.  Compute waves load the inputs of step s+D branch-free (lanes
// outside the mesh read a clamped, valid position) and never store to global memory; the storer wave
// writes step s-1 during step s.  In the diagonal layout all these accesses are contiguous across
// the lanes of a wavefront.
//
// The source of a cell is Q (optional, per ordinate) plus sigma_s * phi / (4 pi) (optional, the
// isotropic scattering source of Smv, TransportOperator.cpp:369-392, fused into the load).
#pragma once

#include <algorithm>
#include <string>

#include "cell_solve.hpp"
#include "diag_layout.hpp"
#include "tsw_internal.hpp"

namespace tsw {

constexpr double kInv4Pi = 0.079577471545947667884;

template <typename T, int P>
struct DiagLds {  // LDS carve-up in units of T
    static constexpr int N = P * P;
    size_t ex, in, out, inv, total;
    __host__ __device__ DiagLds(int cthreads, int nx, int ny, int M) {
        ex = 0;                              // [2][cthreads][P] y-traces between rows
        in = ex + (size_t)2 * cthreads * P;  // [nx][P] y-inflow traces of row 0
        out = in + (size_t)nx * P;           // [2][ny][N] solved cells awaiting the store
        inv = out + (size_t)2 * ny * N;      // [M][N][N] element inverses (material path)
        total = inv + (size_t)M * N * N;
    }
};

// Band exchange: records of P values moved with relaxed device-scope atomics (coherent across the
// chip's L2 slices); an empty record holds ExchangeSentinel<T>::bits.
constexpr int kSpinLimit = 1 << 20;

template <typename T>
__device__ __forceinline__ typename ExchangeSentinel<T>::U to_bits(T v) {
    return __builtin_bit_cast(typename ExchangeSentinel<T>::U, v);
}
template <typename T>
__device__ __forceinline__ T from_bits(typename ExchangeSentinel<T>::U v) {
    return __builtin_bit_cast(T, v);
}
// Global-address-space view of a pointer: accesses compile to global_* instructions, which count
// only against vmcnt, so LDS waits do not also wait for outstanding prefetches.
template <typename T>
using gptr = __attribute__((address_space(1))) T *;
template <typename T>
__device__ __forceinline__ gptr<T> G(T *p) {
    return (gptr<T>)p;
}

template <typename U>
__device__ __forceinline__ U xload(gptr<const U> p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename U>
__device__ __forceinline__ void xstore(gptr<U> p, U v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename U, int P>
struct Record {
    U v[P];
};

// Waits until record `rec` (already loaded once into r) is filled; on the spin limit, or when another
// band has given up (*abort_flag), returns zeros and the caller proceeds.
template <typename T, int P>
__device__ __noinline__ Record<typename ExchangeSentinel<T>::U, P> wait_record(
    gptr<const typename ExchangeSentinel<T>::U> rec, Record<typename ExchangeSentinel<T>::U, P> r, gptr<int> abort_flag) {
    using U = typename ExchangeSentinel<T>::U;
    for (int it = 0;; ++it) {
        bool ok = true;
#pragma unroll
        for (int d = 0; d < P; ++d) ok = ok && r.v[d] != ExchangeSentinel<T>::bits;
        if (ok) return r;
        if (it >= kSpinLimit || ((it & 255) == 255 && xload<int>(abort_flag) != 0)) {
            if (it >= kSpinLimit) xstore<int>(abort_flag, 1);
#pragma unroll
            for (int d = 0; d < P; ++d) r.v[d] = 0;
            return r;
        }
        __builtin_amdgcn_s_sleep(2);
#pragma unroll
        for (int d = 0; d < P; ++d) r.v[d] = xload<U>(rec + d);
    }
}

// MODE bit 1: add Q from `qsrc`; bit 2: add sigma_s * phi / (4 pi).
// Workgroup b*nq + q sweeps rows [b*band_rows, ...) of ordinate q (band-major numbering, so a band's
// upstream neighbour is dispatched before it and never waits on it).
template <typename T, int P, bool MAT, int MODE, int D>
__global__ __launch_bounds__(512) void k_diag_sweep(DiagSweepParams<T> sp1, const DiagSweepParams<T> *__restrict__ batch,
                                                    const int2 *__restrict__ items) {
    constexpr int N = P * P;
    using U = typename ExchangeSentinel<T>::U;
    // one problem (sp1, workgroup = band x ordinate) or a batch (items[workgroup] = {problem, ordinate})
    int item_p = 0, item_q = 0;
    if (batch) {
        const gptr<const int> it = G(reinterpret_cast<const int *>(items)) + 2 * (size_t)blockIdx.x;
        item_p = it[0];
        item_q = it[1];
    }
    const DiagSweepParams<T> sp = batch ? batch[item_p] : sp1;
    if (sp.skip && *G(sp.skip)) return;  // a replayed iteration after the loop has stopped
    const int nb = (batch || sp.nbands < 2) ? 1 : sp.nbands;
    const int q = batch ? item_q : (int)(blockIdx.x % (unsigned)sp.nq);
    const int band = batch ? 0 : (int)(blockIdx.x / (unsigned)sp.nq);
    constexpr bool HASQ = (MODE & 1) != 0, HASPHI = (MODE & 2) != 0;
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const int nx = sp.nx, ny = sp.ny;
    const int BR = nb > 1 ? sp.band_rows : ny;  // rows per band
    const int R0 = band * BR;                    // first row of this band
    const int nrows = BR < ny - R0 ? BR : ny - R0;
    const int cthreads = ((BR + 63) / 64) * 64;  // waves beyond this row count idle at the barriers
    const DiagLds<T, P> L(cthreads, nx, BR, MAT ? sp.M : 0);
    T *ex = lds + L.ex, *lin = lds + L.in, *lout = lds + L.out, *linv = lds + L.inv;
    const DiagMap dm{nx, ny};
    const size_t ncell = (size_t)dm.cells();
    // exchange records: read from the band below (xin), written for the band above (xout)
    const gptr<U> xin = (nb > 1 && band > 0) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band - 1) * nx * P : nullptr;
    const gptr<U> xout = (nb > 1 && band < nb - 1) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band) * nx * P : nullptr;

    const int Q = q >= sp.qo3 ? 3 : (q >= sp.qo2 ? 2 : (q >= sp.qo1 ? 1 : 0));
    const int tid = threadIdx.x;
    const T xi = G(sp.xi)[q], eta = G(sp.eta)[q];
    const bool xneg = xi <= T(0), yneg = eta <= T(0);

    if (MAT) {
        for (int k = tid; k < sp.M * N * N; k += blockDim.x) {
            const int m = k / (N * N), e = k - m * N * N;
            linv[k] = G(sp.ainv)[((size_t)m * sp.nq + q) * N * N + e];
        }
    }
    if (band == 0)
        for (int k = tid; k < nx * P; k += blockDim.x) lin[k] = sp.tin_y ? G(sp.tin_y)[(size_t)q * nx * P + k] : T(0);

    const int steps = nx + nrows - 1;              // local step t is global diagonal R0 + t
    const int nsteps = ((steps + D - 1) / D) * D;  // padded steps have no active cell
    const gptr<T> psi_q = G(sp.psi) + (size_t)q * N * ncell;
    // rows of this band on diagonal s
    auto band_lo = [&](int s) { const int a = dm.rlo(s); return a > R0 ? a : R0; };
    auto band_hi = [&](int s) {
        const int a = s < ny - 1 ? s : ny - 1, b = R0 + nrows - 1;
        return a < b ? a : b;
    };

    if (tid >= (int)blockDim.x - 64) {
        // storer: barrier k (k >= 1) closes compute step k-1, whose cells are written here while
        // the compute waves fill the other ring slot; it also hands the band's top trace on and empties
        // the records this band has consumed
        const int lane = tid - ((int)blockDim.x - 64);
        lds_barrier();
        for (int t = 0; t < nsteps; ++t) {
            lds_barrier();
            if (t >= steps) continue;
            const int s = R0 + t;
            if (xout) {
                const int c = t - (nrows - 1);
                if (c >= 0 && c < nx && lane < P)
                    xstore(xout + (size_t)c * P + lane, to_bits(ex[((size_t)(t & 1) * cthreads + (nrows - 1)) * P + lane]));
            }
            if (xin && t < nx && lane < P) xstore(xin + (size_t)t * P + lane, ExchangeSentinel<T>::bits);
            const T *slot = lout + (size_t)(t & 1) * BR * N;
            const int lo = band_lo(s), hi = band_hi(s);
            const long long p0 = dm.base(s) - dm.rlo(s);
            for (int r = lo + lane; r <= hi; r += 64) {
                const T *src = slot + (size_t)(r - R0) * N;
#pragma unroll
                for (int k = 0; k < N; ++k) psi_q[(size_t)k * ncell + (size_t)(p0 + r)] = src[k];
            }
        }
        return;
    }

    const int rl = tid < cthreads ? tid : cthreads - 1;  // spare waves mirror the last row, stay inactive
    const int r = R0 + rl;                               // mesh row (counted from the inflow edge)
    const bool row_ok = tid < nrows;
    T tx[P];
#pragma unroll
    for (int d = 0; d < P; ++d) tx[d] = (row_ok && sp.tin_x) ? G(sp.tin_x)[((size_t)q * ny + r) * P + d] : T(0);

    // position of this row's cell at local step t; lanes whose row has no cell on that diagonal read
    // the nearest cell of the band on it, so a wavefront's loads stay within one contiguous run
    // (diagonal bounds are wave-uniform scalar arithmetic)
    auto pos_at = [&](int t) -> size_t {
        if (t >= steps) return 0;
        const int s = R0 + t;
        const int lo = band_lo(s), hi = band_hi(s);
        const int rc = r < lo ? lo : (r > hi ? hi : r);
        return (size_t)(dm.base(s) + rc - dm.rlo(s));
    };
    const gptr<const T> qsrc_q = HASQ ? G(sp.qsrc) + (size_t)q * N * ncell : nullptr;
    const gptr<const T> phi_Q = HASPHI ? G(sp.phi) + (size_t)Q * N * ncell : nullptr;
    const gptr<const T> sigs_Q = HASPHI ? G(sp.sig_s) + (size_t)Q * ncell : nullptr;
    const gptr<const int> mat_Q = MAT ? G(sp.mat) + (size_t)Q * ncell : nullptr;
    const gptr<const T> sigt_Q = MAT ? nullptr : G(sp.sig_t) + (size_t)Q * ncell;
    const bool xrow = xin != nullptr && rl == 0;  // this lane takes its y-inflow from the band below

    // prefetch ring: slot u holds the inputs this row meets at a step t with t % D == u
    T pq[D][HASQ ? N : 1], pf[D][HASPHI ? N : 1], ps[D], pm_sig[D];
    int pm[D];
    Record<U, P> px[D];
    auto fetch = [&](int u, int t) {
        const size_t p = pos_at(t);
        if (HASQ) {
#pragma unroll
            for (int k = 0; k < N; ++k) pq[u][k] = qsrc_q[(size_t)k * ncell + p];
        }
        if (HASPHI) {
#pragma unroll
            for (int k = 0; k < N; ++k) pf[u][k] = phi_Q[(size_t)k * ncell + p];
            ps[u] = sigs_Q[p];
        }
        if (MAT) pm[u] = mat_Q[p];
        else pm_sig[u] = sigt_Q[p];
        if (xrow && t < nx) {
#pragma unroll
            for (int d = 0; d < P; ++d) px[u].v[d] = xload<U>(xin + (size_t)t * P + d);
        }
    };
#pragma unroll
    for (int u = 0; u < D; ++u) fetch(u, u);
    lds_barrier();

    // FP32 keeps the current material's inverse in registers and reloads it from LDS only when the
    // row enters another material (the LDS reads then move data only for the lanes that changed)
    constexpr bool REGMAT = MAT && sizeof(T) == 4;
    T Ainv[REGMAT ? N * N : 1];
    int cur_mat = -1;
    const T area = sp.dx * sp.dy;
    for (int t0 = 0; t0 < nsteps; t0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int t = t0 + u;
            const int c = t - rl;
            const bool active = row_ok && c >= 0 && c < nx;
            // consume slot u, then refill it with step t + D (kept apart by scheduling barriers so
            // the slot keeps its registers and its wait covers only its own loads)
            T B[N];
            {
                T src[N];
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    T v = T(0);
                    if (HASQ) v = pq[u][k];
                    if (HASPHI) v += ps[u] * pf[u][k] * T(kInv4Pi);
                    src[k] = v;
                }
                T ty[P];
                if (xrow) {
                    if (active) {
                        bool ok = true;
#pragma unroll
                        for (int d = 0; d < P; ++d) ok = ok && px[u].v[d] != ExchangeSentinel<T>::bits;
                        if (!ok) px[u] = wait_record<T, P>(xin + (size_t)c * P, px[u], G(sp.abort_flag));
                    }
#pragma unroll
                    for (int d = 0; d < P; ++d) ty[d] = active ? from_bits<T>(px[u].v[d]) : T(0);
                } else {
                    const int cc = c < 0 ? 0 : (c >= nx ? nx - 1 : c);
                    const T *tsrc = rl == 0 ? lin + (size_t)cc * P : ex + ((size_t)((t - 1) & 1) * cthreads + (rl - 1)) * P;
#pragma unroll
                    for (int d = 0; d < P; ++d) ty[d] = tsrc[d];
                }
                rhs_from_traces<T, P>(B, src, tx, ty, xi, eta, sp.dx, sp.dy);
            }
            const int mat = pm[u];
            const T sig = pm_sig[u];
            __builtin_amdgcn_sched_barrier(0);
            fetch(u, t + D);
            __builtin_amdgcn_sched_barrier(0);
            if (active) {
                T x[N];
                if (REGMAT) {
                    if (mat != cur_mat) {
                        cur_mat = mat;
                        const T *Mx = linv + (size_t)mat * N * N;
#pragma unroll
                        for (int k = 0; k < N * N; ++k) Ainv[k] = Mx[k];
                    }
#pragma unroll
                    for (int row = 0; row < N; ++row) {
                        T acc = T(0);
#pragma unroll
                        for (int col = 0; col < N; ++col) acc += Ainv[row * N + col] * B[col];
                        x[row] = acc;
                    }
                } else if (MAT) {
                    matvec_lds<T, N>(linv + (size_t)mat * N * N, B, x);
                } else {
                    T A[N * N];
                    assemble_matrix<T, P>(A, xi, eta, sp.dx, sp.dy, (sig + sp.inv_dt) * area);
                    ge_solve_rcp<T, N>(A, B);
#pragma unroll
                    for (int k = 0; k < N; ++k) x[k] = B[k];
                }
                T *o = lout + ((size_t)(t & 1) * BR + rl) * N;
#pragma unroll
                for (int k = 0; k < N; ++k) o[k] = x[k];
                trace_x<T, P>(x, xneg, tx);
                T tyo[P];
                trace_y<T, P>(x, yneg, tyo);
                T *dst_l = ex + ((size_t)(t & 1) * cthreads + rl) * P;
#pragma unroll
                for (int d = 0; d < P; ++d) dst_l[d] = tyo[d];
            }
            lds_barrier();
        }
    }
}

// rows handled by one workgroup
template <typename T>
inline int band_rows_of(const DiagSweepParams<T> &sp) {
    return sp.nbands > 1 ? sp.band_rows : sp.ny;
}

template <typename T, int P, bool MAT, int MODE, int D>
void launch_ds_d(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s, const DiagSweepParams<T> *batch,
                        const int2 *items, int nitems, int max_ny) {
    const int rows = batch ? max_ny : band_rows_of(sp);
    const int threads = ((rows + 63) / 64) * 64 + 64;
    auto kern = k_diag_sweep<T, P, MAT, MODE, D>;
    if (lds > 64 * 1024)
        TSW_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int blocks = batch ? nitems : sp.nq * (sp.nbands > 1 ? sp.nbands : 1);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, s, sp, batch, items);
    TSW_HIP(hipGetLastError());
}

// Prefetch depth D (steps of inputs in flight per row): sp.prefetch, or 2 (FP64) / 3 (FP32) by default.
template <typename T, int P, bool MAT, int MODE>
void launch_ds(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s, const DiagSweepParams<T> *batch = nullptr,
                      const int2 *items = nullptr, int nitems = 0, int max_ny = 0) {
    const int D = sp.prefetch > 0 ? sp.prefetch : (sizeof(T) == 8 ? 2 : 3);
    switch (D) {
        case 2: launch_ds_d<T, P, MAT, MODE, 2>(sp, lds, s, batch, items, nitems, max_ny); break;
        case 3: launch_ds_d<T, P, MAT, MODE, 3>(sp, lds, s, batch, items, nitems, max_ny); break;
        case 4: launch_ds_d<T, P, MAT, MODE, 4>(sp, lds, s, batch, items, nitems, max_ny); break;
        case 6: launch_ds_d<T, P, MAT, MODE, 6>(sp, lds, s, batch, items, nitems, max_ny); break;
        default: throw Error(TSW_E_ARG, "prefetch depth must be 2, 3, 4 or 6");
    }
}

template <typename T, int P, bool MAT>
void launch_ds_mode(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s) {
    const int mode = (sp.qsrc ? 1 : 0) | (sp.phi ? 2 : 0);
    switch (mode) {
        case 1: launch_ds<T, P, MAT, 1>(sp, lds, s); break;
        case 2: launch_ds<T, P, MAT, 2>(sp, lds, s); break;
        case 3: launch_ds<T, P, MAT, 3>(sp, lds, s); break;
        default: throw Error(TSW_E_ARG, "diagonal sweep without a source");
    }
}

template <typename T, int P>
void launch_ds_p(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s) {
    if (sp.ainv) launch_ds_mode<T, P, true>(sp, lds, s);
    else launch_ds_mode<T, P, false>(sp, lds, s);
}

// Entry points per (precision, DG degree P = p+1), instantiated by TSW_SWEEP_INSTANTIATE.
template <typename T, int P>
void launch_diag_sweep_p(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s) {
    launch_ds_p<T, P>(sp, lds, s);
}

template <typename T, int P>
void launch_diag_sweep_batch_p(const DiagSweepParams<T> &rep, const DiagSweepParams<T> *probs, const int2 *items,
                               int nitems, int max_ny, size_t lds, hipStream_t s) {
    const int mode = (rep.qsrc ? 1 : 0) | (rep.phi ? 2 : 0);
    if (rep.ainv) {
        if (mode == 1) launch_ds<T, P, true, 1>(rep, lds, s, probs, items, nitems, max_ny);
        else if (mode == 2) launch_ds<T, P, true, 2>(rep, lds, s, probs, items, nitems, max_ny);
        else launch_ds<T, P, true, 3>(rep, lds, s, probs, items, nitems, max_ny);
    } else {
        if (mode == 1) launch_ds<T, P, false, 1>(rep, lds, s, probs, items, nitems, max_ny);
        else if (mode == 2) launch_ds<T, P, false, 2>(rep, lds, s, probs, items, nitems, max_ny);
        else launch_ds<T, P, false, 3>(rep, lds, s, probs, items, nitems, max_ny);
    }
}

#define TSW_SWEEP_INSTANTIATE(T, P)                                                                         \
    template void launch_diag_sweep_p<T, P>(const DiagSweepParams<T> &, size_t, hipStream_t);              \
    template void launch_diag_sweep_batch_p<T, P>(const DiagSweepParams<T> &, const DiagSweepParams<T> *,   \
                                                  const int2 *, int, int, size_t, hipStream_t);

#define TSW_SWEEP_EXTERN(T, P)                                                                              \
    extern template void launch_diag_sweep_p<T, P>(const DiagSweepParams<T> &, size_t, hipStream_t);       \
    extern template void launch_diag_sweep_batch_p<T, P>(const DiagSweepParams<T> &, const DiagSweepParams<T> *, \
                                                         const int2 *, int, int, size_t, hipStream_t);

}  // namespace tsw
