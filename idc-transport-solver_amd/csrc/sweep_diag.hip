// Staggered sweep and its companion operators on the diagonal-major layout (diag_layout.hpp).
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
#include <algorithm>

#include "cell_solve.hpp"
#include "diag_layout.hpp"
#include "tsw_internal.hpp"

namespace tsw {

namespace {
constexpr double kInv4Pi = 0.079577471545947667884;
}

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
static int band_rows_of(const DiagSweepParams<T> &sp) {
    return sp.nbands > 1 ? sp.band_rows : sp.ny;
}

template <typename T>
size_t diag_sweep_lds_bytes(const DiagSweepParams<T> &sp, int p) {
    const int BR = band_rows_of(sp);
    if (BR > 448) return 0;
    const int cthreads = ((BR + 63) / 64) * 64;
    const int M = sp.ainv ? sp.M : 0;
    size_t total;
    switch (p) {
        case 0: total = DiagLds<T, 1>(cthreads, sp.nx, BR, M).total; break;
        case 1: total = DiagLds<T, 2>(cthreads, sp.nx, BR, M).total; break;
        case 2: total = DiagLds<T, 3>(cthreads, sp.nx, BR, M).total; break;
        case 3: total = DiagLds<T, 4>(cthreads, sp.nx, BR, M).total; break;
        default: return 0;
    }
    const size_t bytes = total * sizeof(T);
    return bytes <= max_lds_per_block() ? bytes : 0;
}

template <typename T, int P, bool MAT, int MODE>
static void launch_ds(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s, const DiagSweepParams<T> *batch = nullptr,
                      const int2 *items = nullptr, int nitems = 0, int max_ny = 0) {
    constexpr int D = sizeof(T) == 8 ? 2 : 3;
    const int rows = batch ? max_ny : band_rows_of(sp);
    const int threads = ((rows + 63) / 64) * 64 + 64;
    auto kern = k_diag_sweep<T, P, MAT, MODE, D>;
    if (lds > 64 * 1024)
        TSW_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    const int blocks = batch ? nitems : sp.nq * (sp.nbands > 1 ? sp.nbands : 1);
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), lds, s, sp, batch, items);
    TSW_HIP(hipGetLastError());
}

template <typename T, int P, bool MAT>
static void launch_ds_mode(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s) {
    const int mode = (sp.qsrc ? 1 : 0) | (sp.phi ? 2 : 0);
    switch (mode) {
        case 1: launch_ds<T, P, MAT, 1>(sp, lds, s); break;
        case 2: launch_ds<T, P, MAT, 2>(sp, lds, s); break;
        case 3: launch_ds<T, P, MAT, 3>(sp, lds, s); break;
        default: throw Error(TSW_E_ARG, "diagonal sweep without a source");
    }
}

template <typename T, int P>
static void launch_ds_p(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s) {
    if (sp.ainv) launch_ds_mode<T, P, true>(sp, lds, s);
    else launch_ds_mode<T, P, false>(sp, lds, s);
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
        case 0: launch_ds_p<T, 1>(sp, lds, s); break;
        case 1: launch_ds_p<T, 2>(sp, lds, s); break;
        case 2: launch_ds_p<T, 3>(sp, lds, s); break;
        case 3: launch_ds_p<T, 4>(sp, lds, s); break;
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
    const int mode = (rep.qsrc ? 1 : 0) | (rep.phi ? 2 : 0);
    const bool mat = rep.ainv != nullptr;
#define TSW_B(PP)                                                                                           \
    do {                                                                                                    \
        if (mat) {                                                                                          \
            if (mode == 1) launch_ds<T, PP, true, 1>(rep, lds, s, probs, items, nitems, max_ny);            \
            else if (mode == 2) launch_ds<T, PP, true, 2>(rep, lds, s, probs, items, nitems, max_ny);       \
            else launch_ds<T, PP, true, 3>(rep, lds, s, probs, items, nitems, max_ny);                      \
        } else {                                                                                            \
            if (mode == 1) launch_ds<T, PP, false, 1>(rep, lds, s, probs, items, nitems, max_ny);           \
            else if (mode == 2) launch_ds<T, PP, false, 2>(rep, lds, s, probs, items, nitems, max_ny);      \
            else launch_ds<T, PP, false, 3>(rep, lds, s, probs, items, nitems, max_ny);                     \
        }                                                                                                   \
    } while (0)
    switch (p) {
        case 0: TSW_B(1); break;
        case 1: TSW_B(2); break;
        case 2: TSW_B(3); break;
        case 3: TSW_B(4); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree not built");
    }
#undef TSW_B
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
// them through LDS: on the reference side each cell's ordinates of the quadrant are one contiguous
// run, on the diagonal side each (q, k) plane receives R consecutive positions.
template <typename TR, typename TD>
__global__ __launch_bounds__(256) void k_psi_convert(TR *ref, TD *diag, DiagMap dm, int nq, int N, int qo1, int qo2,
                                                     int qo3, int R, int dir) {
    extern __shared__ unsigned char smem_raw[];
    double *tile = reinterpret_cast<double *>(smem_raw);  // [R][nqQ*N]
    const int s = blockIdx.x, Q = blockIdx.z;
    const int r0 = dm.rlo(s) + (int)blockIdx.y * R;
    const int rhi = s < dm.ny - 1 ? s : dm.ny - 1;
    if (r0 > rhi) return;
    const int nr = (rhi - r0 + 1) < R ? (rhi - r0 + 1) : R;
    const int q0 = Q == 0 ? 0 : (Q == 1 ? qo1 : (Q == 2 ? qo2 : qo3));
    const int q1 = Q == 0 ? qo1 : (Q == 1 ? qo2 : (Q == 2 ? qo3 : nq));
    const int run = (q1 - q0) * N;  // values per cell on the reference side
    const size_t ncell = (size_t)dm.cells();
    const long long p0 = dm.pos(s, r0);
    const bool xn = quad_xneg(Q), yn = quad_yneg(Q);
    auto ref_run = [&](int t) -> TR * {  // start of cell (s, r0 + t)'s run for this quadrant
        int i, j;
        dm.mesh(s - (r0 + t), r0 + t, xn, yn, i, j);
        return ref + (((size_t)i * (dm.ny + 2) + j) * nq + q0) * N;
    };
    if (dir == 0) {
        for (int t = 0; t < nr; ++t) {
            const TR *src = ref_run(t);
            for (int e = threadIdx.x; e < run; e += blockDim.x) tile[(size_t)t * run + e] = (double)src[e];
        }
        __syncthreads();
        for (int e = threadIdx.x; e < run * nr; e += blockDim.x) {
            const int t = e % nr, qk = e / nr;  // consecutive threads: consecutive positions
            const int q = q0 + qk / N, k = qk % N;
            diag[(size_t)q * N * ncell + (size_t)k * ncell + (size_t)(p0 + t)] = TD(tile[(size_t)t * run + qk]);
        }
    } else {
        for (int e = threadIdx.x; e < run * nr; e += blockDim.x) {
            const int t = e % nr, qk = e / nr;
            const int q = q0 + qk / N, k = qk % N;
            tile[(size_t)t * run + qk] = (double)diag[(size_t)q * N * ncell + (size_t)k * ncell + (size_t)(p0 + t)];
        }
        __syncthreads();
        for (int t = 0; t < nr; ++t) {
            TR *dst = ref_run(t);
            for (int e = threadIdx.x; e < run; e += blockDim.x) dst[e] = TR(tile[(size_t)t * run + e]);
        }
    }
}

template <typename TR, typename TD>
void launch_psi_convert(const Geom &g, TR *ref, TD *diag, int dir, hipStream_t s) {
    const DiagMap dm{g.nx, g.ny};
    int max_run = 0;
    for (int Q = 0; Q < 4; ++Q) max_run = std::max(max_run, (g.quad_off[Q + 1] - g.quad_off[Q]) * g.N());
    if (max_run == 0) return;
    int R = 32;
    while (R > 1 && (size_t)R * max_run * sizeof(double) > 32 * 1024) R /= 2;
    const int maxlen = g.nx < g.ny ? g.nx : g.ny;
    const dim3 grid(g.nx + g.ny - 1, (maxlen + R - 1) / R, 4);
    hipLaunchKernelGGL((k_psi_convert<TR, TD>), grid, dim3(256), (size_t)R * max_run * sizeof(double), s, ref, diag,
                       dm, g.nq, g.N(), g.quad_off[1], g.quad_off[2], g.quad_off[3], R, dir);
    TSW_HIP(hipGetLastError());
}
template void launch_psi_convert<double, double>(const Geom &, double *, double *, int, hipStream_t);
template void launch_psi_convert<double, float>(const Geom &, double *, float *, int, hipStream_t);

// per-cell array [i][j][k] (k < K, halo layout) -> per-quadrant diagonal [Q][k][pos]
template <typename TI, typename TO>
__global__ void k_cell_to_diag(const TI *cell, TO *out, DiagMap dm, int K) {
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
void launch_cell_to_diag(const Geom &g, const TI *cell, TO *out, int K, hipStream_t s) {
    const DiagMap dm{g.nx, g.ny};
    hipLaunchKernelGGL((k_cell_to_diag<TI, TO>), dim3(g.nx + g.ny - 1, 4), dim3(rows_block(g.ny)), 0, s, cell, out, dm,
                       K);
    TSW_HIP(hipGetLastError());
}
template void launch_cell_to_diag<double, double>(const Geom &, const double *, double *, int, hipStream_t);
template void launch_cell_to_diag<double, float>(const Geom &, const double *, float *, int, hipStream_t);
template void launch_cell_to_diag<int32_t, int32_t>(const Geom &, const int32_t *, int32_t *, int, hipStream_t);

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
                                                      int qo2, int qo3, int nq) {
    const int Q = blockIdx.y;
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ncell * N) return;
    const int q0 = Q == 0 ? 0 : (Q == 1 ? qo1 : (Q == 2 ? qo2 : qo3));
    const int q1 = Q == 0 ? qo1 : (Q == 1 ? qo2 : (Q == 2 ? qo3 : nq));
    double acc = 0.0;
    for (int q = q0; q < q1; ++q) acc += w[q] * (double)psi[(size_t)q * N * ncell + t];
    phiQ[(size_t)Q * N * ncell + t] = acc;
}

__global__ void k_phi_from_quadrants(const double *__restrict__ phiQ, double *__restrict__ phi, DiagMap dm, int N,
                                     double alpha, double beta) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t ncell = (size_t)dm.cells();
    if (t >= ncell * N) return;
    const int k = (int)(t % N);
    const size_t cidx = t / N;
    const int i = 1 + (int)(cidx / dm.ny), j = 1 + (int)(cidx % dm.ny);
    double acc = 0.0;
    for (int Q = 0; Q < 4; ++Q)
        acc += phiQ[(size_t)Q * N * ncell + (size_t)k * ncell + (size_t)dm.pos_of_cell(i, j, quad_xneg(Q), quad_yneg(Q))];
    double *y = phi + ((size_t)i * (dm.ny + 2) + j) * N + k;
    *y = (beta == 0.0 ? 0.0 : beta * *y) + alpha * acc;
}

template <typename TI>
void launch_pmv_diag(const Geom &g, const TI *psi, const double *w, double *phiQ, double *phi, double alpha, double beta,
                     hipStream_t s) {
    const DiagMap dm{g.nx, g.ny};
    const size_t n = (size_t)dm.cells() * g.N();
    hipLaunchKernelGGL((k_pmv_quadrant<TI>), dim3((unsigned)((n + 255) / 256), 4), dim3(256), 0, s, psi, w, phiQ,
                       (size_t)dm.cells(), g.N(), g.quad_off[1], g.quad_off[2], g.quad_off[3], g.nq);
    hipLaunchKernelGGL(k_phi_from_quadrants, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, phiQ, phi, dm, g.N(),
                       alpha, beta);
    TSW_HIP(hipGetLastError());
}
template void launch_pmv_diag<double>(const Geom &, const double *, const double *, double *, double *, double, double, hipStream_t);
template void launch_pmv_diag<float>(const Geom &, const float *, const double *, double *, double *, double, double, hipStream_t);

// =====================================================================================================
// FP64 angular residual on the diagonal layout, the refinement's r = Q + sigma_s phi/(4 pi) - L psi
// (Lmv, TransportOperator.cpp:1576-1658, written with face traces), stored in TO.
// psi's inflow comes from tin_x/tin_y (traces of the halo cells).  Grid (diagonal, ordinate).
// =====================================================================================================
template <typename TO, int P>
__global__ void k_residual_diag(const double *__restrict__ qsrc, const double *__restrict__ phi_d,
                                const double *__restrict__ sigs_d, const double *__restrict__ sigt_d,
                                const double *__restrict__ psi, const double *__restrict__ tin_x,
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
    const double *pq = psi + (size_t)q * N * ncell;
    double x[N];
#pragma unroll
    for (int k = 0; k < N; ++k) x[k] = pq[(size_t)k * ncell + p];
    double tx[P], ty[P];
    if (c == 0) {
#pragma unroll
        for (int d = 0; d < P; ++d) tx[d] = tin_x[((size_t)q * dm.ny + r) * P + d];
    } else {
        const size_t pu = (size_t)dm.pos(s - 1, r);
        double u[N];
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = pq[(size_t)k * ncell + pu];
        trace_x<double, P>(u, xneg, tx);
    }
    if (r == 0) {
#pragma unroll
        for (int d = 0; d < P; ++d) ty[d] = tin_y[((size_t)q * dm.nx + c) * P + d];
    } else {
        const size_t pu = (size_t)dm.pos(s - 1, r - 1);
        double u[N];
#pragma unroll
        for (int k = 0; k < N; ++k) u[k] = pq[(size_t)k * ncell + pu];
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
        v += sigs_d[pQ] * phi_d[(size_t)Q * N * ncell + (size_t)row * ncell + p] * kInv4Pi;
        pr[(size_t)row * ncell] = TO(v - lx);
    }
}

template <typename TO>
void launch_residual_diag(const Geom &g, const double *qsrc, const double *phi_d, const double *sigs_d,
                          const double *sigt_d, const double *psi, const double *tin_x, const double *tin_y,
                          const double *xi, const double *eta, double inv_dt, TO *res, hipStream_t s) {
    const DiagMap dm{g.nx, g.ny};
    const dim3 grid(g.nx + g.ny - 1, g.nq), block(rows_block(g.ny));
    const int a = g.quad_off[1], b = g.quad_off[2], c = g.quad_off[3];
#define TSW_RES(PP)                                                                                                   \
    hipLaunchKernelGGL((k_residual_diag<TO, PP>), grid, block, 0, s, qsrc, phi_d, sigs_d, sigt_d, psi, tin_x, tin_y, \
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
template void launch_residual_diag<float>(const Geom &, const double *, const double *, const double *, const double *,
                                          const double *, const double *, const double *, const double *, const double *,
                                          double, float *, hipStream_t);
template void launch_residual_diag<double>(const Geom &, const double *, const double *, const double *, const double *,
                                           const double *, const double *, const double *, const double *, const double *,
                                           double, double *, hipStream_t);

}  // namespace tsw
