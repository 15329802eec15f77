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

// 16-byte LDS vectors.  Every LDS row below starts on a 16-byte boundary and is padded to whole vectors,
// so a row moves with ds_read_b128 / ds_write_b128 instead of one access per value.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
    typedef float V __attribute__((ext_vector_type(4)));
    static constexpr int W = 4;
};
template <>
struct Vec16<double> {
    typedef double V __attribute__((ext_vector_type(2)));
    static constexpr int W = 2;
};
template <typename T>
__host__ __device__ constexpr int vround(int n) {
    return (n + Vec16<T>::W - 1) / Vec16<T>::W * Vec16<T>::W;
}

// v[0..LEN) <- p[0..LEN), p 16-byte aligned with vround(LEN) readable entries
template <typename T, int LEN>
__device__ __forceinline__ void lds_get(const T *p, T (&v)[LEN]) {
    using V = typename Vec16<T>::V;
    constexpr int W = Vec16<T>::W;
    const V *pv = reinterpret_cast<const V *>(p);
#pragma unroll
    for (int i = 0; i < (LEN + W - 1) / W; ++i) {
        const V x = pv[i];
#pragma unroll
        for (int j = 0; j < W; ++j)
            if (i * W + j < LEN) v[i * W + j] = x[j];
    }
}
// p[0..LEN) <- v: whole 16-byte vectors, the remainder with one narrower access per element pair or
// element (entries past LEN are left as they are)
template <typename T, int LEN>
__device__ __forceinline__ void lds_put(T *p, const T (&v)[LEN]) {
    using V = typename Vec16<T>::V;
    constexpr int W = Vec16<T>::W;
    V *pv = reinterpret_cast<V *>(p);
#pragma unroll
    for (int i = 0; i < LEN / W; ++i) {
        V x;
#pragma unroll
        for (int j = 0; j < W; ++j) x[j] = v[i * W + j];
        pv[i] = x;
    }
    int k = LEN / W * W;
    if constexpr (sizeof(T) == 4) {
        typedef float V2 __attribute__((ext_vector_type(2)));
        if (LEN - k >= 2) {
            V2 y;
            y[0] = v[k];
            y[1] = v[k + 1];
            *reinterpret_cast<V2 *>(p + k) = y;
            k += 2;
        }
    }
    if (k < LEN) p[k] = v[k];
}

template <typename T, int P>
struct DiagLds {  // LDS carve-up in units of T; every array and row is 16-byte aligned
    static constexpr int N = P * P;
    static constexpr int PS = vround<T>(P), NS = vround<T>(N), MS = vround<T>(N * N);
    size_t ex, in, out, inv, stage, total;
    // stage_waves: compute waves with a source stage of 64 N doubles each (staged reference-layout loads)
    __host__ __device__ DiagLds(int cthreads, int nx, int M, bool direct, int stage_waves = 0) {
        ex = 0;                                 // [2][cthreads][PS] y-traces between rows
        in = ex + (size_t)2 * cthreads * PS;    // [nx][PS] y-inflow traces of row 0
        out = in + (size_t)nx * PS;             // [2][cthreads][NS] solved cells awaiting the store
        inv = out + (direct ? 0 : (size_t)2 * cthreads * NS);  // [M][MS] element inverses (material path)
        stage = inv + (size_t)M * MS;           // [stage_waves][64 N] double
        total = stage + (size_t)stage_waves * 64 * N * (8 / sizeof(T));
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
// Raw buffer access: a resource over one array, a per-lane 32-bit byte offset and a wave-uniform
// 32-bit byte offset (the coefficient plane), so a cell's N coefficients share one address register.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void *base, unsigned long long bytes) {
    const int n = bytes >= 0xFFFFFFFFull ? -1 : (int)(unsigned)bytes;  // num_records in bytes
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, n, 0x00020000);
}
template <typename T>
__device__ __forceinline__ T buf_ld(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff) {
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
}
template <typename T>
__device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, unsigned voff, unsigned soff, T v) {
    if constexpr (sizeof(T) == 8) {
        typedef unsigned V2 __attribute__((ext_vector_type(2)));
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(V2, v), r, (int)voff, (int)soff, 0);
    } else {
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)voff, (int)soff, 0);
    }
}

// A cell's N coefficients of one ordinate in the reference layout: N contiguous doubles at byte offset
// `off` (16-byte aligned when N is even: 16-byte accesses, else 8-byte ones).
template <int N>
__device__ __forceinline__ void ref_ld(__amdgpu_buffer_rsrc_t r, unsigned off, double (&v)[N]) {
    if constexpr (N % 2 == 0) {
        typedef unsigned V4 __attribute__((ext_vector_type(4)));
        typedef unsigned V2 __attribute__((ext_vector_type(2)));
#pragma unroll
        for (int i = 0; i < N / 2; ++i) {
            const V4 x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 16 * i, 0);
            v[2 * i] = __builtin_bit_cast(double, V2{x[0], x[1]});
            v[2 * i + 1] = __builtin_bit_cast(double, V2{x[2], x[3]});
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) v[k] = buf_ld<double>(r, off, 8 * k);
    }
}
template <typename T, int N>
__device__ __forceinline__ void ref_st(__amdgpu_buffer_rsrc_t r, unsigned off, const T (&v)[N]) {
    typedef unsigned V2 __attribute__((ext_vector_type(2)));
    if constexpr (N % 2 == 0) {
        typedef unsigned V4 __attribute__((ext_vector_type(4)));
#pragma unroll
        for (int i = 0; i < N / 2; ++i) {
            const V2 a = __builtin_bit_cast(V2, (double)v[2 * i]), b = __builtin_bit_cast(V2, (double)v[2 * i + 1]);
            __builtin_amdgcn_raw_buffer_store_b128(V4{a[0], a[1], b[0], b[1]}, r, (int)off, 16 * i, 0);
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; ++k) buf_st<double>(r, off, 8 * k, (double)v[k]);
    }
}
// Byte offset of cell (c, r) (sweep coordinates of an ordinate's quadrant) of ordinate q in the reference
// layout [nx+2][ny+2][nq][N] of doubles.
__device__ __forceinline__ unsigned ref_off(int c, int r, int nx, int ny, int nq, int q, int N, bool xneg, bool yneg) {
    const int i = xneg ? nx - c : c + 1, j = yneg ? ny - r : r + 1;
    return (unsigned)((((long long)i * (ny + 2) + j) * nq + q) * N * 8);
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

// Position bookkeeping along the diagonals of a band: with off(s) = base(s) - rlo(s), cell (s, r) sits at
// off(s) + r, and off(s+1) = off(s) + min(s, ny-1) + 1 - rlo(s+1) (diagonal s holds min(s, ny-1) - rlo(s) + 1
// cells).  Wave-uniform scalar arithmetic, advanced one diagonal per step.
struct DiagWalk {
    int s, off;
    __device__ void next(int nx, int ny) {
        const int rlo1 = s + 2 - nx > 0 ? s + 2 - nx : 0;
        off += (s < ny - 1 ? s : ny - 1) + 1 - rlo1;
        ++s;
    }
};

// MODE bit 1: add Q from `qsrc`; bit 2: add sigma_s * phi / (4 pi); bit 4 (with bit 1): Q from `qref` and
// psi to `psi_ref`, both in the reference layout, so no layout conversion is needed.  There a cell's N
// values of the ordinate are one contiguous run, and the runs of a wavefront's cells lie far apart; the
// storer wave writes them as consecutive elements of the concatenated runs (lane l element l + 64 m), so
// the lanes of a store touch a few cache lines rather than one line each.  Bit 8 (with bit 4) loads the
// sources the same way into a per-wave LDS stage that hands each value to the lane of its cell (the
// address unit, not HBM, otherwise sets the step time); without it each lane reads its own run (16-byte
// accesses where N is even).
// Workgroup b*nq + q sweeps rows [b*band_rows, ...) of ordinate q (band-major numbering, so a band's
// upstream neighbour is dispatched before it and never waits on it).  Store variant ST: 0, a storer wave
// writes the cells staged in LDS; 1 (DIRECT, one band only), no storer wave, each compute lane stores its
// own cell, which frees the LDS staging buffer for larger tables; 2 (one wave per band of at most 64
// rows), the compute lanes store their cells, a row takes the y-trace of the row below from the
// neighbouring lane (a wave shift) instead of LDS, and no workgroup barrier is needed.
template <typename T>
__device__ __forceinline__ T lane_up(T v) {  // lane i receives lane i-1's v, lane 0 receives 0
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xf, 0xf, false));
    } else {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, 0x138, 0xf, 0xf, false);
        const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), 0x138, 0xf, 0xf, false);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    }
}

template <typename T, int P, bool MAT, int MODE, int D, int ST>
__global__ __launch_bounds__(512) void k_diag_sweep(DiagSweepParams<T> sp1, const DiagSweepParams<T> *__restrict__ batch,
                                                    const int2 *__restrict__ items) {
    constexpr int N = P * P;
    constexpr bool DIRECT = ST != 0, W1 = ST == 2;
    constexpr bool REF = (MODE & 4) != 0, RSTAGE = REF && (MODE & 8) != 0;
    static_assert(!(REF && W1), "reference-layout I/O is built for the storer-wave and direct-store variants");
    using PQ = typename std::conditional<REF, double, T>::type;  // prefetched source (converted at use)
    using Lds = DiagLds<T, P>;
    constexpr int PS = Lds::PS, NS = Lds::NS, MS = Lds::MS;
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
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const int nx = sp.nx, ny = sp.ny;
    const int BR = nb > 1 ? sp.band_rows : ny;  // rows per band
    const int R0 = band * BR;                    // first row of this band
    const int nrows = BR < ny - R0 ? BR : ny - R0;
    // every compute thread owns one LDS row slot; threads past the band's rows compute on clamped
    // inputs and their results are never read
    const int cthreads = DIRECT ? (int)blockDim.x : (int)blockDim.x - 64;
    const Lds L(cthreads, nx, MAT ? sp.M : 0, DIRECT, RSTAGE ? cthreads / 64 : 0);
    T *ex = lds + L.ex, *lin = lds + L.in, *lout = lds + L.out, *linv = lds + L.inv;
    // this compute wave's source stage (staged reference-layout loads)
    double *rstage = RSTAGE ? reinterpret_cast<double *>(lds + L.stage) + (size_t)(threadIdx.x >> 6) * 64 * N : nullptr;
    const DiagMap dm{nx, ny};
    const size_t ncell = (size_t)dm.cells();
    // exchange records: read from the band below (xin), written for the band above (xout)
    const gptr<U> xin = (nb > 1 && band > 0) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band - 1) * nx * P : nullptr;
    const gptr<U> xout = (nb > 1 && band < nb - 1) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band) * nx * P : nullptr;

    const int Q = q >= sp.qo3 ? 3 : (q >= sp.qo2 ? 2 : (q >= sp.qo1 ? 1 : 0));
    const int tid = threadIdx.x;
    const T xi = G(sp.xi)[q], eta = G(sp.eta)[q];
    const bool xneg = xi <= T(0), yneg = eta <= T(0);

    if (MAT) {  // the table holds dx dy A^{-1}: the material path solves with the source unscaled (see B below)
        for (int k = tid; k < sp.M * N * N; k += blockDim.x) {
            const int m = k / (N * N), e = k - m * N * N;
            linv[(size_t)m * MS + e] = G(sp.ainv)[((size_t)m * sp.nq + q) * N * N + e];
        }
    }
    if (band == 0)
        for (int k = tid; k < nx * P; k += blockDim.x) {
            const int c = k / P, d = k - c * P;
            lin[(size_t)c * PS + d] = sp.tin_y ? G(sp.tin_y)[(size_t)q * nx * P + k] : T(0);
        }

    const int steps = nx + nrows - 1;              // local step t is global diagonal R0 + t
    const int nsteps = ((steps + D - 1) / D) * D;  // padded steps have no active cell
    const unsigned plane = (unsigned)(ncell * sizeof(T));  // bytes of one coefficient plane (host-checked)
    const unsigned long long ref_bytes = REF ? (unsigned long long)(nx + 2) * (ny + 2) * sp.nq * N * 8 : 0;
    const DiagWalk walk0{R0, (int)(dm.base(R0) - dm.rlo(R0))};

    if (!DIRECT && tid >= cthreads) {
        // storer: barrier k (k >= 1) closes compute step k-1, whose cells are written here while
        // the compute waves fill the other ring slot; it also hands the band's top trace on and empties
        // the records this band has consumed
        const int lane = tid - cthreads;
        const __amdgpu_buffer_rsrc_t rpsi = REF ? buf_rsrc(sp.psi_ref, ref_bytes)
                                                : buf_rsrc(sp.psi + (size_t)q * N * ncell, (unsigned long long)N * plane);
        DiagWalk w = walk0;
        lds_barrier();
        for (int t = 0; t < nsteps; ++t) {
            lds_barrier();
            if (t >= steps) continue;
            if (xout) {
                const int c = t - (nrows - 1);
                if (c >= 0 && c < nx && lane < P)
                    xstore(xout + (size_t)c * P + lane, to_bits(ex[((size_t)(t & 1) * cthreads + (nrows - 1)) * PS + lane]));
            }
            if (xin && t < nx && lane < P) xstore(xin + (size_t)t * P + lane, ExchangeSentinel<T>::bits);
            const T *slot = lout + (size_t)(t & 1) * cthreads * NS;
            const int s = w.s;
            const int lo = s - nx + 1 > R0 ? s - nx + 1 : R0;
            const int hi = s < R0 + nrows - 1 ? s : R0 + nrows - 1;
            if constexpr (REF) {  // element e of the runs of rows lo.. (coefficient e % N of row lo + e / N)
                const int E = (hi - lo + 1) * N;
                for (int e = lane; e < E; e += 64) {
                    const int g = e / N, o = e - g * N, r = lo + g;
                    buf_st<double>(rpsi, ref_off(s - r, r, nx, ny, sp.nq, q, N, xneg, yneg) + 8u * (unsigned)o, 0,
                                   (double)slot[(size_t)(r - R0) * NS + o]);
                }
            } else {
                for (int r = lo + lane; r <= hi; r += 64) {
                    T v[N];
                    lds_get<T, N>(slot + (size_t)(r - R0) * NS, v);
                    const unsigned boff = (unsigned)(w.off + r) * (unsigned)sizeof(T);
#pragma unroll
                    for (int k = 0; k < N; ++k) buf_st<T>(rpsi, boff, k * plane, v[k]);
                }
            }
            w.next(nx, ny);
        }
        return;
    }

    const int rl = tid;     // row slot
    const int r = R0 + rl;  // mesh row (counted from the inflow edge)
    const bool row_ok = rl < nrows;
    T tx[P];
#pragma unroll
    for (int d = 0; d < P; ++d) tx[d] = (row_ok && sp.tin_x) ? G(sp.tin_x)[((size_t)q * ny + r) * P + d] : T(0);

    const unsigned long long nplanes = (unsigned long long)N * plane;
    const __amdgpu_buffer_rsrc_t rq = REF ? buf_rsrc(sp.qref, ref_bytes)
                                          : buf_rsrc(HASQ ? sp.qsrc + (size_t)q * N * ncell : nullptr, HASQ ? nplanes : 0);
    const __amdgpu_buffer_rsrc_t rout = REF ? buf_rsrc(DIRECT ? sp.psi_ref : nullptr, DIRECT ? ref_bytes : 0)
                                            : buf_rsrc(DIRECT ? sp.psi + (size_t)q * N * ncell : nullptr, DIRECT ? nplanes : 0);
    DiagWalk cw = walk0;  // diagonal of the current step (DIRECT stores)
    const __amdgpu_buffer_rsrc_t rphi = buf_rsrc(HASPHI ? sp.phi + (size_t)Q * N * ncell : nullptr, HASPHI ? nplanes : 0);
    const __amdgpu_buffer_rsrc_t rsigs = buf_rsrc(HASPHI ? sp.sig_s + (size_t)Q * ncell : nullptr, HASPHI ? plane : 0);
    const __amdgpu_buffer_rsrc_t rmat = buf_rsrc(MAT ? sp.mat + (size_t)Q * ncell : nullptr, MAT ? ncell * 4 : 0);
    const __amdgpu_buffer_rsrc_t rsigt = buf_rsrc(MAT ? nullptr : sp.sig_t + (size_t)Q * ncell, MAT ? 0 : plane);
    // the wave holding the band's first row takes that row's y-inflow from the band below (xwave is
    // wave-uniform, so the exchange code is a scalar branch the other waves skip)
    const int w0 = __builtin_amdgcn_readfirstlane(tid & ~63);  // first row slot of this wave
    // the wave's rows w0 .. w0 + live_n - nx have cells on the diagonals of local steps [w0, w0 + live_n)
    const int live_n = w0 < nrows ? (nrows - 1 - w0 < 63 ? nrows - 1 - w0 : 63) + nx : 0;
    const bool xwave = xin != nullptr && w0 == 0;
    const bool xrow = xwave && rl == 0;

    // prefetch ring: slot u holds the inputs this row meets at a step t with t % D == u.  Lanes whose row
    // has no cell on the fetched diagonal read the nearest cell of the band on it (so a wavefront's loads
    // stay within one contiguous run); padded steps read position 0.
    PQ pq[D][HASQ ? N : 1];
    T pf[D][HASPHI ? N : 1], ps[D], pm_sig[D];
    int pm[D];
    Record<U, P> px[D];
    DiagWalk fw = walk0;  // diagonal of the next fetch
    int ft = 0;           // its local step
    auto fetch = [&](int u) {
        const bool live = ft < steps;
        const int lo = live ? (fw.s - nx + 1 > R0 ? fw.s - nx + 1 : R0) : 0;
        const int hi = live ? (fw.s < R0 + nrows - 1 ? fw.s : R0 + nrows - 1) : 0;
        const int rc = r < lo ? lo : (r > hi ? hi : r);
        const unsigned p = (unsigned)((live ? fw.off : 0) + rc);
        const unsigned bt = p * (unsigned)sizeof(T), bi = p * 4u;
        if constexpr (RSTAGE) {  // element lane + 64 m of the runs of this wave's rows (clamped as above)
            const int wl = tid & 63, wrow = R0 + (tid & ~63);
#pragma unroll
            for (int m = 0; m < N; ++m) {
                const int e = wl + 64 * m, g = e / N, o = e - g * N;
                const int rg = wrow + g < lo ? lo : (wrow + g > hi ? hi : wrow + g);
                const int cl = fw.s - rg < 0 ? 0 : (fw.s - rg >= nx ? nx - 1 : fw.s - rg);
                pq[u][m] = buf_ld<double>(rq, ref_off(cl, rg, nx, ny, sp.nq, q, N, xneg, yneg) + 8u * (unsigned)o, 0);
            }
        } else if constexpr (REF) {  // the cell (clamped into the mesh on padded steps) in the reference layout
            const int cl = fw.s - rc < 0 ? 0 : (fw.s - rc >= nx ? nx - 1 : fw.s - rc);
            ref_ld<N>(rq, ref_off(cl, rc, nx, ny, sp.nq, q, N, xneg, yneg), pq[u]);
        } else if (HASQ) {
#pragma unroll
            for (int k = 0; k < N; ++k) pq[u][k] = buf_ld<T>(rq, bt, k * plane);
        }
        if (HASPHI) {
#pragma unroll
            for (int k = 0; k < N; ++k) pf[u][k] = buf_ld<T>(rphi, bt, k * plane);
            ps[u] = buf_ld<T>(rsigs, bt, 0);
        }
        if (MAT) pm[u] = buf_ld<int>(rmat, bi, 0);
        else pm_sig[u] = buf_ld<T>(rsigt, bt, 0);
        if (xwave && ft < nx) {  // every lane of the wave reads the record; row 0's lane uses it
#pragma unroll
            for (int d = 0; d < P; ++d) px[u].v[d] = xload<U>(xin + (size_t)ft * P + d);
        }
        fw.next(nx, ny);
        ++ft;
    };
#pragma unroll
    for (int u = 0; u < D; ++u) fetch(u);
    lds_barrier();

    // FP32 keeps the current material's inverse in registers and reloads it from LDS only when the
    // row enters another material (the LDS reads then move data only for the lanes that changed)
    constexpr bool REGMAT = MAT && sizeof(T) == 4 && P <= 3;  // a Q3 inverse (256 values) would spill
    T Ainv[REGMAT ? N * N : 1];
    int cur_mat = -1;
    T typrev[W1 ? P : 1];  // W1: this row's outflow y-trace of the previous step
#pragma unroll
    for (int d = 0; d < (W1 ? P : 1); ++d) typrev[d] = T(0);
    const T area = sp.dx * sp.dy;
    // B(d,e) = area src(d,e) + kx(d) Tx(e) + ky(e) Ty(d), the face-trace form of the reference's CalcB
    // (cell_solve.hpp: rhs_from_traces); kx, ky are wave-uniform.  The material path solves
    // (dx dy A^{-1}) (B / (dx dy)), so its source term needs no scaling.
    const T bsc = MAT ? T(1) : area;
    T kxd[P], kye[P];
#pragma unroll
    for (int d = 0; d < P; ++d) {
        const T kd = xneg ? T(-(2 * d + 1)) : T(ipow_m1(d) * (2 * d + 1));
        const T ke = yneg ? T(-(2 * d + 1)) : T(ipow_m1(d) * (2 * d + 1));
        if (MAT) {  // (dy xi kx) / (dx dy) = xi kx / dx, rounded once
            kxd[d] = T((double)xi * (double)kd / (double)sp.dx);
            kye[d] = T((double)eta * (double)ke / (double)sp.dy);
        } else {
            kxd[d] = (sp.dy * xi) * kd;
            kye[d] = (sp.dx * eta) * ke;
        }
    }
    // P_a(-1) = (-1)^a on an outflow face at the cell's low side; kept opaque so the traces below are
    // one multiply-add per term rather than a sign select and an add
    T sx = xneg ? T(-1) : T(1), sy = yneg ? T(-1) : T(1);
    asm volatile("" : "+v"(sx), "+v"(sy));
    for (int t0 = 0; t0 < nsteps; t0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int t = t0 + u;
            const int c = t - rl;
            const bool active = row_ok && (unsigned)c < (unsigned)nx;
            // a wave none of whose rows has a cell on this diagonal skips the step's arithmetic (its
            // results would never be read), so the other waves of the CU get its issue slots
            const bool live = (unsigned)(t - w0) < (unsigned)live_n;
            T B[N];
            if (live) {
                T ty[P];
                const int cc = c < 0 ? 0 : (c >= nx ? nx - 1 : c);
                if constexpr (W1) {  // the row below produced it in the previous step, one lane down
                    T tl[P];
                    lds_get<T, P>(lin + (size_t)cc * PS, tl);
#pragma unroll
                    for (int d = 0; d < P; ++d) {
                        ty[d] = lane_up(typrev[d]);
                        ty[d] = rl == 0 ? tl[d] : ty[d];
                    }
                } else {
                    const T *tsrc = rl == 0 ? lin + (size_t)cc * PS : ex + ((size_t)((t - 1) & 1) * cthreads + (rl - 1)) * PS;
                    lds_get<T, P>(tsrc, ty);
                }
                if constexpr (REGMAT) {  // a material change reloads the inverse now, behind the work on B
                    if (pm[u] != cur_mat) {
                        cur_mat = pm[u];
                        lds_get<T, REGMAT ? N * N : 1>(linv + (size_t)cur_mat * MS, Ainv);
                    }
                }
                if (xwave) {
                    if (xrow && active) {
                        bool ok = true;
#pragma unroll
                        for (int d = 0; d < P; ++d) ok = ok && px[u].v[d] != ExchangeSentinel<T>::bits;
                        if (!ok) px[u] = wait_record<T, P>(xin + (size_t)c * P, px[u], G(sp.abort_flag));
                    }
#pragma unroll
                    for (int d = 0; d < P; ++d) ty[d] = xrow ? (active ? from_bits<T>(px[u].v[d]) : T(0)) : ty[d];
                }
                if constexpr (RSTAGE) {  // the stage hands each lane its own cell's run
                    const int wl = tid & 63;
#pragma unroll
                    for (int m = 0; m < N; ++m) rstage[wl + 64 * m] = pq[u][m];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
#pragma unroll
                    for (int k = 0; k < N; ++k) pq[u][k] = rstage[wl * N + k];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
                }
                T a[N];  // bsc * source
                if (HASPHI) {
                    const T wsc = bsc * ps[u] * T(kInv4Pi);
#pragma unroll
                    for (int k = 0; k < N; ++k)
                        a[k] = HASQ ? fmadd(wsc, pf[u][k], MAT ? T(pq[u][k]) : bsc * T(pq[u][k])) : wsc * pf[u][k];
                } else {
#pragma unroll
                    for (int k = 0; k < N; ++k) a[k] = MAT ? T(pq[u][k]) : bsc * T(pq[u][k]);
                }
#pragma unroll
                for (int d = 0; d < P; ++d)
#pragma unroll
                    for (int e = 0; e < P; ++e) B[e + P * d] = fmadd(kye[e], ty[d], fmadd(kxd[d], tx[e], a[e + P * d]));
            }
            if (live) {
            const int mat = pm[u];
            const T sig = pm_sig[u];
            T x[N];
            if constexpr (REGMAT) {
#pragma unroll
                for (int row = 0; row < N; ++row) {
                    T acc = Ainv[row * N] * B[0];
#pragma unroll
                    for (int col = 1; col < N; ++col) acc = fmadd(Ainv[row * N + col], B[col], acc);
                    x[row] = acc;
                }
            } else if constexpr (MAT) {
                matvec_lds<T, N>(linv + (size_t)mat * MS, B, x);
            } else {
                T A[N * N];
                assemble_matrix<T, P>(A, xi, eta, sp.dx, sp.dy, (sig + sp.inv_dt) * area);
                ge_solve_rcp<T, N>(A, B);
#pragma unroll
                for (int k = 0; k < N; ++k) x[k] = B[k];
            }
            if constexpr (DIRECT) {
                if (active) {
                    if constexpr (REF) {
                        ref_st<T, N>(rout, ref_off(c, r, nx, ny, sp.nq, q, N, xneg, yneg), x);
                    } else {
                        const unsigned boff = (unsigned)(cw.off + r) * (unsigned)sizeof(T);
#pragma unroll
                        for (int k = 0; k < N; ++k) buf_st<T>(rout, boff, k * plane, x[k]);
                    }
                }
            } else {
                lds_put<T, N>(lout + ((size_t)(t & 1) * cthreads + rl) * NS, x);
            }
            T txo[P], tyo[P];  // outflow face traces: sum_a s^a psi(a, .), s = +-1
#pragma unroll
            for (int e = 0; e < P; ++e) {
                T v = x[e];
#pragma unroll
                for (int a = 1; a < P; ++a) v = (a & 1) ? v + sx * x[e + P * a] : v + x[e + P * a];
                txo[e] = v;
            }
#pragma unroll
            for (int d = 0; d < P; ++d) {
                T v = x[P * d];
#pragma unroll
                for (int b = 1; b < P; ++b) v = (b & 1) ? v + sy * x[b + P * d] : v + x[b + P * d];
                tyo[d] = v;
            }
#pragma unroll
            for (int d = 0; d < P; ++d) tx[d] = active ? txo[d] : tx[d];
            if constexpr (W1) {
#pragma unroll
                for (int d = 0; d < P; ++d) typrev[d] = tyo[d];
                // the band's top row hands its trace to the band above; the first row empties the record
                // it consumed (the records hold the sentinel again for the next sweep)
                if (xout && rl == nrows - 1 && active) {
#pragma unroll
                    for (int d = 0; d < P; ++d) xstore(xout + (size_t)c * P + d, to_bits(tyo[d]));
                }
                if (xrow && active) {
#pragma unroll
                    for (int d = 0; d < P; ++d) xstore(xin + (size_t)c * P + d, ExchangeSentinel<T>::bits);
                }
            } else {
                lds_put<T, P>(ex + ((size_t)(t & 1) * cthreads + rl) * PS, tyo);
            }
            }
            if (DIRECT) cw.next(nx, ny);
            // refill slot u with step t + D (scheduling barriers keep the slot in its registers)
            __builtin_amdgcn_sched_barrier(0);
            fetch(u);
            __builtin_amdgcn_sched_barrier(0);
            if constexpr (!W1) lds_barrier();
        }
    }
}

}  // namespace tsw

#include "sweep_split.hpp"

namespace tsw {

// rows handled by one workgroup
template <typename T>
inline int band_rows_of(const DiagSweepParams<T> &sp) {
    return sp.nbands > 1 ? sp.band_rows : sp.ny;
}

template <typename T, int P, bool MAT, int MODE, int D>
void launch_ds_d(const DiagSweepParams<T> &sp, size_t lds, hipStream_t s, const DiagSweepParams<T> *batch,
                        const int2 *items, int nitems, int max_ny) {
    // rows of a workgroup: the band (one problem) or the largest band of the batch's members (max_ny)
    const int rows = batch ? max_ny : band_rows_of(sp);
    if (sp.split) {
        if constexpr (MAT && split_supported<T, P>() && ((MODE & 4) == 0 || MODE == 5)) {
            constexpr int LG = split_lanes<P>(), RPW = 64 / LG;  // lanes per row, rows per wavefront
            const int threads = LG * ((rows + RPW - 1) / RPW * RPW);  // LG lanes for each row slot
            if (threads > 512) throw Error(TSW_E_ARG, "lane-split sweep needs bands of at most 512 / lanes rows");
            if (batch) {
                auto kern = k_diag_sweep_split<T, P, MODE, D, true>;
                if (lds > 64 * 1024)
                    TSW_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                int grid = nitems;
                if (sp.work) {  // persistent: as many workgroups as the device holds at once (the counter is zeroed)
                    int per_cu = 0, dev = 0, cus = 0;
                    TSW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, threads, lds));
                    TSW_HIP(hipGetDevice(&dev));
                    TSW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
                    grid = std::max(1, std::min(nitems, std::max(per_cu, 1) * cus));
                }
                hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, s, sp, batch, items, sp.work, nitems);
            } else {
                if constexpr ((MODE & 4) != 0) throw Error(TSW_E_ARG, "reference-layout lane-split sweep is built for batches");
                auto kern = k_diag_sweep_split<T, P, MODE, D, false>;
                if (lds > 64 * 1024)
                    TSW_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
                hipLaunchKernelGGL(kern, dim3(sp.nq * (sp.nbands > 1 ? sp.nbands : 1)), dim3(threads), lds, s, sp,
                                   nullptr, nullptr, nullptr, 0);
            }
            TSW_HIP(hipGetLastError());
            return;
        } else {
            throw Error(TSW_E_ARG, "lane-split sweep not built for this precision, degree or solve path");
        }
    }
    const bool direct = MAT && sp.direct, wave1 = sp.wave1 != 0;
    if (direct && !batch && sp.nbands > 1) throw Error(TSW_E_ARG, "direct-store sweep with row bands");
    if (wave1 && (batch || rows > 64)) throw Error(TSW_E_ARG, "one-wave sweep needs bands of at most 64 rows");
    const int threads = wave1 ? 64 : ((rows + 63) / 64) * 64 + (direct ? 0 : 64);
    constexpr bool REF = (MODE & 4) != 0;
    if (REF && wave1) throw Error(TSW_E_ARG, "reference-layout sweep without a one-wave variant");
    auto kern = (wave1 && !REF) ? k_diag_sweep<T, P, MAT, MODE, D, REF ? 0 : 2>
                      : (direct ? k_diag_sweep<T, P, MAT, MODE, D, MAT ? 1 : 0> : k_diag_sweep<T, P, MAT, MODE, D, 0>);
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
    if (rep.qref && rep.split) {  // lane-split kernel on the caller's layout (rep.prefetch or the default depth)
        if (rep.phi || !rep.ainv) throw Error(TSW_E_ARG, "reference-layout lane-split batch sweep: sources, material path");
        launch_ds<T, P, true, 5>(rep, lds, s, probs, items, nitems, max_ny);
        return;
    }
    if (rep.qref) {  // reference-layout I/O: sources in place, the default prefetch depth
        constexpr int D = sizeof(T) == 8 ? 2 : 3;
        if (rep.phi) throw Error(TSW_E_ARG, "reference-layout batch sweep: sources only");
        if (rep.ref_stage) {
            if (rep.ainv) launch_ds_d<T, P, true, 13, D>(rep, lds, s, probs, items, nitems, max_ny);
            else launch_ds_d<T, P, false, 13, D>(rep, lds, s, probs, items, nitems, max_ny);
        } else {
            if (rep.ainv) launch_ds_d<T, P, true, 5, D>(rep, lds, s, probs, items, nitems, max_ny);
            else launch_ds_d<T, P, false, 5, D>(rep, lds, s, probs, items, nitems, max_ny);
        }
        return;
    }
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

// Inflow traces of a batch's sources from their reference-layout halos, as launch_inflow computes them
// for one problem: tin_x [nq][ny][P] from the x-inflow halo cell of each row, tin_y [nq][nx][P] from the
// y-inflow halo cell of each column.  Member blockIdx.y; x strides over ordinates x rows (columns).
inline __device__ const double *inflow_source(const DiagSweepParams<double> &p) { return p.qref; }
inline __device__ const double *inflow_source(const DiagSweepParams<float> &p) { return p.qref; }
template <typename T, int P, typename Params>
__global__ void k_inflow_batch(const Params *__restrict__ batch) {
    constexpr int N = P * P;
    const Params sp = batch[blockIdx.y];
    const double *src = inflow_source(sp);
    const int nx = sp.nx, ny = sp.ny, nq = sp.nq, len = nx > ny ? nx : ny;
    for (int t = (int)(blockIdx.x * blockDim.x + threadIdx.x); t < nq * len; t += (int)(gridDim.x * blockDim.x)) {
        const int q = t / len, r = t - q * len;
        const int Q = q >= sp.qo3 ? 3 : (q >= sp.qo2 ? 2 : (q >= sp.qo1 ? 1 : 0));
        const bool xneg = Q >= 2, yneg = (Q & 1) != 0;
        if (r < ny) {  // row r: the x-inflow halo cell
            const int j = yneg ? ny - r : r + 1, i = xneg ? nx + 1 : 0;
            double tr[P];
            trace_x<double, P>(src + (((size_t)i * (ny + 2) + j) * nq + q) * N, xneg, tr);
#pragma unroll
            for (int d = 0; d < P; ++d) const_cast<T *>(sp.tin_x)[((size_t)q * ny + r) * P + d] = T(tr[d]);
        }
        if (r < nx) {  // column r: the y-inflow halo cell
            const int i = xneg ? nx - r : r + 1, j = yneg ? ny + 1 : 0;
            double tr[P];
            trace_y<double, P>(src + (((size_t)i * (ny + 2) + j) * nq + q) * N, yneg, tr);
#pragma unroll
            for (int d = 0; d < P; ++d) const_cast<T *>(sp.tin_y)[((size_t)q * nx + r) * P + d] = T(tr[d]);
        }
    }
}

template <typename T, int P>
void launch_diag_inflow_batch_p(const DiagSweepParams<T> *batch, int nprob, hipStream_t s) {
    hipLaunchKernelGGL((k_inflow_batch<T, P, DiagSweepParams<T>>), dim3(64, nprob), dim3(256), 0, s, batch);
    TSW_HIP(hipGetLastError());
}

#define TSW_SWEEP_INSTANTIATE(T, P)                                                                         \
    template void launch_diag_inflow_batch_p<T, P>(const DiagSweepParams<T> *, int, hipStream_t);          \
    template void launch_diag_sweep_p<T, P>(const DiagSweepParams<T> &, size_t, hipStream_t);              \
    template void launch_diag_sweep_batch_p<T, P>(const DiagSweepParams<T> &, const DiagSweepParams<T> *,   \
                                                  const int2 *, int, int, size_t, hipStream_t);

#define TSW_SWEEP_EXTERN(T, P)                                                                              \
    extern template void launch_diag_inflow_batch_p<T, P>(const DiagSweepParams<T> *, int, hipStream_t);   \
    extern template void launch_diag_sweep_p<T, P>(const DiagSweepParams<T> &, size_t, hipStream_t);       \
    extern template void launch_diag_sweep_batch_p<T, P>(const DiagSweepParams<T> &, const DiagSweepParams<T> *, \
                                                         const int2 *, int, int, size_t, hipStream_t);

}  // namespace tsw
