// Lane-split staggered sweep (k_diag_sweep_split), the material path with LG lanes per mesh row:
// LG = 2 for P = 2 and 4 for P = 3, 4 (for P = 3 the fourth lane repeats the third and stores nothing).
//
// Lane g of row r owns the coefficients k = e + P*g (e < P) of the row's cell.  At a step it loads the
// sources of its coefficients, forms their right-hand sides from the row's x-trace (held by every lane of
// the row) and its own y-trace value, broadcasts them to the row's other lanes with DPP quad permutations,
// applies its P rows of the element inverse to the whole right-hand side, stores its coefficients, and
// gathers the row's solution back to form the outflow traces.  A wave then holds 16 rows instead of 64
// and issues about 175 instructions per step instead of about 250 (FP32 Q2), and a band of 64 rows fills
// four wavefronts, so an ordinate's sweep spreads over four times as many SIMDs.  Every value is computed by the same
// operations, in the same order, as the material path of k_diag_sweep (sums of products are written as
// explicit fused multiply-adds in both), so the two kernels give the same bits.
//
// Bands, exchange records, prefetch ring and the wave-uniform skip of diagonals without a cell follow
// k_diag_sweep; stores are direct (no storer wave), and the band's first and last rows consume and fill
// the exchange records themselves, as the one-wave variant does.
//
// MODE bit 4 (with bit 1, sources only): the source is read from `qref` and psi written to `psi_ref`, both in
// the reference layout [i][j][q][d][e], where lane g's coefficients are one contiguous run of P doubles
// (16-byte accesses, and an 8-byte one for odd P) and the LG lanes of a row cover the cell's run, so no layout
// conversion is needed and a wavefront's accesses are RPW whole cell runs.  BATCH: the workgroup's
// parameters are batch[items[b].x], and items[b].y = band * nq + ordinate, numbered band-major so that a
// band's upstream band is always dispatched first; items with a negative problem index are padding.
#pragma once

namespace tsw {

// Lanes per row of the lane-split sweep for DG order P (P rows of the element inverse per lane).
template <int P>
__host__ __device__ constexpr int split_lanes() {
    return P == 2 ? 2 : 4;
}
// The lane-split sweep keeps a lane's P rows of the inverse in registers: built where they fit
// (Q1-Q3 in both precisions; an FP64 Q3 lane holds 64 doubles, about 230 VGPRs in all, two waves per SIMD).
template <typename T, int P>
__host__ __device__ constexpr bool split_supported() {
    return P >= 2 && P <= 4 && P * P * P * (int)sizeof(T) <= 64 * 8;
}

template <typename T, int P>
struct SplitLds {  // LDS carve-up in units of T, 16-byte aligned arrays and rows
    static constexpr int N = P * P;
    static constexpr int PS = vround<T>(P), RS = vround<T>(P * N);
    size_t ex, in, inv, total;
    // slots: row slots of the workgroup; M: materials ([M][P][RS] inverse rows, lane g's block at g)
    __host__ __device__ SplitLds(int slots, int nx, int M) {
        ex = 0;                               // [2][slots][PS] y-traces between rows
        in = ex + (size_t)2 * slots * PS;     // [nx][PS] y-inflow traces of row 0
        inv = in + (size_t)nx * PS;           // [M][P][RS]
        total = inv + (size_t)M * P * RS;
    }
};

// v read from lane K of its LG-lane group (quad permutation; K is a constant after unrolling)
template <int LG>
__device__ __forceinline__ int group_bcast_i(int v, int K) {
    if constexpr (LG == 4) {
        switch (K) {
            case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xf, 0xf, false);
            case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xf, 0xf, false);
            case 2: return __builtin_amdgcn_mov_dpp(v, 0xAA, 0xf, 0xf, false);
            default: return __builtin_amdgcn_mov_dpp(v, 0xFF, 0xf, 0xf, false);
        }
    } else {  // pairs: lanes (0,1) read lane K, lanes (2,3) read lane 2+K
        switch (K) {
            case 0: return __builtin_amdgcn_mov_dpp(v, 0xA0, 0xf, 0xf, false);
            default: return __builtin_amdgcn_mov_dpp(v, 0xF5, 0xf, 0xf, false);
        }
    }
}
template <int LG, typename T>
__device__ __forceinline__ T group_bcast(T v, int K) {
    if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, group_bcast_i<LG>(__builtin_bit_cast(int, v), K));
    } else {
        const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
        const unsigned lo = (unsigned)group_bcast_i<LG>((int)(unsigned)b, K);
        const unsigned hi = (unsigned)group_bcast_i<LG>((int)(unsigned)(b >> 32), K);
        return __builtin_bit_cast(T, ((unsigned long long)hi << 32) | lo);
    }
}

// P contiguous doubles of a reference-layout run at byte offset off: 16-byte accesses, then an 8-byte one
// for odd P (off is a multiple of 8; buffer accesses of several dwords need only dword alignment)
template <int P>
__device__ __forceinline__ void ref_run_ld(__amdgpu_buffer_rsrc_t r, unsigned off, double (&v)[P]) {
    typedef unsigned V4 __attribute__((ext_vector_type(4)));
    typedef unsigned V2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int i = 0; i < P / 2; ++i) {
        const V4 x = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 16 * i, 0);
        v[2 * i] = __builtin_bit_cast(double, V2{x[0], x[1]});
        v[2 * i + 1] = __builtin_bit_cast(double, V2{x[2], x[3]});
    }
    if constexpr (P % 2 != 0) v[P - 1] = buf_ld<double>(r, off, 8 * (P - 1));
}
// v (as doubles) to P contiguous doubles at byte offset off, in the same pieces as ref_run_ld
template <typename T, int P>
__device__ __forceinline__ void ref_run_st(__amdgpu_buffer_rsrc_t r, unsigned off, const T (&v)[P]) {
    typedef unsigned V4 __attribute__((ext_vector_type(4)));
    typedef unsigned V2 __attribute__((ext_vector_type(2)));
#pragma unroll
    for (int i = 0; i < P / 2; ++i) {
        const V2 a = __builtin_bit_cast(V2, (double)v[2 * i]), b = __builtin_bit_cast(V2, (double)v[2 * i + 1]);
        __builtin_amdgcn_raw_buffer_store_b128(V4{a[0], a[1], b[0], b[1]}, r, (int)off, 16 * i, 0);
    }
    if constexpr (P % 2 != 0) buf_st<double>(r, off, 8 * (P - 1), (double)v[P - 1]);
}

// One item of the lane-split sweep: band wid / nq of ordinate wid % nq of the problem sp, by the whole
// workgroup (every thread calls it; it ends with a workgroup barrier).
template <typename T, int P, int MODE, int D>
__device__ __forceinline__ void split_item(const DiagSweepParams<T> &sp, unsigned wid) {
    constexpr int N = P * P, LG = split_lanes<P>(), RPW = 64 / LG;  // rows per wavefront
    using Lds = SplitLds<T, P>;
    constexpr int PS = Lds::PS, RS = Lds::RS;
    using U = typename ExchangeSentinel<T>::U;
    constexpr bool HASQ = (MODE & 1) != 0, HASPHI = (MODE & 2) != 0, REF = (MODE & 4) != 0;
    static_assert(!REF || MODE == 5, "reference-layout lane-split sweep: sources only");
    using PQ = typename std::conditional<REF, double, T>::type;  // prefetched source (converted at use)
    const int nb = sp.nbands < 2 ? 1 : sp.nbands;
    const int q = (int)(wid % (unsigned)sp.nq);
    const int band = (int)(wid / (unsigned)sp.nq);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const int nx = sp.nx, ny = sp.ny;
    const int BR = nb > 1 ? sp.band_rows : ny;
    const int R0 = band * BR;
    const int nrows = BR < ny - R0 ? BR : ny - R0;
    const int slots = (int)blockDim.x / LG;
    const Lds L(slots, nx, sp.M);
    T *ex = lds + L.ex, *lin = lds + L.in, *linv = lds + L.inv;
    const DiagMap dm{nx, ny};
    const size_t ncell = (size_t)dm.cells();
    const gptr<U> xin = (nb > 1 && band > 0) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band - 1) * nx * P : nullptr;
    const gptr<U> xout = (nb > 1 && band < nb - 1) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band) * nx * P : nullptr;

    const int Q = q >= sp.qo3 ? 3 : (q >= sp.qo2 ? 2 : (q >= sp.qo1 ? 1 : 0));
    const int tid = threadIdx.x;
    const T xi = G(sp.xi)[q], eta = G(sp.eta)[q];
    const bool xneg = xi <= T(0), yneg = eta <= T(0);

    // the table holds dx dy A^{-1}, row-blocked so lane g's P rows are one aligned LDS run
    for (int k = tid; k < sp.M * N * N; k += blockDim.x) {
        const int m = k / (N * N), e = k - m * N * N, row = e / N, col = e - row * N;
        linv[((size_t)m * P + row / P) * RS + (row % P) * N + col] = G(sp.ainv)[((size_t)m * sp.nq + q) * N * N + e];
    }
    if (band == 0)
        for (int k = tid; k < nx * P; k += blockDim.x) {
            const int c = k / P, d = k - c * P;
            lin[(size_t)c * PS + d] = sp.tin_y ? G(sp.tin_y)[(size_t)q * nx * P + k] : T(0);
        }

    const int steps = nx + nrows - 1;
    const int nsteps = ((steps + D - 1) / D) * D;
    const unsigned plane = (unsigned)(ncell * sizeof(T));
    const DiagWalk walk0{R0, (int)(dm.base(R0) - dm.rlo(R0))};

    const int rl = tid / LG, g = tid % LG;
    const int dd = g < P ? g : P - 1;  // the x-degree whose coefficients this lane owns
    const bool owner = g < P;
    const int r = R0 + rl;
    const bool row_ok = rl < nrows;
    T tx[P];
#pragma unroll
    for (int d = 0; d < P; ++d) tx[d] = (row_ok && sp.tin_x) ? G(sp.tin_x)[((size_t)q * ny + r) * P + d] : T(0);

    const unsigned long long nplanes = (unsigned long long)N * plane;
    const unsigned dplane = (unsigned)(dd * P) * plane;  // byte offset of this lane's first plane (< N * plane)
    const unsigned long long ref_bytes = REF ? (unsigned long long)(nx + 2) * (ny + 2) * sp.nq * N * 8 : 0;
    const unsigned dref = 8u * (unsigned)(dd * P);  // byte offset of this lane's run within a cell's run
    const __amdgpu_buffer_rsrc_t rq = REF ? buf_rsrc(sp.qref, ref_bytes)
                                          : buf_rsrc(HASQ ? sp.qsrc + (size_t)q * N * ncell : nullptr, HASQ ? nplanes : 0);
    const __amdgpu_buffer_rsrc_t rout = REF ? buf_rsrc(sp.psi_ref, ref_bytes) : buf_rsrc(sp.psi + (size_t)q * N * ncell, nplanes);
    DiagWalk cw = walk0;
    const __amdgpu_buffer_rsrc_t rphi = buf_rsrc(HASPHI ? sp.phi + (size_t)Q * N * ncell : nullptr, HASPHI ? nplanes : 0);
    const __amdgpu_buffer_rsrc_t rsigs = buf_rsrc(HASPHI ? sp.sig_s + (size_t)Q * ncell : nullptr, HASPHI ? plane : 0);
    const __amdgpu_buffer_rsrc_t rmat = buf_rsrc(sp.mat + (size_t)Q * ncell, ncell * 4);
    const int w0 = __builtin_amdgcn_readfirstlane((tid & ~63) / LG);  // first row slot of this wave
    const int live_n = w0 < nrows ? (nrows - 1 - w0 < RPW - 1 ? nrows - 1 - w0 : RPW - 1) + nx : 0;
    const bool xwave = xin != nullptr && w0 == 0;
    const bool xrow = xwave && rl == 0;

    PQ pq[D][HASQ ? P : 1];
    T pf[D][HASPHI ? P : 1], ps[D];
    int pm[D];
    U px[D];
    DiagWalk fw = walk0;
    int ft = 0;
    auto fetch = [&](int u) {
        const bool live = ft < steps;
        const int lo = live ? (fw.s - nx + 1 > R0 ? fw.s - nx + 1 : R0) : 0;
        const int hi = live ? (fw.s < R0 + nrows - 1 ? fw.s : R0 + nrows - 1) : 0;
        const int rc = r < lo ? lo : (r > hi ? hi : r);
        const unsigned p = (unsigned)((live ? fw.off : 0) + rc);
        const unsigned bt = p * (unsigned)sizeof(T), bi = p * 4u;
        if constexpr (REF) {  // the cell (clamped into the mesh on padded steps) in the reference layout
            const int cl = fw.s - rc < 0 ? 0 : (fw.s - rc >= nx ? nx - 1 : fw.s - rc);
            ref_run_ld<P>(rq, ref_off(cl, rc, nx, ny, sp.nq, q, N, xneg, yneg) + dref, pq[u]);
        } else if (HASQ) {
#pragma unroll
            for (int e = 0; e < P; ++e) pq[u][e] = buf_ld<T>(rq, bt + dplane, e * plane);
        }
        if (HASPHI) {
#pragma unroll
            for (int e = 0; e < P; ++e) pf[u][e] = buf_ld<T>(rphi, bt + dplane, e * plane);
            ps[u] = buf_ld<T>(rsigs, bt, 0);
        }
        pm[u] = buf_ld<int>(rmat, bi, 0);
        if (xwave && ft < nx) px[u] = xload<U>(xin + (size_t)ft * P + dd);
        fw.next(nx, ny);
        ++ft;
    };
#pragma unroll
    for (int u = 0; u < D; ++u) fetch(u);
    lds_barrier();

    T Arow[P * N];  // rows dd*P .. dd*P + P-1 of the current material's dx dy A^{-1}
    int cur_mat = -1;
    T kxo, kye[P];  // face couplings divided by dx dy (rounded once), this lane's x-degree for kx
    {
        const T kd = xneg ? T(-(2 * dd + 1)) : T(ipow_m1(dd) * (2 * dd + 1));
        kxo = T((double)xi * (double)kd / (double)sp.dx);
#pragma unroll
        for (int e = 0; e < P; ++e) {
            const T ke = yneg ? T(-(2 * e + 1)) : T(ipow_m1(e) * (2 * e + 1));
            kye[e] = T((double)eta * (double)ke / (double)sp.dy);
        }
    }
    T sx = xneg ? T(-1) : T(1), sy = yneg ? T(-1) : T(1);
    asm volatile("" : "+v"(sx), "+v"(sy));
    for (int t0 = 0; t0 < nsteps; t0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int t = t0 + u;
            const int c = t - rl;
            const bool active = row_ok && (unsigned)c < (unsigned)nx;
            const bool live = (unsigned)(t - w0) < (unsigned)live_n;
            if (live) {
                const int cc = c < 0 ? 0 : (c >= nx ? nx - 1 : c);
                T ty = rl == 0 ? lin[(size_t)cc * PS + dd] : ex[((size_t)((t - 1) & 1) * slots + (rl - 1)) * PS + dd];
                if (pm[u] != cur_mat) {
                    cur_mat = pm[u];
                    lds_get<T, P * N>(linv + ((size_t)cur_mat * P + dd) * RS, Arow);
                }
                if (xwave) {
                    if (xrow && active && px[u] == ExchangeSentinel<T>::bits) {
                        Record<U, 1> rec{{px[u]}};
                        px[u] = wait_record<T, 1>(xin + (size_t)c * P + dd, rec, G(sp.abort_flag)).v[0];
                    }
                    ty = xrow ? (active ? from_bits<T>(px[u]) : T(0)) : ty;
                }
                T Bo[P];  // right-hand side of coefficients e + P dd
#pragma unroll
                for (int e = 0; e < P; ++e) {
                    T a;
                    if (HASPHI) {
                        const T wsc = T(1) * ps[u] * T(kInv4Pi);
                        a = HASQ ? fmadd(wsc, pf[u][e], T(pq[u][e])) : wsc * pf[u][e];
                    } else {
                        a = T(pq[u][e]);
                    }
                    Bo[e] = fmadd(kye[e], ty, fmadd(kxo, tx[e], a));
                }
                T B[N];
#pragma unroll
                for (int d = 0; d < P; ++d)
#pragma unroll
                    for (int e = 0; e < P; ++e) B[e + P * d] = group_bcast<LG>(Bo[e], d);
                T xo[P];
#pragma unroll
                for (int e = 0; e < P; ++e) {
                    T acc = Arow[e * N] * B[0];
#pragma unroll
                    for (int col = 1; col < N; ++col) acc = fmadd(Arow[e * N + col], B[col], acc);
                    xo[e] = acc;
                }
                if (active && owner) {
                    if constexpr (REF) {
                        ref_run_st<T, P>(rout, ref_off(c, r, nx, ny, sp.nq, q, N, xneg, yneg) + dref, xo);
                    } else {
                        const unsigned boff = (unsigned)(cw.off + r) * (unsigned)sizeof(T) + dplane;
#pragma unroll
                        for (int e = 0; e < P; ++e) buf_st<T>(rout, boff, e * plane, xo[e]);
                    }
                }
                T x[N];
#pragma unroll
                for (int d = 0; d < P; ++d)
#pragma unroll
                    for (int e = 0; e < P; ++e) x[e + P * d] = group_bcast<LG>(xo[e], d);
#pragma unroll
                for (int e = 0; e < P; ++e) {  // outflow x-trace: sum_a sx^a psi(a, e)
                    T v = x[e];
#pragma unroll
                    for (int a = 1; a < P; ++a) v = (a & 1) ? v + sx * x[e + P * a] : v + x[e + P * a];
                    tx[e] = active ? v : tx[e];
                }
                T tyo = xo[0];  // outflow y-trace of x-degree dd: sum_b sy^b psi(dd, b)
#pragma unroll
                for (int b = 1; b < P; ++b) tyo = (b & 1) ? tyo + sy * xo[b] : tyo + xo[b];
                ex[((size_t)(t & 1) * slots + rl) * PS + dd] = tyo;
                if (xout && rl == nrows - 1 && active && owner) xstore(xout + (size_t)c * P + dd, to_bits(tyo));
                if (xrow && active && owner) xstore(xin + (size_t)c * P + dd, ExchangeSentinel<T>::bits);
            }
            cw.next(nx, ny);
            __builtin_amdgcn_sched_barrier(0);
            fetch(u);
            __builtin_amdgcn_sched_barrier(0);
            lds_barrier();
        }
    }
}

// One problem (sp1, workgroup = band x ordinate), a batch with one item per workgroup (items[b] = {problem,
// band * nq + ordinate}, a negative problem marking padding), or a persistent batch launch (work != null):
// each workgroup takes the next item index from *work until nwork are taken, so items start in list
// order (band-major: a band's upstream band has always been taken, by a running workgroup, before it).
template <typename T, int P, int MODE, int D, bool BATCH>
__global__ __launch_bounds__(512) void k_diag_sweep_split(DiagSweepParams<T> sp1, const DiagSweepParams<T> *__restrict__ batch,
                                                          const int2 *__restrict__ items, int *__restrict__ work, int nwork) {
    if constexpr (!BATCH) {
        if (sp1.skip && *G(sp1.skip)) return;  // a replayed iteration after the loop has stopped
        split_item<T, P, MODE, D>(sp1, blockIdx.x);
    } else if (work == nullptr) {
        const gptr<const int> it = G(reinterpret_cast<const int *>(items)) + 2 * (size_t)blockIdx.x;
        const int item_p = it[0], item_v = it[1];
        if (item_p < 0) return;  // padding of an XCD-ordered item list
        const DiagSweepParams<T> sp = batch[item_p];
        if (sp.skip && *G(sp.skip)) return;
        split_item<T, P, MODE, D>(sp, (unsigned)item_v);
    } else {
        __shared__ int next;
        for (;;) {
            if (threadIdx.x == 0) next = atomicAdd(work, 1);
            __syncthreads();
            const int k = __builtin_amdgcn_readfirstlane(next);  // workgroup-uniform: scalar item loads
            __syncthreads();  // every thread has read `next` before it is taken again
            if (k >= nwork) break;  // workgroup-uniform: every wave leaves
            const gptr<const int> it = G(reinterpret_cast<const int *>(items)) + 2 * (size_t)k;
            const int item_p = it[0], item_v = it[1];
            if (item_p < 0) continue;
            const DiagSweepParams<T> sp = batch[item_p];
            if (sp.skip && *G(sp.skip)) continue;
            split_item<T, P, MODE, D>(sp, (unsigned)item_v);
        }
    }
}

}  // namespace tsw
