// Ordinate-lane sweep (k_lanes_sweep) on the reference layout [i][j][q][d][e].
//
// The lanes of a wavefront hold J consecutive ordinates of one quadrant ("ordinate block") times RW mesh
// rows (RW * J <= 64), and the waves of a workgroup hold consecutive row groups, so a workgroup sweeps a
// band of W * RW rows for one ordinate block.  At step t the lane of (ordinate q, row r) solves cell
// (c = t - r, r) of its quadrant's orientation: a step is one diagonal, and the workgroup barrier ending it
// is the reference wavefront's diagonal barrier (SweepPatternUniCartWavefront.cpp:253-296).  Because the
// ordinates of a cell are contiguous in the reference layout, a row group reads its cell's J * N source
// values and writes its J * N results as one contiguous run: the kernel works on the caller's arrays with
// no layout conversion.  The x-upwind trace stays in registers, the y-upwind trace passes through LDS
// (between row groups) or through the band exchange records (between bands, as in k_diag_sweep).
//
// Cell I/O is coalesced through LDS stages: lane l of a wave loads elements l, l + 64, ... of the
// concatenation of its row groups' runs, so consecutive lanes touch consecutive values, and a per-wave
// stage hands each value to the lane of its ordinate and coefficient.  Loaded per lane instead, the 64
// lanes of an access sit N values apart and the address unit, not HBM, sets the step time.  Results go the
// other way through a double-buffered stage to a storer wave (the workgroup's last wave), which writes step
// t-1 during step t: compute waves then issue no stores, so their prefetch loads are not queued behind
// stores (vector memory operations complete in issue order).
//
// Element solve: the material table dx dy A^{-1} per (material, ordinate); a lane keeps its ordinate's
// matrix for the current material in registers and reloads it (from LDS when the block's table fits,
// else from global memory) when its row enters another material.  The right-hand side, the matrix-vector
// product and the outflow traces are the same operations in the same order as in k_diag_sweep's material
// path, so both kernels give the same bits.
//
// MODE bit 1: source Q read from `src` (its halos give the inflow, as face traces tin_x / tin_y); bit 2:
// scattering source sigma_s phi / (4 pi) (Smv, TransportOperator.cpp:369-392) from the FP64 scalar flux;
// bit 4: instead of psi, the moments sum_{q in block} w_q psi_q of every cell are written per block
// (Pmv, TransportOperator.cpp:140-158, summed inside the wavefront: the lanes of a row group hold the
// block's ordinates of one cell), so a GMRES matvec writes no angular flux.
#pragma once

#include "cell_solve.hpp"
#include "sweep_kernel.hpp"
#include "tsw_internal.hpp"

namespace tsw {

template <typename T, int P>
struct LanesLds {  // LDS carve-up in bytes, every array 16-byte aligned
    static constexpr int N = P * P;
    static constexpr int PS = vround<T>(P);
    size_t ex, red, stage, out, inv, total;
    // slots = rows of the band * J; compute waves of the workgroup; M, J: material table [M][J][N*N] when
    // `table`; mom: moments instead of psi (no storer stage)
    __host__ __device__ LanesLds(int slots, int rows, int waves, int J, int M, bool mom, bool table) {
        ex = 0;                                                          // [2][slots][PS] T
        red = ex + ((size_t)2 * slots * PS * sizeof(T) + 15) / 16 * 16;  // [2][rows][N][J] double
        stage = red + (mom ? (size_t)2 * rows * N * J * sizeof(double) : 0);  // [waves][64 N] T
        out = stage + ((size_t)waves * 64 * N * sizeof(T) + 15) / 16 * 16;    // [2][waves][64 N] T
        inv = out + (mom ? 0 : ((size_t)2 * waves * 64 * N * sizeof(T) + 15) / 16 * 16);
        total = inv + (table ? ((size_t)M * J * N * N * sizeof(T) + 15) / 16 * 16 : 0);
    }
};

// vmcnt(0) (expcnt, lgkmcnt left at their maxima): wait for every outstanding vector memory operation.
constexpr int kWaitVmcnt0 = 0x0F70;

// One item of the ordinate-lane sweep: band `band` of ordinate block `bi` of the problem sp, by the whole
// workgroup (the storer wave returns after its last store, the compute waves after their last step).
template <typename T, int P, int MODE, int D>
__device__ __forceinline__ void lanes_item(const LanesParams<T> &sp, int bi, int band) {
    constexpr int N = P * P, NN = N * N;
    constexpr bool HASQ = (MODE & 1) != 0, HASPHI = (MODE & 2) != 0, MOM = (MODE & 4) != 0;
    using Lds = LanesLds<T, P>;
    constexpr int PS = Lds::PS;
    using U = typename ExchangeSentinel<T>::U;
    const int nb = sp.nbands < 2 ? 1 : sp.nbands;
    const LaneBlock bk = sp.blk[bi];
    const int nx = sp.nx, ny = sp.ny, nq = sp.nq, J = sp.J, RW = sp.RW;
    const bool xneg = bk.quad >= 2, yneg = (bk.quad & 1) != 0;
    const int BR = nb > 1 ? sp.band_rows : ny;
    const int R0 = band * BR;
    const int nrows = BR < ny - R0 ? BR : ny - R0;
    const int tid = (int)threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int gi = lane / J, j = lane - gi * J;
    const int rl = wv * RW + gi;  // row of the band (lanes past RW * J hold none)
    const bool grp_ok = gi < RW && rl < nrows;
    const bool row_ok = grp_ok && j < bk.n;
    const int q = bk.q0 + (j < bk.n ? j : 0);  // spare lanes mirror the block's first ordinate
    const int rr = R0 + (rl < nrows ? rl : nrows - 1);
    const int jc = yneg ? ny - rr : rr + 1;  // reference-layout j of the row
    const int waves = (int)(blockDim.x >> 6) - (MOM ? 0 : 1);  // compute waves (then the storer wave)
    const int rows = waves * RW;
    const int slots = rows * J;
    const int slot = gi < RW ? rl * J + j : j;            // LDS trace slot of the lane
    const int slot_up = gi < RW && rl > 0 ? slot - J : slot;  // the row below's slot
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const Lds L(slots, rows, waves, J, sp.M, MOM, sp.inv_lds != 0);
    T *ex = reinterpret_cast<T *>(smem_raw + L.ex);
    double *red = reinterpret_cast<double *>(smem_raw + L.red);
    T *stage = reinterpret_cast<T *>(smem_raw + L.stage) + (size_t)wv * 64 * N;  // this wave's stage
    T *outst = reinterpret_cast<T *>(smem_raw + L.out);                          // [2][waves][64 N]
    T *linv = reinterpret_cast<T *>(smem_raw + L.inv);

    if (sp.inv_lds)  // the block's slice of the material table: [M][J][N*N]
        for (int k = tid; k < sp.M * J * NN; k += (int)blockDim.x) {
            const int m = k / (J * NN), rest = k - m * J * NN, jj = rest / NN, e = rest - jj * NN;
            const int qq = bk.q0 + (jj < bk.n ? jj : 0);
            linv[k] = G(sp.ainv)[((size_t)m * nq + qq) * NN + e];
        }

    const int steps = nx + nrows - 1;
    const int nsteps = ((steps + D - 1) / D) * D;
    const int dstep_b = (int)((xneg ? -1LL : 1LL) * (ny + 2) * nq * N * (long long)sizeof(double));
    const unsigned long long psi_bytes = (unsigned long long)(nx + 2) * (ny + 2) * nq * N * sizeof(double);
    // Coalesced I/O: element m of lane l of compute wave w is element l + 64 m of w's runs (run g = row group
    // g's cell, J N values from the block's first ordinate; elements past the block's n N values, past the
    // band's rows or past RW runs are not moved).  Its byte offset in the cell of column c is
    // io_off + c * dstep_b (io_off < 2^31, checked by the host), its row of the band io_row.
    const int run = J * N, nrun = bk.n * N;
    auto io_elem = [&](int w, int m, int &off, int &row, bool &ok) {
        const int e = lane + 64 * m, g = e / run, o = e - g * run, rg = w * RW + g;
        ok = g < RW && o < nrun && rg < nrows;
        row = rg;
        const int rrg = R0 + (rg < nrows ? rg : nrows - 1);
        const long long c0 = (long long)(xneg ? nx : 1) * (ny + 2) + (yneg ? ny - rrg : rrg + 1);
        off = (int)(((c0 * nq + bk.q0) * N + o) * (long long)sizeof(double));
    };
    if (!MOM && wv == waves) {  // the storer wave (at most kLanesStoreWaves compute waves): step t-1 during step t
        const __amdgpu_buffer_rsrc_t rpsi = buf_rsrc(sp.psi, psi_bytes);
        int soff[kLanesStoreWaves][N], srow[kLanesStoreWaves][N];
        bool sok[kLanesStoreWaves][N];
#pragma unroll
        for (int w = 0; w < kLanesStoreWaves; ++w)
#pragma unroll
            for (int m = 0; m < N; ++m) {
                io_elem(w, m, soff[w][m], srow[w][m], sok[w][m]);
                sok[w][m] = sok[w][m] && w < waves;
            }
        lds_barrier();  // the compute waves' first barrier (after their first prefetch)
        auto store = [&](int t) {
            const T *sb = outst + (size_t)(t & 1) * waves * 64 * N;
#pragma unroll
            for (int w = 0; w < kLanesStoreWaves; ++w)
#pragma unroll
                for (int m = 0; m < N; ++m) {
                    const int cm = t - srow[w][m];
                    if (sok[w][m] && (unsigned)cm < (unsigned)nx)
                        buf_st<double>(rpsi, (unsigned)soff[w][m] + (unsigned)cm * (unsigned)dstep_b, 0,
                                       (double)sb[(size_t)w * 64 * N + lane + 64 * m]);
                }
        };
        for (int t = 0; t < nsteps; ++t) {
            if (t > 0) store(t - 1);
            lds_barrier();
        }
        store(nsteps - 1);
        return;
    }

    const gptr<U> xin = (nb > 1 && band > 0) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band - 1) * nx * P : nullptr;
    const gptr<U> xout = (nb > 1 && band < nb - 1) ? G(reinterpret_cast<U *>(sp.xtr)) + ((size_t)q * (nb - 1) + band) * nx * P : nullptr;
    const gptr<const U> tiny = (HASQ && band == 0 && sp.tin_y) ? G(reinterpret_cast<const U *>(sp.tin_y)) + (size_t)q * nx * P : nullptr;
    const bool yrow = rl == 0;  // takes its y-inflow from tin_y (band 0) or the record of the band below
    const bool ywave = __builtin_amdgcn_readfirstlane(wv) == 0;

    // this lane's row: cell (c, row) sits at reference cell (ic, jc), ic = c + 1 or nx - c
    const long long cell0 = ((long long)(xneg ? nx : 1) * (ny + 2) + jc);  // reference cell of c = 0
    const int mstep = xneg ? -(ny + 2) : (ny + 2);

    const int w0 = __builtin_amdgcn_readfirstlane(wv * RW);  // first row of this wave
    const int wrows = w0 < nrows ? (nrows - w0 < RW ? nrows - w0 : RW) : 0;
    const int live_n = wrows > 0 ? wrows - 1 + nx : 0;  // steps in [w0, w0 + live_n) touch this wave's rows

    int io_off[HASQ ? N : 1], io_row[HASQ ? N : 1];  // this lane's source elements
    if (HASQ) {
#pragma unroll
        for (int m = 0; m < N; ++m) {
            bool ok;
            io_elem(wv, m, io_off[m], io_row[m], ok);
        }
    }
    const __amdgpu_buffer_rsrc_t rsrc = buf_rsrc(HASQ ? sp.src : nullptr, HASQ ? psi_bytes : 0);
    // this lane's values in the stage: lane = gi J + j, so the row groups' runs are the first RW J N entries
    // and lanes past them (which compute nothing useful) write to the unused tail
    const int own = lane * N;

    T tx[P];  // x-inflow trace of the row (tin_x), then the outflow trace of the last cell
#pragma unroll
    for (int d = 0; d < P; ++d)
        tx[d] = (HASQ && sp.tin_x) ? G(sp.tin_x)[((size_t)q * ny + rr) * P + d] : T(0);

    double pq[D][HASQ ? N : 1];  // the source, in the coalesced arrangement (lane element m)
    double pf[D][HASPHI ? N : 1], ps[D];
    int pm[D];
    U py[D][P];
    int ft = 0;
    auto fetch = [&](int u) {
        const int c = ft - rl;
        const int cc = c < 0 ? 0 : (c >= nx ? nx - 1 : c);
        const long long cell = cell0 + (long long)cc * mstep;
        if (HASQ) {  // elements of cells outside the mesh read a valid address or, past the array, zero
#pragma unroll
            for (int m = 0; m < N; ++m) {
                const int cm = ft - io_row[m];
                pq[u][m] = buf_ld<double>(rsrc, (unsigned)io_off[m] + (unsigned)cm * (unsigned)dstep_b, 0);
            }
        }
        if (HASPHI) {
            const gptr<const double> p = G(sp.phi) + cell * N;
#pragma unroll
            for (int k = 0; k < N; ++k) pf[u][k] = p[k];
            ps[u] = G(sp.sig_s)[cell];
        }
        pm[u] = G(sp.mat)[cell];
        if (ywave && yrow) {
            if (xin) {
#pragma unroll
                for (int d = 0; d < P; ++d) py[u][d] = xload<U>(xin + (size_t)cc * P + d);
            } else if (tiny) {
#pragma unroll
                for (int d = 0; d < P; ++d) py[u][d] = tiny[(size_t)cc * P + d];
            }
        }
        ++ft;
    };
#pragma unroll
    for (int u = 0; u < D; ++u) fetch(u);
    lds_barrier();

    T A[NN];  // dx dy A^{-1} of the current material, this lane's ordinate
    int cur_mat = -1;
    T kx[P], ky[P];  // face couplings divided by dx dy (rounded once)
    {
        const T xi = G(sp.xi)[q], eta = G(sp.eta)[q];
#pragma unroll
        for (int d = 0; d < P; ++d) {
            const T kd = xneg ? T(-(2 * d + 1)) : T(ipow_m1(d) * (2 * d + 1));
            const T ke = yneg ? T(-(2 * d + 1)) : T(ipow_m1(d) * (2 * d + 1));
            kx[d] = T((double)xi * (double)kd / (double)sp.dx);
            ky[d] = T((double)eta * (double)ke / (double)sp.dy);
        }
    }
    T sx = xneg ? T(-1) : T(1), sy = yneg ? T(-1) : T(1);
    asm volatile("" : "+v"(sx), "+v"(sy));
    const double wq = MOM ? G(sp.w)[q] : 0.0;
    double *mom = MOM ? sp.mom + (size_t)bi * (size_t)(nx + 2) * (ny + 2) * N : nullptr;

    // moments of step t-1 (MOM): lane j of a row group sums coefficient k = j, j + J, ... over the block
    auto reduce = [&](int t) {
        const int c = t - rl;
        if (grp_ok && (unsigned)c < (unsigned)nx) {
            const double *rb = red + ((size_t)(t & 1) * rows + rl) * N * J;
            const long long cell = cell0 + (long long)c * mstep;
            for (int k = j; k < N; k += J) {
                double s = rb[k * J];
                for (int jj = 1; jj < bk.n; ++jj) s += rb[k * J + jj];
                mom[cell * N + k] = s;
            }
        }
    };

    for (int t0 = 0; t0 < nsteps; t0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int t = t0 + u;
            if (MOM && t > 0) reduce(t - 1);
            const int c = t - rl;
            const bool active = row_ok && (unsigned)c < (unsigned)nx;
            const bool live = (unsigned)(t - w0) < (unsigned)live_n;
            if (live) {
                const int cc = c < 0 ? 0 : (c >= nx ? nx - 1 : c);
                T ty[P];
                if (yrow) {
                    if (xin) {
                        bool empty = false;  // the record is filled value by value: wait for all P
#pragma unroll
                        for (int d = 0; d < P; ++d) empty = empty || py[u][d] == ExchangeSentinel<T>::bits;
                        if (active && empty) {
                            Record<U, P> rec;
#pragma unroll
                            for (int d = 0; d < P; ++d) rec.v[d] = py[u][d];
                            rec = wait_record<T, P>(xin + (size_t)cc * P, rec, G(sp.abort_flag));
#pragma unroll
                            for (int d = 0; d < P; ++d) py[u][d] = rec.v[d];
                        }
#pragma unroll
                        for (int d = 0; d < P; ++d) ty[d] = active ? from_bits<T>(py[u][d]) : T(0);
                    } else {
#pragma unroll
                        for (int d = 0; d < P; ++d) ty[d] = tiny ? from_bits<T>(py[u][d]) : T(0);
                    }
                } else {
                    lds_get<T, P>(ex + ((size_t)((t - 1) & 1) * slots + slot_up) * PS, ty);
                }
                if (pm[u] != cur_mat) {
                    cur_mat = pm[u];
                    if (sp.inv_lds) {
                        const T *src = linv + ((size_t)cur_mat * J + j) * NN;
#pragma unroll
                        for (int k = 0; k < NN; ++k) A[k] = src[k];
                    } else {
                        const gptr<const T> src = G(sp.ainv) + ((size_t)cur_mat * nq + q) * NN;
#pragma unroll
                        for (int k = 0; k < NN; ++k) A[k] = src[k];
                        // Wait here, on the rare reload: vector memory operations complete in issue order,
                        // so a wait for A after the branch (where the compiler would put it) would also
                        // wait for the prefetches in flight on every step, reloaded or not.
                        __builtin_amdgcn_s_waitcnt(kWaitVmcnt0);
                    }
                }
                T q_own[HASQ ? N : 1];  // this lane's source values, through the stage
                if (HASQ) {
#pragma unroll
                    for (int m = 0; m < N; ++m) stage[lane + 64 * m] = T(pq[u][m]);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
#pragma unroll
                    for (int k = 0; k < N; ++k) q_own[k] = stage[own + k];
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
                }
                T B[N];
#pragma unroll
                for (int d = 0; d < P; ++d)
#pragma unroll
                    for (int e = 0; e < P; ++e) {
                        const int k = e + P * d;
                        T a;
                        if (HASPHI) {
                            const T wsc = T(1) * T(ps[u]) * T(kInv4Pi);
                            a = HASQ ? fmadd(wsc, T(pf[u][k]), q_own[k]) : wsc * T(pf[u][k]);
                        } else {
                            a = q_own[k];
                        }
                        B[k] = fmadd(ky[e], ty[d], fmadd(kx[d], tx[e], a));
                    }
                T x[N];
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    T acc = A[k * N] * B[0];
#pragma unroll
                    for (int col = 1; col < N; ++col) acc = fmadd(A[k * N + col], B[col], acc);
                    x[k] = acc;
                }
                if (MOM) {
                    double *rb = red + ((size_t)(t & 1) * rows + (gi < RW ? rl : 0)) * N * J;
                    if (grp_ok) {
#pragma unroll
                        for (int k = 0; k < N; ++k) rb[k * J + j] = wq * (double)x[k];
                    }
                } else {  // to the storer wave (stored during the next step)
                    T *ob = outst + ((size_t)(t & 1) * waves + wv) * 64 * N;
#pragma unroll
                    for (int k = 0; k < N; ++k) ob[own + k] = x[k];
                }
#pragma unroll
                for (int e = 0; e < P; ++e) {  // outflow x-trace: sum_a sx^a psi(a, e)
                    T v = x[e];
#pragma unroll
                    for (int a = 1; a < P; ++a) v = (a & 1) ? v + sx * x[e + P * a] : v + x[e + P * a];
                    tx[e] = active ? v : tx[e];
                }
                T to[P];  // outflow y-trace: sum_b sy^b psi(d, b)
#pragma unroll
                for (int d = 0; d < P; ++d) {
                    T v = x[P * d];
#pragma unroll
                    for (int b = 1; b < P; ++b) v = (b & 1) ? v + sy * x[b + P * d] : v + x[b + P * d];
                    to[d] = v;
                }
                if (gi < RW) lds_put<T, P>(ex + ((size_t)(t & 1) * slots + slot) * PS, to);
                if (xout && rl == nrows - 1 && active) {
#pragma unroll
                    for (int d = 0; d < P; ++d) xstore(xout + (size_t)c * P + d, to_bits(to[d]));
                }
                if (xin && yrow && active) {
#pragma unroll
                    for (int d = 0; d < P; ++d) xstore(xin + (size_t)c * P + d, ExchangeSentinel<T>::bits);
                }
            }
            __builtin_amdgcn_sched_barrier(0);
            fetch(u);
            __builtin_amdgcn_sched_barrier(0);
            lds_barrier();
        }
    }
    if (MOM) reduce(nsteps - 1);
}

// One problem (sp1; workgroup = band x ordinate block) or, with BATCH, a batch: items[k] = {problem, block,
// band} with parameters batch[problem], band-major (a band's upstream neighbour comes first); one item per
// workgroup, or (work != null) a persistent launch whose workgroups take item indices in order from *work
// until nwork are taken, so an item's upstream band has always been taken by a running workgroup.
template <typename T, int P, int MODE, int D, bool BATCH>
__global__ __launch_bounds__(256) void k_lanes_sweep(LanesParams<T> sp1, const LanesParams<T> *__restrict__ batch,
                                                     const int4 *__restrict__ items, int *__restrict__ work,
                                                     int nwork) {
    if constexpr (!BATCH) {
        if (sp1.skip && *G(sp1.skip)) return;  // a replayed iteration after the loop has stopped
        lanes_item<T, P, MODE, D>(sp1, (int)(blockIdx.x % (unsigned)sp1.nblk), (int)(blockIdx.x / (unsigned)sp1.nblk));
    } else if (work == nullptr) {
        const int4 item = items[blockIdx.x];
        const LanesParams<T> sp = batch[item.x];
        if (sp.skip && *G(sp.skip)) return;
        lanes_item<T, P, MODE, D>(sp, item.y, item.z);
    } else {
        __shared__ int next;
        for (;;) {
            if (threadIdx.x == 0) next = atomicAdd(work, 1);
            __syncthreads();
            const int k = __builtin_amdgcn_readfirstlane(next);  // workgroup-uniform: scalar item loads
            __syncthreads();  // every thread has read `next` before it is taken again
            if (k >= nwork) break;  // workgroup-uniform: every wave leaves
            const int4 item = items[k];
            const LanesParams<T> sp = batch[item.x];
            if (sp.skip && *G(sp.skip)) continue;
            lanes_item<T, P, MODE, D>(sp, item.y, item.z);
        }
    }
}

template <typename T, int P, int MODE, int D>
void launch_lanes_dd(const LanesParams<T> &sp, int threads, size_t lds, hipStream_t s, const LanesParams<T> *batch,
                     const int4 *items, int nitems) {
    auto go = [&](auto kern, int blocks) {
        if (lds > 64 * 1024)
            TSW_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        int grid = blocks;
        if (batch && sp.work) {  // persistent: as many workgroups as the device holds at once (counter zeroed)
            int per_cu = 0, dev = 0, cus = 0;
            TSW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)kern, threads, lds));
            TSW_HIP(hipGetDevice(&dev));
            TSW_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
            grid = std::max(1, std::min(blocks, std::max(per_cu, 1) * cus));
        }
        hipLaunchKernelGGL(kern, dim3(grid), dim3(threads), lds, s, sp, batch, items, batch ? sp.work : nullptr,
                           blocks);
        TSW_HIP(hipGetLastError());
    };
    if (!batch) {
        go(k_lanes_sweep<T, P, MODE, D, false>, sp.nblk * (sp.nbands > 1 ? sp.nbands : 1));
    } else if constexpr (MODE == 1) {  // batches sweep sources in place
        go(k_lanes_sweep<T, P, MODE, D, true>, nitems);
    } else {
        throw Error(TSW_E_ARG, "ordinate-lane batch: sources only");
    }
}

// Prefetch depth D (steps of cell inputs in flight per lane): sp.prefetch (FP64 2 or 4, FP32 3 or 6), by
// default 2 in FP64 and 3 in FP32 (the deeper rings measured 3-9% slower on the 280x280 Q2 S12 lattice:
// the step is set by the lane's own instruction stream, not by memory latency).
template <typename T, int P, int MODE>
void launch_lanes_d(const LanesParams<T> &sp, int threads, size_t lds, hipStream_t s, const LanesParams<T> *batch,
                    const int4 *items, int nitems) {
    constexpr bool F64 = sizeof(T) == 8;
    const int D = sp.prefetch > 0 ? sp.prefetch : (F64 ? 2 : 3);
    if (D == (F64 ? 2 : 3)) launch_lanes_dd<T, P, MODE, F64 ? 2 : 3>(sp, threads, lds, s, batch, items, nitems);
    else if (D == (F64 ? 4 : 6)) launch_lanes_dd<T, P, MODE, F64 ? 4 : 6>(sp, threads, lds, s, batch, items, nitems);
    else throw Error(TSW_E_ARG, F64 ? "ordinate-lane prefetch depth must be 2 or 4 in FP64"
                                    : "ordinate-lane prefetch depth must be 3 or 6 in FP32");
}

// Entry point per (precision, DG order P); MODE from the params (src, phi, mom), those of `sp` for a batch
// (whose members share the mode).
template <typename T, int P>
void launch_lanes_sweep_p(const LanesParams<T> &sp, int threads, size_t lds, hipStream_t s,
                          const LanesParams<T> *batch, const int4 *items, int nitems) {
    const int mode = (sp.src ? 1 : 0) | (sp.phi ? 2 : 0) | (sp.mom ? 4 : 0);
    switch (mode) {
        case 1: launch_lanes_d<T, P, 1>(sp, threads, lds, s, batch, items, nitems); break;
        case 3: launch_lanes_d<T, P, 3>(sp, threads, lds, s, batch, items, nitems); break;
        case 5: launch_lanes_d<T, P, 5>(sp, threads, lds, s, batch, items, nitems); break;
        case 6: launch_lanes_d<T, P, 6>(sp, threads, lds, s, batch, items, nitems); break;
        case 7: launch_lanes_d<T, P, 7>(sp, threads, lds, s, batch, items, nitems); break;
        default: throw Error(TSW_E_ARG, "ordinate-lane sweep: unsupported source / output combination");
    }
}

inline __device__ const double *inflow_source(const LanesParams<double> &p) { return p.src; }
inline __device__ const double *inflow_source(const LanesParams<float> &p) { return p.src; }

template <typename T, int P>
void launch_lanes_inflow_batch_p(const LanesParams<T> *batch, int nprob, hipStream_t s) {
    hipLaunchKernelGGL((k_inflow_batch<T, P, LanesParams<T>>), dim3(64, nprob), dim3(256), 0, s, batch);
    TSW_HIP(hipGetLastError());
}

#define TSW_LANES_INSTANTIATE(T, P)                                                                          \
    template void launch_lanes_sweep_p<T, P>(const LanesParams<T> &, int, size_t, hipStream_t,               \
                                             const LanesParams<T> *, const int4 *, int);                     \
    template void launch_lanes_inflow_batch_p<T, P>(const LanesParams<T> *, int, hipStream_t);

}  // namespace tsw
