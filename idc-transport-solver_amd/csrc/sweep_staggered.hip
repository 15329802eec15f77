// "hip-staggered" traversal: one launch per sweep, one workgroup per ordinate, one compute thread per
// mesh row, plus one storer wave.
//
// The reference's wavefront (SweepPatternUniCartWavefront.cpp:253-296) processes diagonal i+j = const
// with a barrier between diagonals.  Here each ordinate's whole sweep is one workgroup: compute
// thread r owns row r of the mesh (rows counted from the ordinate's inflow edge) and at step s solves
// the cell in column c = s - r (columns counted from the inflow edge), so step s is one diagonal and
// the workgroup barrier closing the step is the diagonal barrier.  The x-upwind face trace stays in
// the thread's registers (same row, previous step); the y-upwind face trace comes from thread r-1
// through LDS.  Ordinates are independent: no synchronisation between workgroups.
//
// Memory: compute waves never store to global memory.  They read their sources D steps ahead with
// branch-free loads (columns outside the mesh are clamped to the row's first/last cell, which stay
// cache resident), so the wait for a prefetch never includes other outstanding memory operations.
// Solutions go to an LDS ring; the extra storer wave writes step s-1's cells to HBM during step s and
// never waits on memory.
#include "cell_solve.hpp"
#include "tsw_internal.hpp"

namespace tsw {

template <typename T, int P>
struct StagLayout {  // LDS carve-up, all in units of T
    static constexpr int N = P * P;
    size_t ex, in, out, inv, total;
    __host__ __device__ StagLayout(int cthreads, int nx, int ny, int M) {
        ex = 0;                                   // [2][cthreads][P]  y-traces between rows
        in = ex + (size_t)2 * cthreads * P;       // [nx][P]           y-inflow traces of row 0
        out = in + (size_t)nx * P;                // [2][ny][N]        solved cells awaiting the store
        inv = out + (size_t)2 * ny * N;           // [M][N][N]         element inverses (material path)
        total = inv + (size_t)M * N * N;
    }
};

// MAT: the element solve multiplies by the precomputed inverse of the cell's material matrix (LDS);
// otherwise the matrix is assembled and eliminated in registers.  D: prefetch distance in steps.
template <typename T, int P, bool MAT, int D>
__global__ __launch_bounds__(512) void k_sweep_staggered(SweepParams<T> sp) {
    constexpr int N = P * P;
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);
    const int nx = sp.nx, ny = sp.ny;
    const int cthreads = (int)blockDim.x - 64;  // compute threads; the last wave stores
    const StagLayout<T, P> L(cthreads, nx, ny, MAT ? sp.M : 0);
    T *ex = lds + L.ex, *lin = lds + L.in, *lout = lds + L.out, *linv = lds + L.inv;

    const int q = blockIdx.x;
    const int tid = threadIdx.x;
    const T xi = sp.xi[q], eta = sp.eta[q];
    const bool xneg = xi <= T(0), yneg = eta <= T(0);
    const int j_in = yneg ? ny + 1 : 0;
    const size_t sj = (size_t)sp.nq * N, si = (size_t)(ny + 2) * sj, qoff = (size_t)q * N;
    const T area = sp.dx * sp.dy;

    if (MAT) {
        for (int k = tid; k < sp.M * N * N; k += blockDim.x) {
            const int m = k / (N * N), e = k - m * N * N;
            linv[k] = sp.ainv[((size_t)m * sp.nq + q) * N * N + e];
        }
    }
    for (int c = tid; c < nx; c += blockDim.x) {
        const int i = xneg ? nx - c : c + 1;
        T t[P];
        trace_y<T, P>(sp.psi + (size_t)i * si + (size_t)j_in * sj + qoff, yneg, t);
#pragma unroll
        for (int d = 0; d < P; ++d) lin[(size_t)c * P + d] = t[d];
    }

    // steps beyond nx+ny-1 have no active cell; padding to a multiple of D keeps the unrolled loop
    // free of early exits
    const int nsteps = ((nx + ny - 1 + D - 1) / D) * D;
    if (tid >= cthreads) {
        // ---------------- storer wave ----------------
        // Barrier k (k >= 1) closes compute step k-1; after it this wave stores that step's cells
        // while the compute waves work on step k, writing the other ring slot.
        const int lane = tid - cthreads;
        lds_barrier();
        for (int st = 0; st < nsteps; ++st) {
            lds_barrier();
            const T *slot = lout + (size_t)(st & 1) * ny * N;
            for (int r = lane; r < ny; r += 64) {
                const int c = st - r;
                if (c < 0 || c >= nx) continue;
                const int i = xneg ? nx - c : c + 1;
                const int j = yneg ? ny - r : r + 1;
                T *dst = sp.psi + (size_t)i * si + (size_t)j * sj + qoff;
                const T *src = slot + (size_t)r * N;
#pragma unroll
                for (int k = 0; k < N; ++k) dst[k] = src[k];
            }
        }
        return;
    }

    // ---------------- compute waves ----------------
    const int r = tid;
    const bool row_ok = r < ny;
    const int j = row_ok ? (yneg ? ny - r : r + 1) : 1;
    T tx[P];
    if (row_ok) trace_x<T, P>(sp.psi + (size_t)(xneg ? nx + 1 : 0) * si + (size_t)j * sj + qoff, xneg, tx);
    auto col_i = [&](int c) {
        c = c < 0 ? 0 : (c >= nx ? nx - 1 : c);
        return xneg ? nx - c : c + 1;
    };
    // prefetch ring: slot u holds the inputs of the cell this row meets at step s with s % D == u
    T pre[D][N];
    T pre_sig[D];
    int pre_mat[D];
#pragma unroll
    for (int u = 0; u < D; ++u) {
        const int i = col_i(u - r);
        const T *cell = sp.psi + (size_t)i * si + (size_t)j * sj + qoff;
#pragma unroll
        for (int k = 0; k < N; ++k) pre[u][k] = cell[k];
        if (MAT) pre_mat[u] = sp.mat[(size_t)i * (ny + 2) + j];
        else pre_sig[u] = sp.sigma[(size_t)i * (ny + 2) + j];
    }
    lds_barrier();

    for (int s0 = 0; s0 < nsteps; s0 += D) {
#pragma unroll
        for (int u = 0; u < D; ++u) {
            const int s = s0 + u;
            const int c = s - r;
            const bool active = row_ok && c >= 0 && c < nx;
            // Consume slot u, then refill it with the cell D steps ahead (branch-free, clamped column).
            // The scheduling barriers keep the refill after the last read of the old contents, so
            // the slot keeps its registers and the wait for it covers only its own loads.
            T B[N];
            {
                T ty[P];
                const int cc = c < 0 ? 0 : (c >= nx ? nx - 1 : c);
                const T *tsrc = r == 0 ? lin + (size_t)cc * P : ex + ((size_t)((s - 1) & 1) * cthreads + (r - 1)) * P;
#pragma unroll
                for (int d = 0; d < P; ++d) ty[d] = tsrc[d];
                rhs_from_traces<T, P>(B, pre[u], tx, ty, xi, eta, sp.dx, sp.dy);
            }
            const T sig = pre_sig[u];
            const int mat = pre_mat[u];
            __builtin_amdgcn_sched_barrier(0);
            {
                const int i = col_i(c + D);
                const T *cell = sp.psi + (size_t)i * si + (size_t)j * sj + qoff;
#pragma unroll
                for (int k = 0; k < N; ++k) pre[u][k] = cell[k];
                if (MAT) pre_mat[u] = sp.mat[(size_t)i * (ny + 2) + j];
                else pre_sig[u] = sp.sigma[(size_t)i * (ny + 2) + j];
            }
            __builtin_amdgcn_sched_barrier(0);
            if (active) {
                T x[N];
                if (MAT) {
                    matvec_lds<T, N>(linv + (size_t)mat * N * N, B, x);
                } else {
                    T A[N * N];
                    assemble_matrix<T, P>(A, xi, eta, sp.dx, sp.dy, (sig + sp.inv_dt) * area);
                    ge_solve_rcp<T, N>(A, B);
#pragma unroll
                    for (int k = 0; k < N; ++k) x[k] = B[k];
                }
                T *o = lout + ((size_t)(s & 1) * ny + r) * N;
#pragma unroll
                for (int k = 0; k < N; ++k) o[k] = x[k];
                trace_x<T, P>(x, xneg, tx);
                T tyo[P];
                trace_y<T, P>(x, yneg, tyo);
                T *dst_l = ex + ((size_t)(s & 1) * cthreads + r) * P;
#pragma unroll
                for (int d = 0; d < P; ++d) dst_l[d] = tyo[d];
            }
            lds_barrier();
        }
    }
}

template <typename T>
size_t staggered_lds_bytes(const SweepParams<T> &sp, int p) {
    const int P = p + 1;
    if (sp.ny > 448) return 0;
    const int cthreads = ((sp.ny + 63) / 64) * 64;
    size_t total;
    switch (P) {
        case 1: total = StagLayout<T, 1>(cthreads, sp.nx, sp.ny, sp.ainv ? sp.M : 0).total; break;
        case 2: total = StagLayout<T, 2>(cthreads, sp.nx, sp.ny, sp.ainv ? sp.M : 0).total; break;
        case 3: total = StagLayout<T, 3>(cthreads, sp.nx, sp.ny, sp.ainv ? sp.M : 0).total; break;
        case 4: total = StagLayout<T, 4>(cthreads, sp.nx, sp.ny, sp.ainv ? sp.M : 0).total; break;
        default: return 0;
    }
    const size_t bytes = total * sizeof(T);
    return bytes <= max_lds_per_block() ? bytes : 0;
}

template <typename T, int P, bool MAT>
static void launch_stag(const SweepParams<T> &sp, size_t lds, hipStream_t s) {
    constexpr int D = sizeof(T) == 8 ? 2 : 3;
    const int threads = ((sp.ny + 63) / 64) * 64 + 64;
    auto kern = k_sweep_staggered<T, P, MAT, D>;
    if (lds > 64 * 1024) TSW_HIP(hipFuncSetAttribute((const void *)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(kern, dim3(sp.nq), dim3(threads), lds, s, sp);
    TSW_HIP(hipGetLastError());
}

template <typename T, int P>
static void sweep_staggered_p(const SweepParams<T> &sp, hipStream_t s) {
    const size_t lds = staggered_lds_bytes<T>(sp, P - 1);
    if (sp.ainv) launch_stag<T, P, true>(sp, lds, s);
    else launch_stag<T, P, false>(sp, lds, s);
}

template <typename T>
void launch_sweep_staggered(const SweepParams<T> &sp, int p, hipStream_t s) {
    if (staggered_lds_bytes<T>(sp, p) == 0)
        throw Error(TSW_E_UNSUPPORTED, "hip-staggered needs ny <= 448 and its tables within the LDS of one CU");
    switch (p) {
        case 0: sweep_staggered_p<T, 1>(sp, s); break;
        case 1: sweep_staggered_p<T, 2>(sp, s); break;
        case 2: sweep_staggered_p<T, 3>(sp, s); break;
        case 3: sweep_staggered_p<T, 4>(sp, s); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree " + std::to_string(p) + " not built (0..3)");
    }
}

template size_t staggered_lds_bytes<double>(const SweepParams<double> &, int);
template size_t staggered_lds_bytes<float>(const SweepParams<float> &, int);
template void launch_sweep_staggered<double>(const SweepParams<double> &, int, hipStream_t);
template void launch_sweep_staggered<float>(const SweepParams<float> &, int, hipStream_t);

}  // namespace tsw
