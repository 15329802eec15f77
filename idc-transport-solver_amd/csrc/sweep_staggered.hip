// "hip-staggered" traversal: one launch per sweep, one workgroup per ordinate, one thread per mesh row.
//
// The reference's wavefront (SweepPatternUniCartWavefront.cpp:253-296) processes diagonal i+j = const
// with a barrier between diagonals.  Here each ordinate's whole sweep is one workgroup: thread r owns
// row r of the mesh (rows counted from the ordinate's inflow edge) and at step s solves the cell in
// column c = s - r (columns counted from the inflow edge), so step s is exactly one diagonal and the
// workgroup barrier at the end of the step is the diagonal barrier.  The x-upwind face trace stays in
// the thread's registers (same row, previous step); the y-upwind face trace comes from thread r-1
// through LDS (double buffered, so one barrier per step).  Ordinates are independent, so there is no
// synchronisation between workgroups.  Requires ny <= 512 (two waves per SIMD keep 256 VGPRs per thread).
#include "cell_solve.hpp"
#include "tsw_internal.hpp"

namespace tsw {

// MAT: cells carry a material index and the element solve is a product with the precomputed
// inverse of that material's matrix for this ordinate, staged in LDS; otherwise each cell's matrix
// is assembled and eliminated in registers.
template <typename T, int P, bool MAT>
__global__ __launch_bounds__(512) void k_sweep_staggered(SweepParams<T> sp) {
    constexpr int N = P * P;
    extern __shared__ unsigned char smem_raw[];
    T *lds = reinterpret_cast<T *>(smem_raw);        // [2][blockDim.x][P]  y-traces between rows
    T *lds_in = lds + (size_t)2 * blockDim.x * P;    // [nx][P]  y-inflow traces of the first row
    T *lds_inv = lds_in + (size_t)sp.nx * P;         // [M][N][N] inverses of this ordinate (MAT only)

    const int q = blockIdx.x;
    const int r = threadIdx.x;
    const int nx = sp.nx, ny = sp.ny;
    const T xi = sp.xi[q], eta = sp.eta[q];
    const bool xneg = xi <= T(0), yneg = eta <= T(0);
    const bool row_ok = r < ny;
    const int j = yneg ? ny - r : r + 1;       // mesh row of this thread
    const int j_in = yneg ? ny + 1 : 0;        // inflow halo row (used by thread 0)
    const size_t sj = (size_t)sp.nq * N, si = (size_t)(ny + 2) * sj;
    const T area = sp.dx * sp.dy;
    const size_t qoff = (size_t)q * N;
    const int di = xneg ? -1 : 1;              // step between consecutive cells of the row

    if (MAT) {
        for (int k = r; k < sp.M * N * N; k += blockDim.x) {
            const int m = k / (N * N), e = k - m * N * N;
            lds_inv[k] = sp.ainv[((size_t)m * sp.nq + q) * N * N + e];
        }
    }
    // y-inflow traces of the halo row, one per column, staged in LDS by the whole workgroup
    for (int c = r; c < nx; c += blockDim.x) {
        const int i = xneg ? nx - c : c + 1;
        T t[P];
        trace_y<T, P>(sp.psi + (size_t)i * si + (size_t)j_in * sj + qoff, yneg, t);
#pragma unroll
        for (int d = 0; d < P; ++d) lds_in[(size_t)c * P + d] = t[d];
    }

    T tx[P];  // x-face trace entering the next cell of this row
    // Prefetched inputs of this thread's next cell (column 0 first).
    T nsrc[N], nsig = T(0);
    int nmat = 0;
    int i_next = xneg ? nx : 1;
    if (row_ok) {
        trace_x<T, P>(sp.psi + (size_t)(xneg ? nx + 1 : 0) * si + (size_t)j * sj + qoff, xneg, tx);
        const T *cell = sp.psi + (size_t)i_next * si + (size_t)j * sj + qoff;
#pragma unroll
        for (int k = 0; k < N; ++k) nsrc[k] = cell[k];
        if (MAT) nmat = sp.mat[(size_t)i_next * (ny + 2) + j];
        else nsig = sp.sigma[(size_t)i_next * (ny + 2) + j];
    }
    lds_barrier();

    const int nsteps = nx + ny - 1;
    for (int s = 0; s < nsteps; ++s) {
        const int c = s - r;
        const bool active = row_ok && c >= 0 && c < nx;
        if (active) {
            const int i = i_next;
            T src[N];
            const T sig = nsig;
            const int mat = nmat;
#pragma unroll
            for (int k = 0; k < N; ++k) src[k] = nsrc[k];
            // issue the loads of the next cell of this row before working on this one
            if (c + 1 < nx) {
                i_next = i + di;
                const T *cell = sp.psi + (size_t)i_next * si + (size_t)j * sj + qoff;
#pragma unroll
                for (int k = 0; k < N; ++k) nsrc[k] = cell[k];
                if (MAT) nmat = sp.mat[(size_t)i_next * (ny + 2) + j];
                else nsig = sp.sigma[(size_t)i_next * (ny + 2) + j];
            }
            T ty[P];
            const T *src_l = r == 0 ? lds_in + (size_t)c * P : lds + ((size_t)((s - 1) & 1) * blockDim.x + (r - 1)) * P;
#pragma unroll
            for (int d = 0; d < P; ++d) ty[d] = src_l[d];
            T B[N];
            rhs_from_traces<T, P>(B, src, tx, ty, xi, eta, sp.dx, sp.dy);
            if (MAT) {
                T x[N];
                matvec_lds<T, N>(lds_inv + (size_t)mat * N * N, B, x);
#pragma unroll
                for (int k = 0; k < N; ++k) B[k] = x[k];
            } else {
                T A[N * N];
                assemble_matrix<T, P>(A, xi, eta, sp.dx, sp.dy, (sig + sp.inv_dt) * area);
                ge_solve_rcp<T, N>(A, B);
            }
            T *cell = sp.psi + (size_t)i * si + (size_t)j * sj + qoff;
#pragma unroll
            for (int k = 0; k < N; ++k) cell[k] = B[k];
            trace_x<T, P>(B, xneg, tx);
            T tyo[P];
            trace_y<T, P>(B, yneg, tyo);
            T *dst_l = lds + ((size_t)(s & 1) * blockDim.x + r) * P;
#pragma unroll
            for (int d = 0; d < P; ++d) dst_l[d] = tyo[d];
        }
        lds_barrier();
    }
}

template <typename T>
size_t staggered_lds_bytes(const SweepParams<T> &sp, int p) {
    const int P = p + 1, N = P * P;
    const size_t threads = ((size_t)(sp.ny + 63) / 64) * 64;
    if (sp.ny > 512) return 0;
    const size_t bytes = ((2 * threads + sp.nx) * P + (size_t)(sp.ainv ? sp.M : 0) * N * N) * sizeof(T);
    return bytes <= 64 * 1024 ? bytes : 0;
}

template <typename T, int P>
static void sweep_staggered_p(const SweepParams<T> &sp, hipStream_t s) {
    const int threads = ((sp.ny + 63) / 64) * 64;
    const size_t lds = staggered_lds_bytes<T>(sp, P - 1);
    if (sp.ainv)
        hipLaunchKernelGGL((k_sweep_staggered<T, P, true>), dim3(sp.nq), dim3(threads), lds, s, sp);
    else
        hipLaunchKernelGGL((k_sweep_staggered<T, P, false>), dim3(sp.nq), dim3(threads), lds, s, sp);
    TSW_HIP(hipGetLastError());
}

template <typename T>
void launch_sweep_staggered(const SweepParams<T> &sp, int p, hipStream_t s) {
    if (staggered_lds_bytes<T>(sp, p) == 0)
        throw Error(TSW_E_UNSUPPORTED, "hip-staggered needs ny <= 512 and its tables within 64 KiB of LDS");
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
