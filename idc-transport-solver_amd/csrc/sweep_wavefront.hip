// "hip-wavefront" traversal: one launch per spatial diagonal, all four quadrants in the same launch.
//
// Mirrors SweepPatternUniCartWavefront::Sweep (SweepPatternUniCartWavefront.cpp:253-296): for
// diag = 1 .. nx+ny-1, every (cell on the diagonal, ordinate) pair is independent; the quadrant of
// the ordinate flips the cell index so each quadrant walks the mesh from its own inflow corner.
// Thread t of a launch handles cell c = t / nq of the diagonal and ordinate q = t % nq, so the
// lanes of a wavefront cover consecutive ordinates of a quadrant, whose N coefficients are
// contiguous in the [i][j][q][d][e] layout.  Upwind values are the neighbours written by the
// previous launch (L2 resident); the source of the cell is read before it is overwritten.
#include "cell_solve.hpp"
#include "tsw_internal.hpp"

namespace tsw {

template <typename T, int P>
__global__ __launch_bounds__(256) void k_sweep_diag(SweepParams<T> sp, int diag, int i_hi, int len) {
    constexpr int N = P * P;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= len * sp.nq) return;
    const int c = t / sp.nq;
    const int q = t - c * sp.nq;
    int i = i_hi - c;
    int j = diag - i + 1;
    if (q >= sp.qo2) i = sp.nx - i + 1;                     // quadrants 2,3: xi < 0
    if ((q >= sp.qo1 && q < sp.qo2) || q >= sp.qo3) j = sp.ny - j + 1;  // quadrants 1,3: eta < 0

    const T xi = sp.xi[q], eta = sp.eta[q];
    const int xo = xi <= T(0) ? 1 : -1;
    const int yo = eta <= T(0) ? 1 : -1;
    const size_t stride_j = (size_t)sp.nq * N;
    const size_t stride_i = (size_t)(sp.ny + 2) * stride_j;
    T *self = sp.psi + (size_t)i * stride_i + (size_t)j * stride_j + (size_t)q * N;
    const T *upx = self + xo * (ptrdiff_t)stride_i;
    const T *upy = self + yo * (ptrdiff_t)stride_j;

    T B[N];
    assemble_rhs<T, P>(B, self, upx, upy, xi, eta, sp.dx, sp.dy);
    T A[N * N];
    if (!sp.xs) {
        const T sig = sp.sigma[(size_t)i * (sp.ny + 2) + j];
        assemble_matrix<T, P>(A, xi, eta, sp.dx, sp.dy, (sig + sp.inv_dt) * sp.dx * sp.dy);
    } else {
        // degree-p sigma_t (SweepLinearSolveSpec.cpp:300-321): dx dy / dt on the diagonal, then
        // dx dy / 4 (2d+1)(2e+1) sigma(i,j,d,e,u,v) in row (d,e), column (u,v)
        assemble_matrix<T, P>(A, xi, eta, sp.dx, sp.dy, sp.dx * sp.dy * sp.inv_dt);
        const T *tc = sp.xs + ((size_t)i * (sp.ny + 2) + j) * N * N;
        const T quarter = sp.dx * sp.dy / T(4);
#pragma unroll
        for (int uv = 0; uv < N; ++uv)
#pragma unroll
            for (int de = 0; de < N; ++de) {
                const int d = de / P, e = de - P * d;
                A[de + N * uv] += quarter * T((2 * d + 1) * (2 * e + 1)) * tc[uv * N + de];
            }
    }
    ge_solve<T, N>(A, B);
#pragma unroll
    for (int k = 0; k < N; ++k) self[k] = B[k];
}

template <typename T, int P>
static void sweep_wavefront_p(const SweepParams<T> &sp, hipStream_t s) {
    const int num_diags = sp.nx + sp.ny - 1;
    for (int diag = 1; diag <= num_diags; ++diag) {
        const int i_hi = diag < sp.nx ? diag : sp.nx;
        const int i_lo = diag - sp.ny + 1 > 1 ? diag - sp.ny + 1 : 1;
        const int len = i_hi - i_lo + 1;
        const int threads = len * sp.nq;
        const int blocks = (threads + 255) / 256;
        hipLaunchKernelGGL((k_sweep_diag<T, P>), dim3(blocks), dim3(256), 0, s, sp, diag, i_hi, len);
    }
    TSW_HIP(hipGetLastError());
}

template <typename T>
void launch_sweep_wavefront(const SweepParams<T> &sp, int p, hipStream_t s) {
    switch (p) {
        case 0: sweep_wavefront_p<T, 1>(sp, s); break;
        case 1: sweep_wavefront_p<T, 2>(sp, s); break;
        case 2: sweep_wavefront_p<T, 3>(sp, s); break;
        case 3: sweep_wavefront_p<T, 4>(sp, s); break;
        default: throw Error(TSW_E_UNSUPPORTED, "DG degree " + std::to_string(p) + " not built (0..3)");
    }
}

template void launch_sweep_wavefront<double>(const SweepParams<double> &, int, hipStream_t);
template void launch_sweep_wavefront<float>(const SweepParams<float> &, int, hipStream_t);

}  // namespace tsw
