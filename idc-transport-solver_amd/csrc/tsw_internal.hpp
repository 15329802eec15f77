// Internal declarations shared by the libtsw.so translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "tsw.h"

namespace tsw {

// Error raised inside the library; converted to a status code at the C-ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define TSW_HIP(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            throw ::tsw::Error(TSW_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

// Owning device allocation.
struct DBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    DBuf(DBuf &&o) noexcept : ptr(o.ptr), bytes(o.bytes) {
        o.ptr = nullptr;
        o.bytes = 0;
    }
    DBuf &operator=(DBuf &&o) noexcept {
        if (this != &o) {
            release();
            ptr = o.ptr;
            bytes = o.bytes;
            o.ptr = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DBuf() { release(); }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    // Grows (never shrinks) to at least n bytes; contents are not preserved.
    void *reserve(size_t n) {
        if (n > bytes) {
            release();
            TSW_HIP(hipMalloc(&ptr, n));
            bytes = n;
        }
        return ptr;
    }
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Mesh and ordinate description shared by every kernel launch.
struct Geom {
    int nx = 0, ny = 0, p = 0, nq = 0;
    double ax = 0, bx = 1, ay = 0, by = 1;
    double dx = 1, dy = 1;
    int quad_off[5] = {0, 0, 0, 0, 0};
    int N() const { return (p + 1) * (p + 1); }
    size_t ncell_h() const { return (size_t)(nx + 2) * (ny + 2); }  // cells incl. halo
    size_t npsi() const { return ncell_h() * nq * N(); }
    size_t nphi() const { return ncell_h() * N(); }
};

// Kernel-side view of the ordinates and mesh in precision T.
template <typename T> struct SweepParams {
    T *psi;           // in place: source in, angular flux out (halo = inflow)
    const T *sigma;   // [nx+2][ny+2] cell means of sigma_t
    const T *xi;      // [nq]
    const T *eta;     // [nq]
    int nx, ny, nq;
    int qo1, qo2, qo3;  // quadrant offsets 1..3
    T dx, dy, inv_dt;
    const T *xs;      // degree-p sigma_t: its triple-product tensor (launch_xs_tensor), replaces sigma; or null
};

// Triple products of Legendre polynomials on [-1, 1], tpi[a + P(b + P c)] = int P_a P_b P_c, P <= 4
// (ComputeTPI, global.cpp:67-107).
struct Tpi {
    double v[64];
};
// Tensor of a degree-p cross section from its Legendre coefficients sig[cell][u][v] (cells incl. halo):
// t[cell][d*P+e][a*P+b] = sum_{u,v} sig[cell][u][v] tpi(a,d,u) tpi(b,e,v) (CrossSection::ComputeTensor,
// CrossSection.cpp:165-189; TensorIndex CrossSection.hpp:397-478), so the coefficient sigma(i,j,d,e,u,v)
// of CalcA, Smv and Lmv sits at t[cell][u*P+v][d*P+e].
void launch_xs_tensor(const Geom &g, const double *sig, const Tpi &tpi, double *t, hipStream_t s);
// out[cell] = sig[cell][0][0] (cell means of a degree-p cross section)
void launch_xs_means(const Geom &g, const double *sig, double *out, hipStream_t s);

// Sweep on the diagonal-major layout (diag_layout.hpp).  Source of a cell: qsrc (if set) plus
// sig_s * phi / (4 pi) (if phi is set).  Inflow traces tin_x/tin_y (null: vacuum).  Element solve by
// the material table (mat + ainv) or per-cell elimination with sig_t.
template <typename T> struct DiagSweepParams {
    const T *xi, *eta;     // [nq]
    int nx, ny, nq, qo1, qo2, qo3;
    T dx, dy, inv_dt;
    const T *qsrc;         // [nq][N][ncell] or null (may alias psi: in-place sweep)
    const T *phi;          // [4][N][ncell] or null
    const T *sig_s;        // [4][ncell] (with phi)
    const T *tin_x;        // [nq][ny][P] or null
    const T *tin_y;        // [nq][nx][P] or null
    const int32_t *mat;    // [4][ncell] (with ainv)
    const T *ainv;         // [M][nq][N][N] or null
    int M;
    const T *sig_t;        // [4][ncell] (without ainv)
    T *psi;                // out [nq][N][ncell]
    // Row bands: nbands > 1 splits each ordinate's rows into bands of band_rows rows, one workgroup
    // each, pipelined through xtr ([nq][nbands-1][nx][P] y-traces of each band's last row, holding the
    // sentinel pattern when empty); a band that waits past the spin limit sets *abort_flag.
    int band_rows, nbands;
    T *xtr;
    int *abort_flag;
    int prefetch;          // prefetch depth in steps (0: default for the precision)
    int direct;            // 1: no storer wave, the compute waves store psi (one band only; used when
                           // only that LDS layout, without the store staging buffer, fits)
    int wave1;             // 1: one wave per band of at most 64 rows, traces passed between lanes
    int split;             // 1: lane-split kernel (2 or 4 lanes per row, material path, direct stores)
    const int *skip;       // device flag: the launch does nothing while *skip != 0 (null: always runs)
    // Reference-layout I/O (batch members): Q from qref and psi to psi_ref ([nx+2][ny+2][nq][N] doubles,
    // may alias, 16-byte aligned) instead of qsrc / psi
    const double *qref;
    double *psi_ref;
    int ref_stage;         // 1: reference-layout sources loaded through a per-wave LDS stage (coalesced)
    // Batched lane-split launch only (host-side dispatch): work counter of a persistent launch, whose
    // workgroups take items in order from *work until all are taken (null: one item per workgroup)
    int *work;
};

// Ordinate-lane sweep on the reference layout (sweep_lanes.hpp).  An ordinate block is ordinates
// [q0, q0 + n) of quadrant `quad`; a row group of a wavefront holds one block's ordinates of one row.
struct LaneBlock {
    int q0, n, quad, pad;
};
template <typename T> struct LanesParams {
    const T *xi, *eta;          // [nq]
    int nx, ny, nq, qo1, qo2, qo3;
    T dx, dy;
    const LaneBlock *blk;       // [nblk]
    int nblk, J, RW;            // lanes per row group (>= every block's n), row groups per wave (RW * J <= 64)
    int prefetch;               // steps of inputs in flight per lane (0: default for the precision)
    const double *src;          // bit 1: Q [nx+2][ny+2][nq][N]
    const T *tin_x, *tin_y;     // with src: inflow traces [nq][ny][P], [nq][nx][P] of its halos
    double *psi;                // out unless mom: interior of [nx+2][ny+2][nq][N] (may alias src)
    const int32_t *mat;         // [nx+2][ny+2]
    const T *ainv;              // [M][nq][N][N]: dx dy A^{-1}
    int M, inv_lds;             // inv_lds: stage the blocks' tables in LDS
    const double *phi;          // bit 2: [nx+2][ny+2][N]
    const double *sig_s;        // bit 2: [nx+2][ny+2]
    double *mom;                // bit 4: [nblk][nx+2][ny+2][N], sum over the block of w_q psi_q (interior)
    const double *w;            // bit 4: [nq]
    int band_rows, nbands;      // rows per band (a workgroup of band_rows / RW waves) and bands per block
    T *xtr;                     // [nq][nbands-1][nx][P] band exchange records
    int *abort_flag;
    const int *skip;            // device flag: the launch does nothing while *skip != 0 (null: always runs)
    int *work;                  // batch launches (host-side dispatch): counter of a persistent launch, or null
};
constexpr int kLanesStoreWaves = 3;  // compute waves per workgroup when a storer wave writes psi
template <typename T, int P>
void launch_lanes_sweep_p(const LanesParams<T> &sp, int threads, size_t lds, hipStream_t s,
                          const LanesParams<T> *batch = nullptr, const int4 *items = nullptr, int nitems = 0);
// tin_x / tin_y of every member of a batch from its src (device parameter table of nprob members)
template <typename T, int P>
void launch_lanes_inflow_batch_p(const LanesParams<T> *batch, int nprob, hipStream_t s);
// OrdinateFlux::ReflectBoundaries on the edges of `mask` (1 x-min, 2 x-max, 4 y-min, 8 y-max); map [2][nq]
void launch_reflect_boundaries(const Geom &g, double *psi, const int *map, int mask, hipStream_t s);
// DensityFunction::SynchronizeHalos across periodic boundaries (the x- and y-halos of rows and columns 1..n)
void launch_periodic_halos(const Geom &g, double *psi, hipStream_t s);
// dst <- src except the ghost cells of the reflecting edges in `mask`
void launch_copy_outside_reflect(const Geom &g, const double *src, double *dst, int mask, hipStream_t s);
// y = beta xb + alpha sum_b mom[b] on the interior of phi-layout arrays (mom: nblk partials of g.nphi())
void launch_mom_sum(const Geom &g, int nblk, const double *mom, const double *xb, double *y, double alpha,
                    double beta, hipStream_t s, const int *skip = nullptr);

// Device-side state of a source iteration replayed from a graph (SISolver.cpp:219-241).
struct SiControl {
    int done;     // set when the loop test fails; every kernel of an iteration then does nothing
    int iters;    // iterations run
    int max_its;
    int pad;
    double tol;   // absolute l2 change of the scalar-flux coefficients
    double err;   // last change
};
void launch_si_check(const Geom &g, const double *cur, const double *last, double *partials, int nblocks,
                     SiControl *ctl, hipStream_t s);

// One multigrid level of the DSA diffusion operator (dsa.hip): device arrays.
struct MgLevel {
    int nx, ny;
    double *cx, *cy, *sv;  // face conductances ((nx+1)*ny, nx*(ny+1)), absorption x volume (nx*ny)
    double *x, *b, *y;     // pre-smoothed iterate, right-hand side, post-smoothed result (nx*ny each)
    const double *inv;     // coarsest level: dense inverse of its operator (row-major n x n), else null
};
constexpr int kMgMaxLevels = 16;
void launch_dsa_rhs(const Geom &g, const double *x, const double *sigma_s, double *b, hipStream_t s);
void launch_dsa_add(const Geom &g, const double *d, double *y, hipStream_t s);
// two tile-wise red-black sweeps on level L from in (null: zero; coarse non-null: in plus the
// prolongated coarse result coarse->y) into out, colour c0 first; the fused residual-and-restriction
void launch_mg_smooth(const MgLevel &L, const double *in, double *out, const MgLevel *coarse, int c0, hipStream_t s);
void launch_mg_restrict_residual(const MgLevel &f, const MgLevel &c, hipStream_t s);
// LDS doubles the single-workgroup V-cycle needs for levels l0..nlev-1 (host arrays of levels)
size_t mg_fused_lds_doubles(const MgLevel *lv, int l0, int nlev);
void launch_mg_fused(const MgLevel *lv_dev, int l0, int nlev, size_t lds_bytes, bool from_y, hipStream_t s);

// Bit pattern of an empty band-exchange record: a signalling NaN, which arithmetic never produces.
template <typename T> struct ExchangeSentinel;
template <> struct ExchangeSentinel<float> {
    using U = uint32_t;
    static constexpr U bits = 0x7F8ADBADu;
};
template <> struct ExchangeSentinel<double> {
    using U = uint64_t;
    static constexpr U bits = 0x7FF4DEADBEEFCAFEull;
};
template <typename T>
void launch_exchange_fill(T *xtr, size_t n, hipStream_t s);
template <typename T>
void launch_diag_sweep(const DiagSweepParams<T> &sp, int p, hipStream_t s);
template <typename T>
size_t diag_sweep_lds_bytes(const DiagSweepParams<T> &sp, int p);
// LDS of a lane-split sweep workgroup of `rows` row slots over an nx-column mesh with M materials (0: the
// lane-split kernel is not built for (T, p) or does not fit)
template <typename T>
size_t split_sweep_lds_bytes(int p, int rows, int nx, int M);
// Batched sweep over work items {problem, ordinate}; `rep` fixes degree-independent choices (mode,
// material path), probs/items are device arrays, lds >= diag_sweep_lds_bytes of every problem.
template <typename T>
void launch_diag_sweep_batch(const DiagSweepParams<T> &rep, const DiagSweepParams<T> *probs, const int2 *items,
                             int nitems, int max_ny, size_t lds, int p, hipStream_t s);
// tin_x / tin_y of every member of a batch from the halos of its qref (device table of nprob members)
template <typename T>
void launch_diag_inflow_batch(const DiagSweepParams<T> *probs, int nprob, int p, hipStream_t s);
// dir 0: reference psi -> diagonal psi; dir 1: diagonal -> reference (interior cells only)
template <typename TR, typename TD>
void launch_psi_convert(const Geom &g, TR *ref, TD *diag, int dir, hipStream_t s);
// per-cell [i][j][K] -> per-quadrant diagonal [4][K][ncell]
template <typename TI, typename TO>
void launch_cell_to_diag(const Geom &g, const TI *cell, TO *out, int K, hipStream_t s, const int *skip = nullptr);
// inflow face traces of the halo cells of reference-layout psi
template <typename TO>
void launch_inflow(const Geom &g, const double *ref, TO *tin_x, TO *tin_y, hipStream_t s);
// phi[i][j][k] = beta xb + alpha sum_q w_q psi (diagonal psi; xb null: xb = phi); phiQ: [4][N][ncell] scratch
template <typename TI>
void launch_pmv_diag(const Geom &g, const TI *psi, const double *w, double *phiQ, double *phi, double alpha,
                     double beta, hipStream_t s, const int *skip = nullptr, const double *xb = nullptr);
// res = qsrc + sig_s phi/(4 pi) - (L + 1/dt) psi on the diagonal layout (FP64 inputs)
template <typename TO, typename TP = double>
void launch_residual_diag(const Geom &g, const double *qsrc, const double *phi_d, const double *sigs_d,
                          const double *sigt_d, const TP *psi, const double *tin_x, const double *tin_y,
                          const double *xi, const double *eta, double inv_dt, TO *res, hipStream_t s);
// reference psi (interior) <- (double) a + (double) b, a and b diagonal-layout FP32 arrays
void launch_psi_store_sum(const Geom &g, double *ref, const float *a, const float *b, hipStream_t s);
// reference psi -> diagonal psi in FP64 and FP32 from one read
void launch_psi_load_dual(const Geom &g, const double *ref, double *d64, float *d32, hipStream_t s);

// Largest dynamic LDS allocation one workgroup may make on the current device (bytes).
size_t max_lds_per_block();

// ---- launch wrappers (defined in the .hip files) ----
template <typename T>
void launch_sweep_wavefront(const SweepParams<T> &sp, int p, hipStream_t s);
template <typename T>
void launch_material_inverses(const Geom &g, const double *xi, const double *eta, const double *sig_m, int M,
                              double inv_dt, T *out, hipStream_t s);

template <typename TI, typename TO>
void launch_pmv(const Geom &g, const TI *psi, TO *phi, const TO *w, TO alpha, TO beta, int all_cells, hipStream_t s);
// halo=1: every halo cell of a per-cell array (per_cell values each) gets value; halo=0: every interior cell
template <typename T>
void launch_fill_region(const Geom &g, T *a, size_t per_cell, int halo, T value, hipStream_t s);
void launch_accumulate_f2d(const float *in, double *out, size_t n, hipStream_t s);
// xs (optional): tensor of a degree-p sigma (launch_xs_tensor), used instead of the cell means
template <typename T>
void launch_smv(const Geom &g, const T *sigma_s, const T *phi, T *psi, T alpha, T beta, int all_cells,
                hipStream_t s, const T *xs = nullptr);
template <typename T>
void launch_lmv(const Geom &g, const T *xi, const T *eta, const T *sigma_t, const T *x, T *y, T alpha, T beta,
                hipStream_t s, const T *xs = nullptr);
void launch_convert_d2f(const double *in, float *out, size_t n, hipStream_t s);
void launch_convert_f2d(const float *in, double *out, size_t n, hipStream_t s);
// out = (double) in on the interior cells of a per-cell array; halo cells of out are left as they are
void launch_convert_f2d_interior(const Geom &g, size_t per_cell, const float *in, double *out, hipStream_t s);

// y[k] = a*x[k] + b*y[k]
void launch_axpby(size_t n, double a, const double *x, double b, double *y, hipStream_t s);
// y = sum_j a[j] x[j] (j < m <= kMax) over psi-layout arrays in one pass; halo cells take only the terms
// whose bit is set in halo_terms.  kMax covers the largest IDC source: 3 + 12 residual terms (12 nodes,
// plus the left end point for Gauss-Legendre nodes).
struct LinComb {
    static constexpr int kMax = 16;
    int m = 0;
    unsigned halo_terms = 0;
    double a[kMax] = {};
    const double *x[kMax] = {};
    void add(double coef, const double *arr, bool on_halo = false) {
        if (m >= kMax) throw Error(TSW_E_ARG, "linear combination of more than 16 arrays");
        if (on_halo) halo_terms |= 1u << m;
        a[m] = coef;
        x[m++] = arr;
    }
};
void launch_lincomb(const Geom &g, size_t per_cell, const LinComb &L, double *y, hipStream_t s);
// flag[0] <- 1 if a and b differ bitwise anywhere in n doubles, else 0
void launch_differs(size_t n, const double *a, const double *b, int *flag, hipStream_t s);
// out[i] = sum_k X_i[k] * y[k], i < m, X_i = X + i*ld; uses `partials` scratch; result on device
void launch_multi_dot(size_t n, int m, const double *X, size_t ld, const double *y, double *out, double *partials,
                      int nblocks, hipStream_t s);
// y[k] -= sum_i h[i] X_i[k]
void launch_multi_axpy(size_t n, int m, const double *X, size_t ld, const double *h, double *y, hipStream_t s);
// v /= sqrt(*sq) when that is positive; *hn = sqrt(*sq) (sq, hn: device scalars)
void launch_normalize(size_t n, const double *sq, double *v, double *hn, hipStream_t s);
// out = sum over interior coefficients of (a-b)^2
void launch_sqdist_interior(const Geom &g, const double *a, const double *b, double *out, double *partials,
                            int nblocks, hipStream_t s);
int reduction_blocks();

// adjoint path (adjoint.hip)
void launch_reverse_ordinates(const Geom &g, const double *in, double *out, int kmode, const double *w, int qmode,
                              bool w_of_out, hipStream_t s);
void launch_cell_grads(const Geom &g, const double *psi, const double *zeta, const double *phi, double *gst,
                       double *gss, hipStream_t s, double scale = 1.0, bool accumulate = false);
void launch_kt_apply(const Geom &g, const double *y, double *out, const double *ss, const double *w, double alpha,
                     hipStream_t s);
void launch_inflow_grad(const Geom &g, const double *zeta, double *out, const double *xi, const double *eta,
                        hipStream_t s);

}  // namespace tsw
