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
    // optional material table: per-cell material index and element-matrix inverses [M][nq][N][N]
    const int32_t *mat = nullptr;
    const T *ainv = nullptr;
    int M = 0;
};

// Largest dynamic LDS allocation one workgroup may make on the current device (bytes).
size_t max_lds_per_block();

// ---- launch wrappers (defined in the .hip files) ----
template <typename T>
void launch_sweep_wavefront(const SweepParams<T> &sp, int p, hipStream_t s);
template <typename T>
void launch_sweep_staggered(const SweepParams<T> &sp, int p, hipStream_t s);
// bytes of LDS the staggered kernel needs for this problem (0 if it cannot run)
template <typename T>
size_t staggered_lds_bytes(const SweepParams<T> &sp, int p);
template <typename T>
void launch_material_inverses(const Geom &g, const double *xi, const double *eta, const double *sig_m, int M,
                              double inv_dt, T *out, hipStream_t s);

template <typename TI, typename TO>
void launch_pmv(const Geom &g, const TI *psi, TO *phi, const TO *w, TO alpha, TO beta, int all_cells, hipStream_t s);
// halo=1: every halo cell of a per-cell array (per_cell values each) gets value; halo=0: every interior cell
template <typename T>
void launch_fill_region(const Geom &g, T *a, size_t per_cell, int halo, T value, hipStream_t s);
void launch_accumulate_f2d(const float *in, double *out, size_t n, hipStream_t s);
template <typename T>
void launch_smv(const Geom &g, const T *sigma_s, const T *phi, T *psi, T alpha, T beta, int all_cells,
                hipStream_t s);
template <typename T>
void launch_lmv(const Geom &g, const T *xi, const T *eta, const T *sigma_t, const T *x, T *y, T alpha, T beta,
                hipStream_t s);
void launch_convert_d2f(const double *in, float *out, size_t n, hipStream_t s);
void launch_convert_f2d(const float *in, double *out, size_t n, hipStream_t s);

// y[k] = a*x[k] + b*y[k]
void launch_axpby(size_t n, double a, const double *x, double b, double *y, hipStream_t s);
// out[i] = sum_k X_i[k] * y[k], i < m, X_i = X + i*ld; uses `partials` scratch; result on device
void launch_multi_dot(size_t n, int m, const double *X, size_t ld, const double *y, double *out, double *partials,
                      int nblocks, hipStream_t s);
// y[k] -= sum_i h[i] X_i[k]
void launch_multi_axpy(size_t n, int m, const double *X, size_t ld, const double *h, double *y, hipStream_t s);
// out = sum over interior coefficients of (a-b)^2
void launch_sqdist_interior(const Geom &g, const double *a, const double *b, double *out, double *partials,
                            int nblocks, hipStream_t s);
int reduction_blocks();

}  // namespace tsw
