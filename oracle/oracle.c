/*
 * oracle.c -- CPU restatement of the reference's 2D RKDG sweep, Pmv/Smv/Lmv, source iteration and
 * sweep-driven Peierls GMRES (crockatt/idc-transport-solver).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path, never part of it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library.
 *
 * Written from the mathematics the reference implements; each function cites the reference lines
 * whose behaviour (including floating-point operation order where it is cheap to keep) it follows.
 * Layouts and conventions: see oracle.h.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define ORC_PI 3.14159265358979323846

/* (-1)^k as in global.hpp:490. */
static inline double sgn_pow(int k) { return (k % 2) ? -1.0 : 1.0; }

static inline size_t psi_idx(int ny, int nq, int p, int q, int i, int j, int d, int e)
{
    const size_t P = (size_t)(p + 1);
    /* OrdinateFlux.hpp:639-700: e + (p+1)*(d + (p+1)*(q + nq*(j + (ny+2)*i))) */
    return (size_t)e + P * ((size_t)d + P * ((size_t)q + (size_t)nq * ((size_t)j + (size_t)(ny + 2) * (size_t)i)));
}

static inline size_t phi_idx(int ny, int p, int i, int j, int d, int e)
{
    const size_t P = (size_t)(p + 1);
    return (size_t)e + P * ((size_t)d + P * ((size_t)j + (size_t)(ny + 2) * (size_t)i));
}

/* ---------------------------------------------------------------------------------------------- */
/* Cross sections of DG degree p (init.cpp:741 with piecewise_cross false).  While the degree set   */
/* by orc_set_xs_degree is > 0, every sigma argument below is a [i][j][u][v] coefficient array and    */
/* enters through the tensor of Legendre triple products (CrossSection::ComputeTensor,              */
/* CrossSection.cpp:165-189).                                                                        */
/* ---------------------------------------------------------------------------------------------- */
static int g_xs_deg = 0;

void orc_set_xs_degree(int degree) { g_xs_deg = degree; }

/* Legendre P_n(x) by the three-term recurrence, and Gauss-Legendre nodes/weights by Newton from the
 * Chebyshev-like guess (own quadrature; the reference reads tabulated rules, Tabulated.cpp:29). */
static double legendre(int n, double x)
{
    double p0 = 1.0, p1 = x;
    if (n == 0) return 1.0;
    for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
    }
    return p1;
}

static void gauss_legendre(int m, double *x, double *w)
{
    for (int k = 0; k < m; ++k) {
        double t = -cos(ORC_PI * (k + 0.75) / (m + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            const double pm = legendre(m, t), pm1 = legendre(m - 1, t);
            dp = m * (pm1 - t * pm) / (1.0 - t * t);
            const double dt = pm / dp;
            t -= dt;
            if (fabs(dt) < 1e-16) break;
        }
        const double pm = legendre(m, t), pm1 = legendre(m - 1, t);
        dp = m * (pm1 - t * pm) / (1.0 - t * t);
        x[k] = t;
        w[k] = 2.0 / ((1.0 - t * t) * dp * dp);
    }
}

/* TPI[a + P(b + P c)] = int_{-1}^{1} P_a P_b P_c, by max(2p, 2) Gauss points, entries below 5e-16 set
 * to zero (ComputeTPI, global.cpp:67-107; ITPI global.hpp:644). */
static void tpi_table(int p, double *tpi)
{
    const int P = p + 1, m = 2 * p > 2 ? 2 * p : 2;
    double xg[16], wg[16];
    gauss_legendre(m, xg, wg);
    for (int a = 0; a <= p; ++a)
        for (int b = 0; b <= p; ++b)
            for (int c = 0; c <= p; ++c) {
                double v = 0.0;
                for (int i = 0; i < m; ++i) v += wg[i] * legendre(a, xg[i]) * legendre(b, xg[i]) * legendre(c, xg[i]);
                if (fabs(v) < 5e-16) v = 0.0;
                tpi[a + P * (b + P * c)] = v;
            }
}

/* Tensor of a degree-p cross section, t[cell*N*N + (d*P+e)*N + a*P+b] = sigma(i,j,a,b,d,e) =
 * sum_{u,v} sigma(i,j,u,v) TPI(a,d,u) TPI(b,e,v) (CrossSection.cpp:179-188, index map TensorIndex
 * CrossSection.hpp:397-478).  Caller frees. */
static double *xs_tensor(int nx, int ny, int p, const double *sig)
{
    const int P = p + 1, N = P * P;
    const size_t ncell = (size_t)(nx + 2) * (ny + 2);
    double tpi[64];
    tpi_table(p, tpi);
    double *t = (double *)calloc(ncell * N * N, sizeof(double));
    for (size_t c = 0; c < ncell; ++c)
        for (int a = 0; a <= p; ++a)
            for (int b = 0; b <= p; ++b)
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e) {
                        double *out = &t[c * N * N + (size_t)(d * P + e) * N + a * P + b];
                        for (int u = 0; u <= p; ++u)
                            for (int v = 0; v <= p; ++v)
                                *out += sig[c * N + u * P + v] * tpi[a + P * (d + P * u)] * tpi[b + P * (e + P * v)];
                    }
    return t;
}

/* sigma(i,j,d,e,u,v) of the tensor above (CrossSection::operator(), CrossSection.hpp:481-535) */
static inline double xs_at(const double *t, int ny, int p, int i, int j, int d, int e, int u, int v)
{
    const int P = p + 1, N = P * P;
    const size_t c = (size_t)j + (size_t)(ny + 2) * i;
    return t[c * N * N + (size_t)(u * P + v) * N + d * P + e];
}

void orc_xs_tensor(int nx, int ny, int p, const double *sigma, double *tensor)
{
    const int N = (p + 1) * (p + 1);
    double *t = xs_tensor(nx, ny, p, sigma);
    memcpy(tensor, t, sizeof(double) * (size_t)(nx + 2) * (ny + 2) * N * N);
    free(t);
}

/* ---------------------------------------------------------------------------------------------- */
/* Angular matrices.  SweepLinearSolveSpec.cpp:98-158 (loop over quadrants, lanes, d, e).          */
/* ---------------------------------------------------------------------------------------------- */
void orc_aq(int p, int nq, const double *xi, const double *eta, double dx, double dy, double *Aq)
{
    const int P = p + 1, N = P * P;
    memset(Aq, 0, sizeof(double) * (size_t)nq * N * N);
    for (int q = 0; q < nq; ++q) {
        double *A = Aq + (size_t)q * N * N;
        const double xi_coeff = dy * xi[q];   /* dx(1) * xi, :117 */
        const double eta_coeff = dx * eta[q]; /* dx(0) * eta, :118 */
#define AQ(d, e, a, b) A[((e) + P * (d)) + N * ((b) + P * (a))]
        for (int d = 0; d <= p; ++d) {
            for (int e = 0; e <= p; ++e) {
                /* xi direction, :124-138 */
                for (int a = d - 1; a >= 0; a -= 2) AQ(d, e, a, e) -= 2.0 * xi_coeff * (2 * d + 1);
                if (xi[q] < 0) {
                    for (int a = 0; a <= p; ++a) AQ(d, e, a, e) += sgn_pow(d + a + 1) * xi_coeff * (2 * d + 1);
                } else {
                    for (int a = 0; a <= p; ++a) AQ(d, e, a, e) += xi_coeff * (2 * d + 1);
                }
                /* eta direction, :141-155 */
                for (int b = e - 1; b >= 0; b -= 2) AQ(d, e, d, b) -= 2.0 * eta_coeff * (2 * e + 1);
                if (eta[q] < 0) {
                    for (int b = 0; b <= p; ++b) AQ(d, e, d, b) += sgn_pow(e + b + 1) * eta_coeff * (2 * e + 1);
                } else {
                    for (int b = 0; b <= p; ++b) AQ(d, e, d, b) += eta_coeff * (2 * e + 1);
                }
            }
        }
#undef AQ
    }
}

/* ---------------------------------------------------------------------------------------------- */
/* Unpivoted GE.  SIMD_LinearAlgebra.hpp:78-115, one lane.                                        */
/* ---------------------------------------------------------------------------------------------- */
void orc_ge_solve(int N, double *A, double *b)
{
#define M(i, j) A[(i) + N * (j)]
    for (int j = 0; j < N; ++j) {
        for (int i = j + 1; i < N; ++i) {
            M(i, j) /= -M(j, j);
            b[i] += M(i, j) * b[j];
        }
        for (int k = j + 1; k < N; ++k)
            for (int i = j + 1; i < N; ++i) M(i, k) += M(i, j) * M(j, k);
    }
    for (int i = N - 1; i >= 0; --i) {
        b[i] /= M(i, i);
        for (int j = 0; j < i; ++j) b[j] -= M(j, i) * b[i];
    }
#undef M
}

/* One cell, one ordinate: CalcB (SweepLinearSolveSpec.cpp:473-598), CalcA (:284-298),
 * SolveSystem, StoreResult (SweepCellSolve.cpp:86-99). */
static void cell_solve(int nx, int ny, int p, int nq, const double *xi, const double *eta, double dx,
                       double dy, const double *Aq, const double *sigma_t, const double *tensor, double inv_dt,
                       int q, int i, int j, double *psi, double *A, double *B, double *W0, double *W1)
{
    (void)nx;
    const int P = p + 1, N = P * P;
    const int xneg = xi[q] <= 0.0, yneg = eta[q] <= 0.0; /* :504-505 */
    const int xoff = xneg ? +1 : -1, yoff = yneg ? +1 : -1;

    for (int d = 0; d <= p; ++d)
        for (int e = 0; e <= p; ++e) {
            W0[e + P * d] = psi[psi_idx(ny, nq, p, q, i + xoff, j, d, e)] * (dy * xi[q]);
            W1[e + P * d] = psi[psi_idx(ny, nq, p, q, i, j + yoff, d, e)] * (dx * eta[q]);
        }
    const double dx_coeff = dx * dy;
    for (int d = 0; d <= p; ++d)
        for (int e = 0; e <= p; ++e) B[e + P * d] = dx_coeff * psi[psi_idx(ny, nq, p, q, i, j, d, e)];
    for (int d = 0; d <= p; ++d)
        for (int e = 0; e <= p; ++e)
            for (int a = 0; a <= p; ++a) {
                const double c = xneg ? sgn_pow(a + 1) * (2 * d + 1) : sgn_pow(d) * (2 * d + 1);
                B[e + P * d] += W0[e + P * a] * c;
            }
    for (int d = 0; d <= p; ++d)
        for (int e = 0; e <= p; ++e)
            for (int b = 0; b <= p; ++b) {
                const double c = yneg ? sgn_pow(b + 1) * (2 * e + 1) : sgn_pow(e) * (2 * e + 1);
                B[e + P * d] += W1[b + P * d] * c;
            }

    memcpy(A, Aq + (size_t)q * N * N, sizeof(double) * N * N);
    if (!tensor) {
        const double sig = sigma_t[(size_t)j + (size_t)(ny + 2) * i];
        const double coeff = (sig + inv_dt) * dx * dy;
        for (int r = 0; r < N; ++r) A[r + N * r] += coeff;
    } else { /* general sigma, :300-321 (1/dt = 0 for steady state) */
        const double dt_coeff = dx * dy * inv_dt, dx_coeff = dx * dy / 4.0;
        for (int d = 0; d <= p; ++d)
            for (int e = 0; e <= p; ++e) {
                A[(e + P * d) * (N + 1)] += dt_coeff;
                for (int u = 0; u <= p; ++u)
                    for (int v = 0; v <= p; ++v)
                        A[(e + P * d) + N * (v + P * u)] += dx_coeff * (2 * d + 1) * (2 * e + 1) *
                                                            xs_at(tensor, ny, p, i, j, d, e, u, v);
            }
    }

    orc_ge_solve(N, A, B);

    for (int d = 0; d <= p; ++d)
        for (int e = 0; e <= p; ++e) psi[psi_idx(ny, nq, p, q, i, j, d, e)] = B[e + P * d];
}

/* ---------------------------------------------------------------------------------------------- */
/* Wavefront traversal.  SweepPatternUniCartWavefront.cpp:253-296 (all quadrants at once),        */
/* diagonal length SweepPatternUniCart.cpp:238-254, quadrant signs OrdinateSet.cpp:658-700.       */
/* ---------------------------------------------------------------------------------------------- */
void orc_sweep(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
               double dx, double dy, const double *sigma_t, double dt, double *psi)
{
    const int P = p + 1, N = P * P;
    double *Aq = (double *)malloc(sizeof(double) * (size_t)nq * N * N);
    double *A = (double *)malloc(sizeof(double) * N * N);
    double *B = (double *)malloc(sizeof(double) * N);
    double *W0 = (double *)malloc(sizeof(double) * N);
    double *W1 = (double *)malloc(sizeof(double) * N);
    orc_aq(p, nq, xi, eta, dx, dy, Aq);
    const double inv_dt = (dt > 0.0 && !isinf(dt)) ? 1.0 / dt : 0.0;
    double *tensor = g_xs_deg > 0 ? xs_tensor(nx, ny, p, sigma_t) : NULL;

    const int num_diags = nx + ny - 1;
    for (int diag = 1; diag <= num_diags; ++diag) {
        /* ComputeDiagLength: cells (i,j), 1<=i<=nx, 1<=j<=ny, i+j-1 == diag */
        const int i_hi = diag < nx ? diag : nx;
        const int i_lo = diag - ny + 1 > 1 ? diag - ny + 1 : 1;
        const int diag_length = i_hi - i_lo + 1;
        for (int cell = 1; cell <= diag_length; ++cell) {
            for (int quad = 0; quad < 4; ++quad) {
                for (int q = quad_off[quad]; q < quad_off[quad + 1]; ++q) {
                    int i = i_hi - cell + 1;
                    int j = diag - i + 1;
                    if (quad >= 2) i = nx - i + 1;      /* xi < 0 */
                    if (quad & 1) j = ny - j + 1;       /* eta < 0 */
                    cell_solve(nx, ny, p, nq, xi, eta, dx, dy, Aq, sigma_t, tensor, inv_dt, q, i, j, psi, A, B,
                               W0, W1);
                }
            }
        }
    }
    free(Aq); free(A); free(B); free(W0); free(W1); free(tensor);
}

/* ---------------------------------------------------------------------------------------------- */
/* Reflecting boundaries.                                                                         */
/* ---------------------------------------------------------------------------------------------- */
/* ReflectMap (OrdinateSet.cpp:1093-1140): map[dim * nq + q] = the ordinate k whose coordinate dim is
 * the negative of q's and whose other coordinates equal q's (to 1e-14), searched from k = 0.  Only xi
 * and eta are known here, so a set in which two ordinates share (xi, eta) (an unreduced set's +-mu
 * pairs) is refused: returns -1 (also when an ordinate has no mirror), else 0. */
int orc_reflect_map(int nq, const double *xi, const double *eta, int *map)
{
    const double TOL = 1.0e-14;
    for (int q = 0; q < nq; ++q)
        for (int k = 0; k < q; ++k)
            if (fabs(xi[q] - xi[k]) < TOL && fabs(eta[q] - eta[k]) < TOL) return -1;
    for (int dim = 0; dim < 2; ++dim)
        for (int q = 0; q < nq; ++q) {
            map[dim * nq + q] = -1;
            for (int k = 0; k < nq; ++k) {
                const double cx = dim == 0 ? -xi[k] : xi[k], cy = dim == 1 ? -eta[k] : eta[k];
                if (fabs(xi[q] - cx) < TOL && fabs(eta[q] - cy) < TOL) {
                    map[dim * nq + q] = k;
                    break;
                }
            }
            if (map[dim * nq + q] < 0) return -1;
        }
    return 0;
}

/* OrdinateFlux::ReflectBoundaries (RKDG/OrdinateFlux.cpp:1181-1420): the ghost cells of every edge in
 * `mask` (1 x-min, 2 x-max, 4 y-min, 8 y-max, BoundaryEdge in global.hpp:93-120) take the mirrored
 * ordinate's values of the adjacent interior cell, the coefficients odd in the reflected direction
 * negated.  As there, x-min covers quadrants 0 and 1, x-max 2 and 3, y-min 0 and 2, and y-max all four. */
void orc_reflect_boundaries(int nx, int ny, int p, int nq, const int *quad_off, const int *map, int mask, double *psi)
{
    static const int quads[4][4] = {{0, 1, -1, -1}, {2, 3, -1, -1}, {0, 2, -1, -1}, {0, 1, 2, 3}};
    for (int edge = 0; edge < 4; ++edge) {
        if (!(mask & (1 << edge))) continue;
        const int dim = edge / 2;
        for (int k = 0; k < 4 && quads[edge][k] >= 0; ++k) {
            const int quad = quads[edge][k];
            for (int q = quad_off[quad]; q < quad_off[quad + 1]; ++q) {
                const int m = map[dim * nq + q];
                const int len = dim == 0 ? ny : nx;
                for (int t = 1; t <= len; ++t)
                    for (int d = 0; d <= p; ++d)
                        for (int e = 0; e <= p; ++e) {
                            int gi, gj, si, sj;
                            if (dim == 0) {
                                gi = edge == 0 ? 0 : nx + 1;
                                si = edge == 0 ? 1 : nx;
                                gj = sj = t;
                            } else {
                                gj = edge == 2 ? 0 : ny + 1;
                                sj = edge == 2 ? 1 : ny;
                                gi = si = t;
                            }
                            psi[psi_idx(ny, nq, p, q, gi, gj, d, e)] =
                                sgn_pow(dim == 0 ? d : e) * psi[psi_idx(ny, nq, p, m, si, sj, d, e)];
                        }
            }
        }
    }
}

/* DensityFunction::SynchronizeHalos across periodic boundaries (Abstract/DensityFunction.cpp:927-956):
 * y-halos of columns 1..nx first, then x-halos of rows 1..ny, every value of the cell (corners untouched). */
void orc_periodic_halos(int nx, int ny, int p, int nq, double *psi)
{
    const size_t per = (size_t)nq * (p + 1) * (p + 1);
    for (int i = 1; i <= nx; ++i)
        for (size_t l = 0; l < per; ++l) {
            psi[((size_t)i * (ny + 2) + 0) * per + l] = psi[((size_t)i * (ny + 2) + ny) * per + l];
            psi[((size_t)i * (ny + 2) + ny + 1) * per + l] = psi[((size_t)i * (ny + 2) + 1) * per + l];
        }
    for (int j = 1; j <= ny; ++j)
        for (size_t l = 0; l < per; ++l) {
            psi[((size_t)0 * (ny + 2) + j) * per + l] = psi[((size_t)nx * (ny + 2) + j) * per + l];
            psi[((size_t)(nx + 1) * (ny + 2) + j) * per + l] = psi[((size_t)1 * (ny + 2) + j) * per + l];
        }
}

/* SweepOperator::Apply with reflecting edges (SweepPatternUniCartWavefront.cpp:180-245): the quadrants
 * one after the other in the order 3, 2, 1, 0 (SweepPatternUniCart.cpp:33-40), each a wavefront over
 * its diagonals, and ReflectBoundaries on the in-place result after each quadrant.  A quadrant swept
 * before the mirror of its reflected inflow takes that inflow from the mirror's values at that moment
 * (the source, or the halo the caller gave), as the reference does. */
void orc_sweep_reflect(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
                       double dx, double dy, const double *sigma_t, double dt, int mask, const int *map, double *psi)
{
    const int P = p + 1, N = P * P;
    double *Aq = (double *)malloc(sizeof(double) * (size_t)nq * N * N);
    double *A = (double *)malloc(sizeof(double) * N * N);
    double *B = (double *)malloc(sizeof(double) * N);
    double *W0 = (double *)malloc(sizeof(double) * N);
    double *W1 = (double *)malloc(sizeof(double) * N);
    orc_aq(p, nq, xi, eta, dx, dy, Aq);
    const double inv_dt = (dt > 0.0 && !isinf(dt)) ? 1.0 / dt : 0.0;
    double *tensor = g_xs_deg > 0 ? xs_tensor(nx, ny, p, sigma_t) : NULL;
    static const int order[4] = {3, 2, 1, 0};
    for (int k = 0; k < 4; ++k) {
        const int quad = order[k];
        for (int diag = 1; diag <= nx + ny - 1; ++diag) {
            const int i_hi = diag < nx ? diag : nx;
            const int i_lo = diag - ny + 1 > 1 ? diag - ny + 1 : 1;
            for (int cell = 1; cell <= i_hi - i_lo + 1; ++cell)
                for (int q = quad_off[quad]; q < quad_off[quad + 1]; ++q) {
                    int i = i_hi - cell + 1;
                    int j = diag - i + 1;
                    if (quad >= 2) i = nx - i + 1;
                    if (quad & 1) j = ny - j + 1;
                    cell_solve(nx, ny, p, nq, xi, eta, dx, dy, Aq, sigma_t, tensor, inv_dt, q, i, j, psi, A, B, W0, W1);
                }
        }
        if (mask) orc_reflect_boundaries(nx, ny, p, nq, quad_off, map, mask, psi);
    }
    free(Aq); free(A); free(B); free(W0); free(W1); free(tensor);
}

/* ---------------------------------------------------------------------------------------------- */
/* The same wavefront with the cells x ordinates of each diagonal shared among OpenMP threads, one   */
/* barrier per diagonal, as SweepPatternUniCartWavefront::Sweep distributes them with its `omp for`  */
/* (SweepPatternUniCartWavefront.cpp:146,253-296).  Per-cell arithmetic is that of orc_sweep, so the */
/* result is bitwise identical; this is the CPU baseline's multi-core form.                          */
/* ---------------------------------------------------------------------------------------------- */
void orc_sweep_omp(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
                   double dx, double dy, const double *sigma_t, double dt, double *psi, int nthreads)
{
    const int P = p + 1, N = P * P;
    double *Aq = (double *)malloc(sizeof(double) * (size_t)nq * N * N);
    orc_aq(p, nq, xi, eta, dx, dy, Aq);
    const double inv_dt = (dt > 0.0 && !isinf(dt)) ? 1.0 / dt : 0.0;
    double *tensor = g_xs_deg > 0 ? xs_tensor(nx, ny, p, sigma_t) : NULL;
    /* quadrant of each ordinate */
    int *quad = (int *)malloc(sizeof(int) * (size_t)nq);
    for (int k = 0; k < 4; ++k)
        for (int q = quad_off[k]; q < quad_off[k + 1]; ++q) quad[q] = k;
    const int num_diags = nx + ny - 1;
#pragma omp parallel num_threads(nthreads > 0 ? nthreads : 1)
    {
        double A[256], B[16], W0[16], W1[16];
        for (int diag = 1; diag <= num_diags; ++diag) {
            const int i_hi = diag < nx ? diag : nx;
            const int i_lo = diag - ny + 1 > 1 ? diag - ny + 1 : 1;
            const int diag_length = i_hi - i_lo + 1;
#pragma omp for schedule(static)
            for (long long t = 0; t < (long long)diag_length * nq; ++t) {
                const int cell = (int)(t / nq) + 1, q = (int)(t % nq);
                int i = i_hi - cell + 1;
                int j = diag - i + 1;
                if (quad[q] >= 2) i = nx - i + 1;
                if (quad[q] & 1) j = ny - j + 1;
                cell_solve(nx, ny, p, nq, xi, eta, dx, dy, Aq, sigma_t, tensor, inv_dt, q, i, j, psi, A, B, W0,
                           W1);
            }   /* implicit barrier: the diagonal is complete */
        }
    }
    free(Aq);
    free(quad);
    free(tensor);
}

/* ---------------------------------------------------------------------------------------------- */
/* Transposed sweep.  The reference has no adjoint; this is the transpose of the linear system     */
/* that orc_sweep solves, written out directly.  With row (d,e) of cell c's equation                */
/*   A_c psi_c - Cx psi_xup - Cy psi_yup = dx dy src_c,                                            */
/*   Cx[(d,e),(a,e)] = dy xi (2d+1) kx(d,a),  kx = (-1)^(a+1) if xi <= 0 else (-1)^d,             */
/*   Cy[(d,e),(d,b)] = dx eta (2e+1) ky(e,b), ky = (-1)^(b+1) if eta <= 0 else (-1)^e             */
/* (CalcB, SweepLinearSolveSpec.cpp:545-598), the transpose is solved cell by cell in reverse      */
/* sweep order:  A_c^T mu_c = g_c + Cx^T mu_xdown + Cy^T mu_ydown.                                  */
/* ---------------------------------------------------------------------------------------------- */
static inline double kx_coef(int neg, int d, int a) { return neg ? sgn_pow(a + 1) : sgn_pow(d); }

void orc_sweep_adjoint(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
                       double dx, double dy, const double *sigma_t, double dt, const double *g, double *out)
{
    (void)quad_off;
    const int P = p + 1, N = P * P;
    const size_t npsi = (size_t)(nx + 2) * (ny + 2) * nq * N;
    double *Aq = (double *)malloc(sizeof(double) * (size_t)nq * N * N);
    double *AT = (double *)malloc(sizeof(double) * N * N);
    double *B = (double *)malloc(sizeof(double) * N);
    double *mu = (double *)calloc(npsi, sizeof(double));
    orc_aq(p, nq, xi, eta, dx, dy, Aq);
    const double inv_dt = (dt > 0.0 && !isinf(dt)) ? 1.0 / dt : 0.0;
    memset(out, 0, sizeof(double) * npsi);
    for (int q = 0; q < nq; ++q) {
        const int xneg = xi[q] <= 0.0, yneg = eta[q] <= 0.0;
        const double cx = dy * xi[q], cy = dx * eta[q];
        /* forward order visits i upwind-first; walk it backwards */
        for (int ii = nx; ii >= 1; --ii)
            for (int jj = ny; jj >= 1; --jj) {
                const int i = xneg ? nx - ii + 1 : ii, j = yneg ? ny - jj + 1 : jj;
                const int idn = xneg ? i - 1 : i + 1, jdn = yneg ? j - 1 : j + 1; /* forward-downwind */
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e) {
                        double v = g[psi_idx(ny, nq, p, q, i, j, d, e)];
                        if (idn >= 1 && idn <= nx)
                            for (int dd = 0; dd <= p; ++dd)
                                v += cx * (2 * dd + 1) * kx_coef(xneg, dd, d) * mu[psi_idx(ny, nq, p, q, idn, j, dd, e)];
                        if (jdn >= 1 && jdn <= ny)
                            for (int ee = 0; ee <= p; ++ee)
                                v += cy * (2 * ee + 1) * kx_coef(yneg, ee, e) * mu[psi_idx(ny, nq, p, q, i, jdn, d, ee)];
                        B[e + P * d] = v;
                    }
                const double coeff = (sigma_t[(size_t)j + (size_t)(ny + 2) * i] + inv_dt) * dx * dy;
                const double *A = Aq + (size_t)q * N * N;
                for (int r = 0; r < N; ++r)
                    for (int c = 0; c < N; ++c) AT[r + N * c] = A[c + N * r] + (r == c ? coeff : 0.0);
                orc_ge_solve(N, AT, B);
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e) {
                        mu[psi_idx(ny, nq, p, q, i, j, d, e)] = B[e + P * d];
                        out[psi_idx(ny, nq, p, q, i, j, d, e)] = dx * dy * B[e + P * d];
                    }
            }
        /* gradient with respect to the inflow halo cells: C^T mu over the inflow faces */
        const int ih = xneg ? nx + 1 : 0, ib = xneg ? nx : 1;
        for (int j = 1; j <= ny; ++j)
            for (int a = 0; a <= p; ++a)
                for (int e = 0; e <= p; ++e) {
                    double v = 0.0;
                    for (int d = 0; d <= p; ++d)
                        v += cx * (2 * d + 1) * kx_coef(xneg, d, a) * mu[psi_idx(ny, nq, p, q, ib, j, d, e)];
                    out[psi_idx(ny, nq, p, q, ih, j, a, e)] = v;
                }
        const int jh = yneg ? ny + 1 : 0, jb = yneg ? ny : 1;
        for (int i = 1; i <= nx; ++i)
            for (int d = 0; d <= p; ++d)
                for (int b = 0; b <= p; ++b) {
                    double v = 0.0;
                    for (int e = 0; e <= p; ++e)
                        v += cy * (2 * e + 1) * kx_coef(yneg, e, b) * mu[psi_idx(ny, nq, p, q, i, jb, d, e)];
                    out[psi_idx(ny, nq, p, q, i, jh, d, b)] = v;
                }
    }
    free(Aq); free(AT); free(B); free(mu);
}

/* ---------------------------------------------------------------------------------------------- */
/* Moment contraction and scattering source.  TransportOperator.cpp:140-158 and :355-392.         */
/* ---------------------------------------------------------------------------------------------- */
void orc_pmv(int nx, int ny, int p, int nq, const double *w, double alpha, double beta, const double *psi,
             double *phi, int all_cells)
{
    const int o = all_cells ? 1 : 0;
    for (int i = 1 - o; i <= nx + o; ++i)
        for (int j = 1 - o; j <= ny + o; ++j) {
            for (int d = 0; d <= p; ++d)
                for (int e = 0; e <= p; ++e) phi[phi_idx(ny, p, i, j, d, e)] *= beta;
            for (int q = 0; q < nq; ++q)
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e)
                        phi[phi_idx(ny, p, i, j, d, e)] += alpha * w[q] * psi[psi_idx(ny, nq, p, q, i, j, d, e)];
        }
}

void orc_smv(int nx, int ny, int p, int nq, double alpha, double beta, const double *sigma_s,
             const double *phi, double *psi, int all_cells)
{
    const int o = all_cells ? 1 : 0;
    const int P = p + 1;
    double val[16];
    double *tensor = g_xs_deg > 0 ? xs_tensor(nx, ny, p, sigma_s) : NULL;
    for (int i = 1 - o; i <= nx + o; ++i)
        for (int j = 1 - o; j <= ny + o; ++j) {
            if (!tensor) {
                const double s = sigma_s[(size_t)j + (size_t)(ny + 2) * i];
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e)
                        val[e + P * d] = alpha * (s * phi[phi_idx(ny, p, i, j, d, e)] / (4.0 * ORC_PI));
            } else { /* general sigma, :375-381 */
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e) {
                        double v = 0.0;
                        for (int u = 0; u <= p; ++u)
                            for (int w = 0; w <= p; ++w)
                                v += xs_at(tensor, ny, p, i, j, d, e, u, w) * phi[phi_idx(ny, p, i, j, u, w)];
                        v *= (2 * d + 1) * (2 * e + 1) / (16.0 * ORC_PI);
                        val[e + P * d] = v * alpha;
                    }
            }
            for (int q = 0; q < nq; ++q)
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e) {
                        double *y = &psi[psi_idx(ny, nq, p, q, i, j, d, e)];
                        *y = beta * *y + val[e + P * d];
                    }
        }
    free(tensor);
}

/* ---------------------------------------------------------------------------------------------- */
/* Explicit transport operator.  TransportOperator.cpp:1576-1658 (2D, unblocked).                 */
/* ---------------------------------------------------------------------------------------------- */
void orc_lmv(int nx, int ny, int p, int nq, const double *xi, const double *eta, double dx, double dy,
             double alpha, double beta, const double *sigma_t, const double *x, double *y)
{
#define X(q, i, j, d, e) x[psi_idx(ny, nq, p, q, i, j, d, e)]
#define Y(q, i, j, d, e) y[psi_idx(ny, nq, p, q, i, j, d, e)]
    double *tensor = g_xs_deg > 0 ? xs_tensor(nx, ny, p, sigma_t) : NULL;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            for (int q = 0; q < nq; ++q) {
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e) Y(q, i, j, d, e) *= beta;
                if (!tensor) {
                    const double sig = sigma_t[(size_t)j + (size_t)(ny + 2) * i];
                    for (int d = 0; d <= p; ++d)
                        for (int e = 0; e <= p; ++e) Y(q, i, j, d, e) += alpha * sig * X(q, i, j, d, e);
                } else { /* general sigma, :1587-1596 */
                    for (int d = 0; d <= p; ++d)
                        for (int e = 0; e <= p; ++e)
                            for (int u = 0; u <= p; ++u)
                                for (int v = 0; v <= p; ++v)
                                    Y(q, i, j, d, e) += alpha * (2 * d + 1) * (2 * e + 1) *
                                                        xs_at(tensor, ny, p, i, j, d, e, u, v) * X(q, i, j, u, v) / 4.0;
                }

                double coeff = alpha * xi[q] / dx;
                if (xi[q] < 0) {
                    for (int d = 0; d <= p; ++d)
                        for (int a = 0; a <= p; ++a)
                            for (int e = 0; e <= p; ++e)
                                Y(q, i, j, d, e) += coeff * (2 * d + 1) *
                                                    (sgn_pow(a) * X(q, i + 1, j, a, e) + sgn_pow(d + a + 1) * X(q, i, j, a, e));
                } else {
                    for (int d = 0; d <= p; ++d)
                        for (int a = 0; a <= p; ++a)
                            for (int e = 0; e <= p; ++e)
                                Y(q, i, j, d, e) += coeff * (2 * d + 1) * (X(q, i, j, a, e) + sgn_pow(d + 1) * X(q, i - 1, j, a, e));
                }
                coeff *= -2.0;
                for (int d = 0; d <= p; ++d)
                    for (int a = d - 1; a >= 0; a -= 2)
                        for (int e = 0; e <= p; ++e) Y(q, i, j, d, e) += coeff * (2 * d + 1) * X(q, i, j, a, e);

                coeff = alpha * eta[q] / dy;
                if (eta[q] < 0) {
                    for (int d = 0; d <= p; ++d)
                        for (int e = 0; e <= p; ++e)
                            for (int b = 0; b <= p; ++b)
                                Y(q, i, j, d, e) += coeff * (2 * e + 1) *
                                                    (sgn_pow(b) * X(q, i, j + 1, d, b) + sgn_pow(e + b + 1) * X(q, i, j, d, b));
                } else {
                    for (int d = 0; d <= p; ++d)
                        for (int e = 0; e <= p; ++e)
                            for (int b = 0; b <= p; ++b)
                                Y(q, i, j, d, e) += coeff * (2 * e + 1) * (X(q, i, j, d, b) + sgn_pow(e + 1) * X(q, i, j - 1, d, b));
                }
                coeff *= -2.0;
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e)
                        for (int b = e - 1; b >= 0; b -= 2) Y(q, i, j, d, e) += coeff * (2 * e + 1) * X(q, i, j, d, b);
            }
    free(tensor);
#undef X
#undef Y
}

double orc_l2dist(int nx, int ny, int p, const double *a, const double *b)
{
    const int N = (p + 1) * (p + 1);
    double dist = 0.0;
    for (int i = 1; i <= nx; ++i)
        for (int j = 1; j <= ny; ++j)
            for (int l = 0; l < N; ++l) {
                const size_t k = (size_t)l + (size_t)N * ((size_t)j + (size_t)(ny + 2) * i);
                const double t = a[k] - b[k];
                dist += t * t;
            }
    return sqrt(dist);
}

/* ---------------------------------------------------------------------------------------------- */
/* Source iteration.  SISolver.cpp:206-249.                                                       */
/* ---------------------------------------------------------------------------------------------- */
int orc_si_solve(int nx, int ny, int p, int nq, const double *xi, const double *eta, const double *w,
                 const int *quad_off, double dx, double dy, const double *sigma_t, const double *sigma_s,
                 double dt, double abs_tol, int max_its, double *psi, double *source, int *iters,
                 double *err)
{
    const int N = (p + 1) * (p + 1);
    const size_t npsi = (size_t)(nx + 2) * (ny + 2) * nq * N;
    const size_t nphi = (size_t)(nx + 2) * (ny + 2) * N;
    double *phi_a = (double *)calloc(nphi, sizeof(double));
    double *phi_b = (double *)calloc(nphi, sizeof(double));
    double *cur = phi_a, *last = phi_b;
    int iter = 0;
    double error = 1.79769313486231570e+308;

    orc_pmv(nx, ny, p, nq, w, 1.0, 0.0, psi, cur, 0);
    while (iter < max_its && error > abs_tol) {
        ++iter;
        double *t = cur; cur = last; last = t;
        memcpy(psi, source, sizeof(double) * npsi);                        /* :226 */
        orc_smv(nx, ny, p, nq, 1.0, 1.0, sigma_s, last, psi, 0);           /* :229 */
        orc_sweep(nx, ny, p, nq, xi, eta, quad_off, dx, dy, sigma_t, dt, psi); /* :232 */
        orc_pmv(nx, ny, p, nq, w, 1.0, 0.0, psi, cur, 0);                 /* :236 */
        error = orc_l2dist(nx, ny, p, cur, last);                          /* :239 */
    }
    memcpy(source, psi, sizeof(double) * npsi);                           /* :249 */
    free(phi_a); free(phi_b);
    if (iters) *iters = iter;
    if (err) *err = error;
    return 0;
}

/* ---------------------------------------------------------------------------------------------- */
/* Peierls-form GMRES(m).  PETScPeierlsSolver.cpp:458-505 and IPLS :548-555.                      */
/* ---------------------------------------------------------------------------------------------- */
typedef struct {
    int nx, ny, p, nq;
    const double *xi, *eta, *w;
    const int *quad_off;
    double dx, dy, dt;
    const double *sigma_t, *sigma_s;
    double *psi, *phi;
} peierls_ctx;

static void pack_interior(const peierls_ctx *c, const double *phi, double *v)
{
    const int N = (c->p + 1) * (c->p + 1);
    size_t k = 0;
    for (int i = 1; i <= c->nx; ++i)
        for (int j = 1; j <= c->ny; ++j)
            for (int l = 0; l < N; ++l) v[k++] = phi[(size_t)l + (size_t)N * ((size_t)j + (size_t)(c->ny + 2) * i)];
}

static void unpack_interior(const peierls_ctx *c, const double *v, double *phi)
{
    const int N = (c->p + 1) * (c->p + 1);
    size_t k = 0;
    for (int i = 1; i <= c->nx; ++i)
        for (int j = 1; j <= c->ny; ++j)
            for (int l = 0; l < N; ++l) phi[(size_t)l + (size_t)N * ((size_t)j + (size_t)(c->ny + 2) * i)] = v[k++];
}

/* y = (I - P L^{-1} S) x, zero inflow. */
static void peierls_apply(peierls_ctx *c, const double *x, double *y)
{
    const int N = (c->p + 1) * (c->p + 1);
    const size_t npsi = (size_t)(c->nx + 2) * (c->ny + 2) * c->nq * N;
    const size_t nphi = (size_t)(c->nx + 2) * (c->ny + 2) * N;
    memset(c->phi, 0, sizeof(double) * nphi);
    unpack_interior(c, x, c->phi);
    memset(c->psi, 0, sizeof(double) * npsi);
    orc_smv(c->nx, c->ny, c->p, c->nq, 1.0, 0.0, c->sigma_s, c->phi, c->psi, 1);
    orc_sweep(c->nx, c->ny, c->p, c->nq, c->xi, c->eta, c->quad_off, c->dx, c->dy, c->sigma_t, c->dt, c->psi);
    orc_pmv(c->nx, c->ny, c->p, c->nq, c->w, -1.0, 1.0, c->psi, c->phi, 0);
    pack_interior(c, c->phi, y);
}

static double dot(size_t n, const double *a, const double *b)
{
    double s = 0.0;
    for (size_t k = 0; k < n; ++k) s += a[k] * b[k];
    return s;
}

int orc_peierls_gmres(int nx, int ny, int p, int nq, const double *xi, const double *eta, const double *w,
                      const int *quad_off, double dx, double dy, const double *sigma_t,
                      const double *sigma_s, double dt, double abs_tol, int max_its, int restart,
                      double *source, double *phi_out, int *iters, double *resid)
{
    const int N = (p + 1) * (p + 1);
    const size_t n = (size_t)nx * ny * N;
    const size_t npsi = (size_t)(nx + 2) * (ny + 2) * nq * N;
    const size_t nphi = (size_t)(nx + 2) * (ny + 2) * N;
    const int m = restart;
    peierls_ctx c = {nx, ny, p, nq, xi, eta, w, quad_off, dx, dy, dt, sigma_t, sigma_s,
                     (double *)malloc(sizeof(double) * npsi), (double *)calloc(nphi, sizeof(double))};
    double *b = (double *)malloc(sizeof(double) * n);
    double *x = (double *)calloc(n, sizeof(double));
    double *r = (double *)malloc(sizeof(double) * n);
    double *V = (double *)malloc(sizeof(double) * n * (m + 1));
    double *H = (double *)calloc((size_t)(m + 1) * m, sizeof(double));
    double *cs = (double *)malloc(sizeof(double) * m), *sn = (double *)malloc(sizeof(double) * m);
    double *g = (double *)malloc(sizeof(double) * (m + 1)), *yv = (double *)malloc(sizeof(double) * m);

    /* RHS = P L^{-1} Q (:461-466) */
    memcpy(c.psi, source, sizeof(double) * npsi);
    orc_sweep(nx, ny, p, nq, xi, eta, quad_off, dx, dy, sigma_t, dt, c.psi);
    memset(c.phi, 0, sizeof(double) * nphi);
    orc_pmv(nx, ny, p, nq, w, 1.0, 0.0, c.psi, c.phi, 0);
    pack_interior(&c, c.phi, b);

    int it = 0, converged = 0;
    double beta = 0.0;
    while (!converged && it < max_its) {
        peierls_apply(&c, x, r);
        for (size_t k = 0; k < n; ++k) r[k] = b[k] - r[k];
        beta = sqrt(dot(n, r, r));
        if (beta <= abs_tol) { converged = 1; break; }
        for (size_t k = 0; k < n; ++k) V[k] = r[k] / beta;
        memset(g, 0, sizeof(double) * (m + 1));
        g[0] = beta;
        int jdone = 0;
        for (int j = 0; j < m && it < max_its; ++j) {
            double *vj = V + (size_t)j * n, *vn = V + (size_t)(j + 1) * n;
            peierls_apply(&c, vj, vn);
            ++it;
            for (int i = 0; i <= j; ++i) { /* modified Gram-Schmidt */
                const double h = dot(n, vn, V + (size_t)i * n);
                H[i + (m + 1) * j] = h;
                for (size_t k = 0; k < n; ++k) vn[k] -= h * V[(size_t)i * n + k];
            }
            const double hn = sqrt(dot(n, vn, vn));
            H[j + 1 + (m + 1) * j] = hn;
            if (hn > 0.0) for (size_t k = 0; k < n; ++k) vn[k] /= hn;
            for (int i = 0; i < j; ++i) { /* apply previous rotations */
                const double h0 = H[i + (m + 1) * j], h1 = H[i + 1 + (m + 1) * j];
                H[i + (m + 1) * j] = cs[i] * h0 + sn[i] * h1;
                H[i + 1 + (m + 1) * j] = -sn[i] * h0 + cs[i] * h1;
            }
            const double h0 = H[j + (m + 1) * j], h1 = H[j + 1 + (m + 1) * j];
            const double den = sqrt(h0 * h0 + h1 * h1);
            cs[j] = h0 / den; sn[j] = h1 / den;
            H[j + (m + 1) * j] = den; H[j + 1 + (m + 1) * j] = 0.0;
            g[j + 1] = -sn[j] * g[j];
            g[j] = cs[j] * g[j];
            jdone = j + 1;
            if (fabs(g[j + 1]) <= abs_tol) { converged = 1; break; }
        }
        for (int i = jdone - 1; i >= 0; --i) { /* back substitution and update */
            double s = g[i];
            for (int k = i + 1; k < jdone; ++k) s -= H[i + (m + 1) * k] * yv[k];
            yv[i] = s / H[i + (m + 1) * i];
        }
        for (int i = 0; i < jdone; ++i)
            for (size_t k = 0; k < n; ++k) x[k] += yv[i] * V[(size_t)i * n + k];
        beta = fabs(g[jdone]);
    }

    /* psi = L^{-1}(Q + S phi) (:504-505) */
    memset(c.phi, 0, sizeof(double) * nphi);
    unpack_interior(&c, x, c.phi);
    orc_smv(nx, ny, p, nq, 1.0, 1.0, sigma_s, c.phi, source, 1);
    orc_sweep(nx, ny, p, nq, xi, eta, quad_off, dx, dy, sigma_t, dt, source);
    if (phi_out) memcpy(phi_out, c.phi, sizeof(double) * nphi);
    if (iters) *iters = it;
    if (resid) *resid = beta;

    free(c.psi); free(c.phi); free(b); free(x); free(r); free(V); free(H); free(cs); free(sn); free(g); free(yv);
    return converged ? 0 : -1;
}
