/*
 * oracle.h -- CPU restatement of the reference's 2D RKDG discrete-ordinates sweep and the operators
 * around it (crockatt/idc-transport-solver, src/linear_solvers and src/operators/RKDG).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product path is compared
 * against.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product library (libtsw.so) never links, loads or calls it.
 *
 * Array layouts are the reference's host layouts, 1-cell halo on every side, 1-based interior:
 *   psi   [i][j][q][d][e]  i in 0..nx+1, j in 0..ny+1, q in 0..nq-1, d,e in 0..p
 *         (src/objects/RKDG/OrdinateFlux.hpp:639-700)
 *   phi   [i][j][d][e]     (src/objects/RKDG/DensityFunction.hpp:475-535)
 *   sigma [i][j]           cell means, DG degree 0 cross sections (src/objects/RKDG/CrossSection.hpp:397-478)
 * Ordinates are grouped by quadrant: q in [quad_off[k], quad_off[k+1]) has signs
 *   k=0: xi>0 eta>0   k=1: xi>0 eta<0   k=2: xi<0 eta>0   k=3: xi<0 eta<0
 * (src/utils/Quadrule/OrdinateSet.cpp:658-700).
 *
 * Every function states the reference lines it restates.  Pinning of each function (the checks that
 * do not merely restate the formula) lives in tests/test_oracle.py.
 */
#ifndef TSW_ORACLE_H
#define TSW_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cross-section degree of the following calls (0 default: sigma arrays are cell means).  Degree p (the
 * flux degree): every sigma argument of orc_sweep, orc_sweep_omp, orc_smv, orc_lmv, orc_si_solve and
 * orc_peierls_gmres is a [i][j][u][v] array of Legendre coefficients, entering through the tensor of
 * Legendre triple products (init.cpp:741, CrossSection::ComputeTensor CrossSection.cpp:165-189; CalcA
 * SweepLinearSolveSpec.cpp:300-321, Smv TransportOperator.cpp:375-381, Lmv :1587-1596).  Process-wide
 * setting; orc_sweep_adjoint supports degree 0 only. */
void orc_set_xs_degree(int degree);

/* tensor[((i*(ny+2)+j)*N + d*P+e)*N + a*P+b] = sigma(i,j,a,b,d,e) = sum_{u,v} sigma[i][j][u][v] TPI(a,d,u)
 * TPI(b,e,v), TPI(a,b,c) = int P_a P_b P_c on [-1,1] (ComputeTPI global.cpp:67-107, ComputeTensor
 * CrossSection.cpp:179-188, TensorIndex CrossSection.hpp:397-478). */
void orc_xs_tensor(int nx, int ny, int p, const double *sigma, double *tensor);

/* Angular part of the element matrices, one N x N column-major matrix per ordinate,
 * Aq[q*N*N + row + N*col], row/col = e + (p+1)*d.
 * SweepLinearSolveSpec.cpp:98-158 (2D), index map SweepCellSolve.hpp:310-378. */
void orc_aq(int p, int nq, const double *xi, const double *eta, double dx, double dy, double *Aq);

/* Unpivoted Gaussian elimination + back substitution on one column-major system, in place.
 * SIMD_LinearAlgebra.hpp:76-118 with SIMD_length = 1. */
void orc_ge_solve(int N, double *A, double *b);

/* In-place sweep psi <- L^{-1} psi.  Halo cells of psi hold the inflow values.
 * dt <= 0 or dt = +inf means steady state (1/dt = 0).
 * SweepOperator.cpp:118-135 -> SweepPatternUniCartWavefront.cpp:253-296 ->
 * SweepCellSolveBase.cpp:86-103 (CalcB SweepLinearSolveSpec.cpp:473-598, CalcA :284-298,
 * SolveSystem SIMD_LinearAlgebra.hpp:76-118, StoreResult SweepCellSolve.cpp:86-99). */
void orc_sweep(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
               double dx, double dy, const double *sigma_t, double dt, double *psi);

/* orc_sweep with each diagonal's cells x ordinates shared among `nthreads` OpenMP threads (one barrier
 * per diagonal, SweepPatternUniCartWavefront.cpp:146,253-296); bitwise equal to orc_sweep. */
/* Reflecting edges (mask bits 1 x-min, 2 x-max, 4 y-min, 8 y-max): the mirror map [2][nq] from (xi, eta)
 * (0, or -1 when an ordinate has none or two ordinates share (xi, eta)); OrdinateFlux::ReflectBoundaries;
 * SweepOperator::Apply with reflecting edges (quadrants in the order 3, 2, 1, 0, ghosts reflected after
 * each). */
int orc_reflect_map(int nq, const double *xi, const double *eta, int *map);
/* DensityFunction::SynchronizeHalos across periodic boundaries (every value of the cell, corners untouched). */
void orc_periodic_halos(int nx, int ny, int p, int nq, double *psi);
void orc_reflect_boundaries(int nx, int ny, int p, int nq, const int *quad_off, const int *map, int mask, double *psi);
void orc_sweep_reflect(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
                       double dx, double dy, const double *sigma_t, double dt, int mask, const int *map, double *psi);
void orc_sweep_omp(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
                   double dx, double dy, const double *sigma_t, double dt, double *psi, int nthreads);

/* Transposed sweep: out (interior) <- zeta = L^{-T} g for the map src -> psi of orc_sweep with zero
 * inflow, where L^{-1} src solves A_c psi_c - (upwind couplings) = dx dy src_c.  out's halo cells
 * receive the gradient of <g, psi> with respect to the inflow halo values (zero on outflow sides and
 * corners).  g's halo cells are ignored.  The reference has no adjoint; this transposes the system
 * of CalcA/CalcB (SweepLinearSolveSpec.cpp:284-298,473-598) directly. */
void orc_sweep_adjoint(int nx, int ny, int p, int nq, const double *xi, const double *eta, const int *quad_off,
                       double dx, double dy, const double *sigma_t, double dt, const double *g, double *out);

/* phi <- beta*phi + alpha * sum_q w_q psi_q.  all_cells: 0 = interior only, 1 = halos too.
 * TransportOperator.cpp:140-158. */
void orc_pmv(int nx, int ny, int p, int nq, const double *w, double alpha, double beta,
             const double *psi, double *phi, int all_cells);

/* psi_q <- beta*psi_q + alpha * sigma_s * phi / (4 pi) for every q.  TransportOperator.cpp:369-392. */
void orc_smv(int nx, int ny, int p, int nq, double alpha, double beta, const double *sigma_s,
             const double *phi, double *psi, int all_cells);

/* y <- beta*y + alpha*(Omega.grad + sigma_t) x on interior cells, x's halos give the inflow.
 * TransportOperator.cpp:1570-1660 (2D, non-blocked form). */
void orc_lmv(int nx, int ny, int p, int nq, const double *xi, const double *eta, double dx, double dy,
             double alpha, double beta, const double *sigma_t, const double *x, double *y);

/* sqrt(sum over interior coefficients of (a-b)^2).  DensityFunction.cpp:291-357. */
double orc_l2dist(int nx, int ny, int p, const double *a, const double *b);

/* Source iteration.  psi: solver state (initial guess in, last iterate out); source: Q in, psi out.
 * SISolver.cpp:190-254.  Returns 0. */
int orc_si_solve(int nx, int ny, int p, int nq, const double *xi, const double *eta, const double *w,
                 const int *quad_off, double dx, double dy, const double *sigma_t, const double *sigma_s,
                 double dt, double abs_tol, int max_its, double *psi, double *source, int *iters,
                 double *err);

/* Restarted GMRES(m) on the Peierls form (I - P L^{-1} S) phi = P L^{-1} Q, then
 * source <- L^{-1}(Q + S phi).  PETScPeierlsSolver.cpp:423-514 (solve), :529-572 (matvec IPLS).
 * Convergence: ||r||_2 <= abs_tol (PETSc test with rtol = 1e-60, :139-143).  phi_out: interior
 * scalar flux [i][j][d][e] with halo (zero halos).  Returns 0 on convergence, -1 otherwise. */
int orc_peierls_gmres(int nx, int ny, int p, int nq, const double *xi, const double *eta, const double *w,
                      const int *quad_off, double dx, double dy, const double *sigma_t,
                      const double *sigma_s, double dt, double abs_tol, int max_its, int restart,
                      double *source, double *phi_out, int *iters, double *resid);

#ifdef __cplusplus
}
#endif

#endif
