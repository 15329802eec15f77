/*
 * tsw.h -- C-ABI of the MI355X transport-sweep library (libtsw.so).
 *
 * The reference's sweep path is C++ (crockatt/idc-transport-solver).  Each entry point below
 * replaces one member of its operator interface; the citation names the reference declaration and
 * the behaviour kept.  Conventions:
 *   - every function returns int status: TSW_OK (0) or a negative TSW_E_* code; no exception
 *     crosses the ABI; tsw_last_error(ctx) returns the message of the last failure on ctx;
 *   - arrays use the reference host layouts (1-cell halo on each side, 1-based interior):
 *       psi   [i][j][q][d][e]  (nx+2)*(ny+2)*nq*(p+1)^2 doubles  (OrdinateFlux.hpp:639-700)
 *       phi   [i][j][d][e]     (nx+2)*(ny+2)*(p+1)^2 doubles     (DensityFunction.hpp:475-535)
 *       sigma [i][j]           (nx+2)*(ny+2) doubles, cell means   (CrossSection.hpp:397-478, DG degree 0),
 *             or [i][j][u][v] Legendre coefficients of DG degree p after tsw_set_xs_degree(ctx, p)
 *   - `mem` says where the caller's arrays live: TSW_MEM_HOST (pageable or pinned host memory)
 *     or TSW_MEM_DEVICE (device pointers on the context's device, same layouts);
 *   - all buffers are caller-owned; the context owns only its scratch and solver state;
 *   - every call is complete (device work finished) when it returns, as the reference's CPU
 *     calls are; one context per host thread.
 */
#ifndef TSW_H
#define TSW_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TSW_OK 0
#define TSW_E_ARG (-1)         /* invalid argument (std::invalid_argument / std::out_of_range) */
#define TSW_E_HIP (-2)         /* HIP runtime failure */
#define TSW_E_STATE (-3)       /* call made before tsw_problem_set */
#define TSW_E_NOCONV (-4)      /* iterative solve did not converge (PETScPeierlsSolver.cpp:489-497) */
#define TSW_E_UNSUPPORTED (-5) /* configuration not built (e.g. DG degree > 3) */

#define TSW_FP64 0 /* sweep and element solves in double precision */
#define TSW_FP32 1 /* sweep and element solves in single precision (outer loops stay FP64) */
#define TSW_MIXED 2 /* sweeps (tsw_sweep_apply, tsw_batch_sweep): FP32 sweeps with FP64 iterative refinement,
                     * psi_0 = L32^{-1} src, then refine_steps times psi += L32^{-1}(src - L psi) with the
                     * residual in FP64; on the engines that sweep the caller's FP64 arrays in place
                     * (batch members on the ordinate-lane engine or the staggered kernels on the
                     * caller's layout, tsw_sweep_apply_multi) the FP64 sweep, which moves the same
                     * bytes and reaches the accuracy without a second sweep; solvers treat it as
                     * TSW_FP32 */

#define TSW_MEM_HOST 0
#define TSW_MEM_DEVICE 1

typedef struct tsw_ctx tsw_ctx;

/* Creates a context on HIP device `device`.  Replaces the construction of SweepOperator
 * (SweepOperator.cpp:25-49) minus the mesh, which tsw_problem_set supplies. */
int tsw_create(int device, tsw_ctx **out);

/* Releases the context and its device memory.  Safe on NULL. */
void tsw_destroy(tsw_ctx *ctx);

/* Message of the last failing call on ctx ("" if none).  Valid until the next call on ctx. */
const char *tsw_last_error(const tsw_ctx *ctx);

/* Mesh + ordinate set + angular matrices: DomainDecomposition (DomainDecomposition.hpp:35-62,
 * uniform Cartesian [ax,bx]x[ay,by] with nx*ny cells), OrdinateSet accessors xi/eta/w/Quadrants
 * (OrdinateSet.hpp:236-418) and SweepLinearSolveSpec::ComputeMatrices (SweepLinearSolveSpec.cpp:47-178).
 * quad_off[5]: ordinates q in [quad_off[k], quad_off[k+1]) lie in quadrant k
 * (0: xi>0,eta>0  1: xi>0,eta<0  2: xi<0,eta>0  3: xi<0,eta<0; OrdinateSet.cpp:658-700).
 * p in 0..3.  Resets solver state. */
int tsw_problem_set(tsw_ctx *ctx, int nx, int ny, double ax, double bx, double ay, double by, int p, int nq,
                    const double *xi, const double *eta, const double *w, const int *quad_off);

/* Selects the traversal, as the `sweep_pattern` factory key does (SweepPattern.cpp:314-359):
 *   "hip-wavefront"        one launch per diagonal i+j = const, all quadrants per launch, replayed
 *                          from a captured HIP graph; reference layout;
 *   "hip-wavefront-eager"  the same launches issued directly;
 *   "hip-staggered"        (default) one launch per sweep, one workgroup per ordinate, a diagonal per
 *                          step, on a diagonal-major device layout; element solves by precomputed
 *                          inverses for piecewise-constant materials (<= 64 distinct sigma_t);
 *   "hip-staggered-ge"     the same with per-cell elimination.
 * The staggered patterns split each ordinate's rows into pipelined bands (tsw_set_tuning).
 * Unknown names fail with TSW_E_ARG as the reference's factory throws std::out_of_range. */
int tsw_set_pattern(tsw_ctx *ctx, const char *name);

/* Degree of the cross sections that the following calls on ctx take (sigma_t, sigma_s): 0 (default, and
 * after every tsw_problem_set) cell means [i][j]; the problem's DG degree p, Legendre coefficients
 * [i][j][u][v], (p+1)^2 doubles per cell, the reference's CrossSection of degree p (init.cpp:741 with
 * piecewise_cross false).  They enter through the tensor of Legendre triple products (ComputeTPI
 * global.cpp:67-107, ComputeTensor CrossSection.cpp:165-189): the element matrices gain
 * dx dy/4 (2d+1)(2e+1) sigma(i,j,d,e,u,v) (CalcA, SweepLinearSolveSpec.cpp:300-321), Smv and Lmv project
 * sigma times the flux (TransportOperator.cpp:375-381, :1587-1596), and DSA takes the cell means.  Sweeps
 * and solves then run on the wavefront kernels with per-cell elimination (any pattern).  Not built:
 * tsw_sweep_vjp, tsw_solve_vjp and tsw_idc_adjoint fail with TSW_E_UNSUPPORTED; tsw_batch_* always take
 * cell means.  Other degrees fail with TSW_E_ARG. */
int tsw_set_xs_degree(tsw_ctx *ctx, int degree);

/* Tuning of the staggered sweep.  key "band_rows": rows of an ordinate's mesh handled by one workgroup;
 * a sweep then runs ceil(ny / band_rows) pipelined workgroups per ordinate (the KBA decomposition of
 * SweepPatternUniCartKBA2.cpp:212-453 applied to rows within the device).  0 (default) chooses
 * automatically, -1 keeps one workgroup per ordinate (bands only when ny > 448).  Key "prefetch": steps
 * of cell inputs each row keeps in flight (2, 3, 4 or 6; 0 = 2 in FP64, 3 in FP32).  Key "one_wave"
 * (0 default, 1): bands of at most 64 rows, each swept by a single wavefront that passes the y-traces
 * lane to lane instead of through LDS.  Key "lane_split" (0 default, 1): on the material-table path,
 * 2 (DG degree 1) or 4 (degrees 2, 3) lanes per row, each solving its share of the cell's coefficients
 * (degrees 1-3; degree 0 runs the one-lane kernel).  Key "batch_direct" (0 default or 1, -1): launch groups of
 * tsw_batch_sweep store from the compute lanes instead of a storer wave (smaller workgroups, more of
 * them per CU), or keep the storer wave where the tables fit with it.  Key "autotune" (1 default, 0
 * off): while band_rows, prefetch, one_wave and lane_split are all 0, the first sweep of each kind
 * (mesh, degree, ordinates, precision, source terms, material path) times every candidate (kernel, band
 * rows, prefetch) into scratch memory and keeps the fastest for later sweeps; inside a HIP graph capture
 * the defaults apply.  Key "lanes" (0 default, 1, -1): the ordinate-lane engine (J ordinates of a quadrant
 * block times 64/J rows per wavefront, reading and writing the caller's layout; DG degrees 0-2, material
 * path) runs single-problem sweeps, GMRES and source iteration with 1; for tsw_batch_set, 1 / -1 force /
 * forbid it per DG degree and 0 leaves the choice to the batch autotuner.  "lanes_block" (0 automatic,
 * 1..64): ordinates per block; "lanes_waves" (0 automatic, 1..4): row-group waves per workgroup; "lanes_prefetch" (0
 * default = 2 in FP64, 3 in FP32; or 4 in FP64, 6 in FP32): steps of inputs in flight per lane.  Key
 * "batch_ref" (0 default, 1, -1): tsw_batch_set's staggered groups on the caller's layout (no layout
 * conversions) forced, forbidden, or left to the autotuner, which times, per DG degree, the staggered
 * kernels through conversions, on the caller's layout, the ordinate-lane engine and the lane-split
 * kernel on the caller's layout, and keeps the fastest.  Key "batch_split" (0 default, 1, -1): that
 * lane-split engine (DG degrees 1-3, row bands with exchange records, a band's workgroup at most 512
 * threads) forced, forbidden or autotuned; "batch_split_rows" (0 default = the most rows one workgroup
 * holds: 256 at degree 1, 128 at degrees 2-3, or 64 where the batch autotuner finds that faster; or
 * 16..256): rows per band; "batch_split_prefetch" (0
 * default = 4, or 2, 3, 6): its prefetch depth; "batch_xcd_tile" (0 default = dispatch order, or
 * 2..64): runs of that many consecutive items (neighbouring ordinates of a member) go to one XCD (with
 * "batch_persistent" 0); "batch_persistent" (1 default, 0): each lane-split and ordinate-lane group is one
 * persistent launch of as many workgroups as the device holds, taking items in order from a device counter
 * (0: one workgroup per item).  Key "ref_stage" (0 default, 1): sweeps on the caller's layout (batch groups,
 * tsw_sweep_apply_multi) stage their source loads through LDS where it fits.  Results do not depend on
 * any of these settings (bit for bit).
 * Solver keys: "use_dsa" (0 default, 1): diffusion synthetic acceleration of tsw_gmres_solve and
 * tsw_solve_vjp, the reference's `use_dsa` parameter (PETScPeierlsSolver.cpp:151, shell preconditioner
 * :180-182, apply :588-640), applied on the left as PETSc does, so the residual test is on
 * y = r + C^{-1} sigma_s r; C is the cell-mean finite-volume diffusion operator (D = 1/(3(sigma_t+1/dt)),
 * vacuum Marshak faces) inverted approximately by "dsa_cycles" (1..16, default 1) multigrid V-cycles.
 * "si_graph" (1 default, 0): staggered source iteration replays captured graphs of 8 iterations with
 * the loop test on the device (same iterates and counts as the eager loop).  "refine_steps" (1..8,
 * default 1): FP64 refinement steps of a TSW_MIXED sweep.  "gmres_refine" (0 default, 1): mixed-precision
 * GMRES by iterative refinement instead of the relaxed strategy.  "fp32_safety" (1..1000, default 10): the
 * relaxed strategy admits FP32 products once the residual estimate is <= abs_tol / (fp32_safety d), d the
 * measured relative error of an FP32 product (the final FP64 residual check guards the result).
 * Test key "stall_test" (1): the next banded staggered sweep starts with the row-band stall flag raised,
 * so its bounded waits give up at once and the call fails with TSW_E_HIP as a real stall would (the
 * exchange records are refilled and later sweeps are unaffected).
 * Unknown keys fail with TSW_E_ARG. */
int tsw_set_tuning(tsw_ctx *ctx, const char *key, int value);

/* Reads back "band_rows" (rows per band of the last staggered sweep, 0 = one band per ordinate),
 * "prefetch" (its prefetch depth), "one_wave" (1 if it ran the one-wavefront kernel), "lane_split" (1 if
 * it ran the lane-split kernel), "lanes" (1 if the last sweep, or for a batch context any member, ran on
 * the ordinate-lane engine), "lanes_block" / "lanes_waves" (its shape), "batch_ref" (1 if the last
 * tsw_batch_sweep ran staggered or lane-split groups on the caller's layout), "batch_engine_p0" ..
 * "batch_engine_p3" (the engine tsw_batch_set chose for that DG degree: 0 staggered through conversions,
 * 1 staggered on the caller's layout, 2 ordinate lanes, 3 lane-split on the caller's layout, -1 none in
 * the batch), "batch_split_rows_p0" .. "batch_split_rows_p3" (the lane-split engine's rows per band the
 * batch autotuner chose for that degree, 0: the most a workgroup holds), or the setting of
 * "batch_direct", "ref_stage", "batch_split", "batch_split_rows",
 * "batch_split_prefetch" (the depth in use), "batch_persistent", "batch_xcd_tile", "ref_split", "autotune",
 * "use_dsa", "dsa_cycles", "si_graph", "refine_steps", "gmres_refine", "fp32_safety". */
int tsw_get_tuning(tsw_ctx *ctx, const char *key, int *value);

/* psi <- L^{-1} psi in place; halo cells of psi hold the inflow and are left unchanged; interior
 * cells are overwritten.  dt <= 0 or +inf is steady state.  SweepOperator::Apply
 * (SweepOperator.hpp:126-131, SweepOperator.cpp:118-135).  precision TSW_FP32 converts psi and sigma
 * to single precision for the sweep and back; TSW_MIXED refines that FP32 sweep to FP64 accuracy.
 * With the hip-staggered pattern, DG degree 1-3, the staggered tuning keys at their defaults and a
 * 16-byte aligned device array (host arrays are staged into one), the lane-split kernel sweeps psi in
 * place on this layout with no layout conversion (tuning key "ref_split", 1 default, -1 off); FP32 then
 * computes in single precision from and into the FP64 array, and TSW_MIXED is the FP64 sweep. */
int tsw_sweep_apply(tsw_ctx *ctx, double *psi_inout, const double *sigma_t, double dt, int precision, int mem);

/* Many right-hand sides: psi_inout[r] <- L^{-1} psi_inout[r] for r < n_rhs, each as tsw_sweep_apply would
 * sweep it (same sigma_t and dt).  It has no counterpart in the reference, whose SweepOperator::Apply
 * (SweepOperator.cpp:118-135) takes one source: a caller with several sources of one problem (sources of
 * several detectors, the stages of an implicit step) calls it once instead of Apply per source.  On the
 * "hip-staggered" pattern with ny <= 448 and 16-byte aligned arrays the sweeps run as one launch over
 * (source, ordinate) work items that share the material table and element inverses, reading and writing
 * the caller's arrays directly; results are bitwise those of tsw_sweep_apply with the tuning key "lanes"
 * at -1 in TSW_FP64 and TSW_FP32, and TSW_MIXED is the FP64 sweep there (it moves the same FP64 arrays and
 * its FP32 arithmetic is not faster).  Elsewhere the sources are swept one after the other. */
int tsw_sweep_apply_multi(tsw_ctx *ctx, int n_rhs, double *const *psi_inout, const double *sigma_t, double dt,
                          int precision, int mem);

#define TSW_REFLECT_X_MIN 0x01 /* BoundaryEdge::X_Min (global.hpp:104) */
#define TSW_REFLECT_X_MAX 0x02 /* BoundaryEdge::X_Max */
#define TSW_REFLECT_Y_MIN 0x04 /* BoundaryEdge::Y_Min */
#define TSW_REFLECT_Y_MAX 0x08 /* BoundaryEdge::Y_Max */
#define TSW_PERIODIC 0x100     /* Global::periodic (the deck key periodic) */

/* Reflecting edges of the current problem: the deck keys reflect_x_min / _x_max / _y_min / _y_max that
 * Init turns into Global::reflecting_boundaries (init.cpp:448-477); 0 restores vacuum/inflow edges
 * everywhere (tsw_problem_set also does).  reflect_map [2][nq] gives, per direction (0 x, 1 y) and
 * ordinate, the index of its mirror (OrdinateSet::ReflectMap, OrdinateSet.cpp:1093-1140); null computes it
 * from xi and eta as the reference does (the first ordinate matching within 1e-14), which fails with
 * TSW_E_ARG when two ordinates share (xi, eta) (an unreduced set's +-mu pairs) or one has no mirror.
 * With reflecting edges tsw_sweep_apply sweeps the quadrants one after the other in the order 3, 2, 1, 0
 * (SweepPatternUniCart.cpp:33-40) and reflects the ghost cells after each one
 * (SweepPatternUniCartWavefront.cpp:180-245), on the "hip-staggered" pattern with cell-mean cross sections,
 * ny <= 448 and a 16-byte aligned psi (else TSW_E_UNSUPPORTED); its result, ghost cells included, is that
 * of the reference's sweep.  tsw_si_solve then runs SISolver::Solve's loop with the ghost cells of the
 * iterate reflected before each source copy (SynchronizeHalos, SISolver.cpp:222) and the source copied
 * onto the interior and the other edges' ghost cells (OpDomain::Interior | OpDomain::Boundary).
 * reflect_mask = TSW_PERIODIC (alone: the reference refuses the combination, init.cpp:498-510) sets periodic
 * boundaries: sweeps take the halos as given, as SweepOperator::Apply does, and tsw_si_solve copies the
 * iterate's halos across the periodic boundaries before each iteration and the source onto the interior
 * only.  GMRES, the IDC integrator and the vector-Jacobian products refuse reflecting and periodic
 * boundaries with TSW_E_UNSUPPORTED. */
int tsw_set_boundaries(tsw_ctx *ctx, int reflect_mask, const int *reflect_map);

/* OrdinateFlux::SynchronizeHalos (Abstract/OrdinateFlux.cpp:200-205): ReflectBoundaries on the reflecting
 * edges, then with TSW_PERIODIC the periodic copy of DensityFunction::SynchronizeHalos
 * (Abstract/DensityFunction.cpp:927-956: the y-halos of columns 1..nx from rows ny and 1, then the x-halos
 * of rows 1..ny from columns nx and 1, every value of the cell; corners untouched). */
int tsw_synchronize_halos(tsw_ctx *ctx, double *psi_inout, int mem);

/* OrdinateFlux::ReflectBoundaries (RKDG/OrdinateFlux.cpp:1181-1420) on psi for the edges set by
 * tsw_set_boundaries: ghost cells of x-min (quadrants 0, 1), x-max (2, 3), y-min (0, 2) and y-max (all
 * quadrants, as the reference does) take the mirrored ordinate's values of the adjacent interior cell,
 * coefficients odd in the reflected direction negated.  Nothing happens without reflecting edges. */
int tsw_reflect_boundaries(tsw_ctx *ctx, double *psi_inout, int mem);

/* tsw_sweep_apply without any device allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It has no counterpart in the
 * reference.  psi_inout and sigma_t are device arrays (TSW_MEM_DEVICE layouts, cell-mean cross sections);
 * the sweep runs the wavefront kernels of "hip-wavefront-eager" whatever the pattern, so its results are
 * bitwise those of tsw_sweep_apply on that pattern.  TSW_FP32 and TSW_MIXED stage their FP32 copies and
 * FP64 iterate in a per-GPU scratch array of 128 MiB held by the library for the life of the process
 * (4 bytes per psi value + 4 per cell for TSW_FP32, 20 + 4 for TSW_MIXED, none for TSW_FP64); a problem
 * that needs more fails with TSW_E_ARG before any work.  Every call on a GPU shares that GPU's scratch:
 * calls that overlap on one GPU (several contexts or threads) race.  The context must have been created
 * and given its problem beforehand (those calls allocate); the call itself makes no hipMalloc/hipFree. */
int tsw_sweep_apply_noalloc(tsw_ctx *ctx, double *psi_inout, const double *sigma_t, double dt, int precision);

/* phi <- beta*phi + alpha*sum_q w_q psi_q.  TransportOperator::Pmv (TransportOperator.hpp:25-43,
 * TransportOperator.cpp:37-186).  all_cells: 0 = OpDomain::Interior, 1 = OpDomain::All. */
int tsw_pmv(tsw_ctx *ctx, double alpha, double beta, const double *psi, double *phi, int all_cells, int mem);

/* psi_q <- beta*psi_q + alpha*sigma_s*phi/(4 pi).  TransportOperator::Smv (TransportOperator.cpp:203-443). */
int tsw_smv(tsw_ctx *ctx, double alpha, double beta, const double *sigma_s, const double *phi, double *psi,
            int all_cells, int mem);

/* y <- beta*y + alpha*(Omega.grad + sigma_t) x on interior cells; x's halos give the inflow.
 * TransportOperator::Lmv (TransportOperator.cpp:1215-1683); x == y is the in-place product
 * (Lmv_InPlace, :602-1200, dispatched at :1224-1228): every cell uses the values of y from before the
 * call. */
int tsw_lmv(tsw_ctx *ctx, double alpha, double beta, const double *sigma_t, const double *x, double *y, int mem);

/* Source iteration, ImplicitSolver::Solve under solve_type "source-iteration" (SISolver.cpp:190-254):
 * source holds Q on entry (halos = inflow) and psi on exit.  The solver's angular-flux state
 * persists in ctx between calls (zero after tsw_problem_set), as SISolver::psi_ptr does.
 * Stops when the absolute l2 change of the scalar-flux coefficients is <= abs_tol or after max_its.
 * iters / err (may be NULL) receive the iteration count and the last change. */
int tsw_si_solve(tsw_ctx *ctx, double *source_inout, const double *sigma_t, const double *sigma_s, double dt,
                 double abs_tol, int max_its, int precision, int mem, int *iters, double *err);

/* Sweep-preconditioned Krylov solve on the Peierls form (I - P L^{-1} S) phi = P L^{-1} Q, then
 * source <- L^{-1}(Q + S phi): ImplicitSolver::Solve under solve_type "petsc-peierls-gmres"
 * (PETScPeierlsSolver.cpp:423-514, matvec IPLS :529-572), restarted GMRES(restart) with the
 * absolute residual test ||r||_2 <= abs_tol (:139-143).  phi_out (may be NULL): converged scalar
 * flux, phi layout.  precision TSW_FP64: FP64 sweeps.  TSW_FP32 (staggered patterns): relaxed inexact
 * GMRES, FP64 Krylov vectors and DSA with FP32 sweeps in the Arnoldi steps taken once the residual
 * estimate is <= abs_tol / (10 d), d the relative error of an FP32-sweep product measured once per cross
 * sections, dt and DSA setting (FP64 sweeps for the right-hand side, earlier steps, restarts and the
 * final psi); the final FP64 sweep also gives the true residual, which *resid reports, and the solve
 * resumes while it exceeds abs_tol.  With tuning key "gmres_refine" = 1, and on the wavefront patterns,
 * FP32 sweeps inside an FP64 iterative refinement on the angular residual r = Q + S P psi - L psi (FP64
 * Lmv), repeated until ||P r||_2 <= abs_tol.  Returns TSW_E_NOCONV if max_its is reached. */
int tsw_gmres_solve(tsw_ctx *ctx, double *source_inout, const double *sigma_t, const double *sigma_s, double dt,
                    double abs_tol, int max_its, int restart, int precision, int mem, double *phi_out,
                    int *iters, double *resid);

/* Vector-Jacobian product of tsw_sweep_apply as a map (source with inflow halos, sigma_t) -> psi, for
 * J = <seed, psi> over every coefficient of the array.  The reference has no adjoint; this is the exact
 * transpose of the system its sweep solves (SweepLinearSolveSpec.cpp:284-298 matrix, :473-598 right-hand
 * side, with the upwind couplings transposed).  Outputs:
 *   grad_src   (psi layout, required): interior = zeta = L^{-T} seed; halo cells = seed's halo (the sweep
 *              passes halos through) + the gradient with respect to the inflow values;
 *   grad_sigma_t (cell array, may be NULL): interior = -sum_{q,d,e} zeta psi, halo cells 0; needs `psi`,
 *              the forward sweep's output.
 * seed's interior and halo cells are both read; grad_src may alias seed.  precision as tsw_sweep_apply. */
int tsw_sweep_vjp(tsw_ctx *ctx, const double *psi, const double *sigma_t, double dt, const double *seed,
                  double *grad_src, double *grad_sigma_t, int precision, int mem);

/* Vector-Jacobian product of the scattering solve (tsw_gmres_solve / tsw_si_solve) as a map
 * (Q with inflow halos, sigma_t, sigma_s) -> psi: the adjoint problem
 * zeta = L^{-T}(seed + K^T zeta), (K^T zeta)_q = w_q sigma_s sum_q' zeta_q' / (4 pi), is solved by the same
 * Peierls GMRES (restart, abs_tol, max_its, precision as tsw_gmres_solve; iters/resid may be NULL).
 *   grad_src: interior = zeta = dJ/dQ, halos = seed's halo + dJ/d(inflow);
 *   grad_sigma_t (may be NULL): -sum zeta psi;  grad_sigma_s (may be NULL): sum_{q,k} zeta_qk phi_k / (4 pi),
 *   phi = P psi.  Both need `psi`, the converged forward solution.
 * Returns TSW_E_NOCONV if the adjoint solve does not converge. */
int tsw_solve_vjp(tsw_ctx *ctx, const double *psi, const double *sigma_t, const double *sigma_s, double dt,
                  const double *seed, double *grad_src, double *grad_sigma_t, double *grad_sigma_s, double abs_tol,
                  int max_its, int restart, int precision, int mem, int *iters, double *resid);

/* IDC node families, the reference's idc_nodes_type (Quadrule::NodesType, Quadrule.hpp:30-70) */
#define TSW_IDC_LOBATTO 0         /* "gauss-lobatto": both ends of the step are nodes */
#define TSW_IDC_LEGENDRE 1        /* "gauss-legendre": interior nodes; output by collocation */
#define TSW_IDC_RADAU 2           /* "gauss-radau": Radau IIA, the right end is a node */
#define TSW_IDC_CHEBYSHEV 3       /* "chebyshev": interior nodes; output by collocation */
#define TSW_IDC_CHEBYSHEV_RADAU 4 /* "chebyshev-radau": the right end is a node */

/* IDC forms, the reference's idc_type (IDCIntegrator::IDCType, idc.cpp:33-50) */
#define TSW_IDC_ERROR 0  /* "error": corrections solve for the error and add it to the stage */
#define TSW_IDC_UPDATE 1 /* "update": corrections solve for the updated stage directly */

/* IDC integrator configuration, the reference's idc_num_nodes, idc_num_corr, idc_nodes_type and
 * idc_use_collocation parameters (idc.cpp:223, SetNodes :2305-2376): num_nodes quadrature nodes on each
 * step (2..12 Lobatto, 1..12 for the other families), num_corrections correction sweeps (0..32).
 * Collocation forms the result from the integration matrix; it is always used for families without the
 * right end (as the reference's Step_Nonhybrid does, :933-936).  Defaults: 4 Lobatto nodes, 3
 * corrections, no collocation, error form.  Only the non-hybrid integrator is built. */
int tsw_idc_set(tsw_ctx *ctx, int num_nodes, int num_corrections, int nodes_type, int use_collocation);

/* IDC form (TSW_IDC_ERROR default, TSW_IDC_UPDATE) for the following steps, the reference's idc_type.
 * Both forms give the same step up to the implicit solves' tolerance; the update form's solves start
 * from the stage's previous iterate in the reference (SetInitialGuess, idc.cpp:893-901), which changes
 * only their iteration counts.  tsw_idc_adjoint transposes the step, which is the same linear map in
 * both forms, so it ignores this setting.  Other values fail with TSW_E_ARG. */
int tsw_idc_set_type(tsw_ctx *ctx, int idc_type);

/* Stage nodes on [0, 1] and integration matrix of an IDC configuration (host-only, needs no device):
 * num_stages = num_nodes + 1 for families without the left end (t = 0 becomes stage 0, as
 * num_stages = num_nodes + !left_endpoint in SetNodes), num_nodes otherwise; nodes[S] ascending; W[S*S]
 * row-major, row 0 the integrals over [0, 1] of the nodes' Lagrange basis, row n > 0 those over
 * [t_{n-1}, t_n] (ComputeQuadratureWeights, columns of non-node stages zero).  Roots are found by Newton
 * in quadruple precision and rounded once.  Writes *num_stages; fills nodes and W when nodes is non-NULL
 * (capacity >= num_stages).  Bad types or counts fail with TSW_E_ARG. */
int tsw_idc_nodes(int nodes_type, int num_nodes, int capacity, int *num_stages, double *nodes, double *W);

/* One IDC step psi^n -> psi^{n+1} in place, IDCIntegrator::Step_Nonhybrid (idc.cpp:846-940) in the form
 * tsw_idc_set_type chose: implicit-Euler prediction over the nodes, corrections whose sources integrate the
 * residuals (L - S P) psi with the nodes' integration matrix (ComputeSource :1686-1740), result at the
 * last node or by collocation.  psi's halos are taken as zero (as the reference assumes) and returned
 * zero; source holds Q with the inflow in its halos, constant over the step.  Each implicit solve is the
 * Peierls GMRES of tsw_gmres_solve (restart 30, abs_tol, max_its, precision: FP32 sweeps with FP64 Krylov
 * vectors for TSW_FP32); residuals and correction sums are FP64.  TSW_E_NOCONV if a solve fails. */
int tsw_idc_step(tsw_ctx *ctx, double *psi_inout, const double *source, const double *sigma_t,
                 const double *sigma_s, double dt, double abs_tol, int max_its, int precision, int mem);

/* nsteps IDC steps of size dt in place (TimeIntegrator stepping loop); arguments as tsw_idc_step. */
int tsw_idc_run(tsw_ctx *ctx, double *psi_inout, const double *source, const double *sigma_t, const double *sigma_s,
                double dt, int nsteps, double abs_tol, int max_its, int precision, int mem);

/* Adjoint of tsw_idc_run for J = <seed, psi^{nsteps}> (interior coefficients): grad_psi0 (psi layout,
 * halos 0) receives dJ/dpsi^0; grad_sigma_t / grad_sigma_s (cell arrays, may be NULL) receive dJ/dsigma
 * summed over every implicit solve and residual of every step.  Each step's transpose reverses its
 * operations: transposed solves by the adjoint problem's GMRES, (L - S P)^T for the residuals.  For the
 * cross-section gradients the forward trajectory is run once keeping psi every `checkpoint_every`
 * steps, and each checkpoint segment is recomputed before it is stepped back through; without them no
 * forward state is needed (the step is linear in psi) and psi0 may be NULL.  The reference has no
 * adjoint time integrator. */
int tsw_idc_adjoint(tsw_ctx *ctx, const double *psi0, const double *source, const double *sigma_t,
                    const double *sigma_s, double dt, int nsteps, int checkpoint_every, const double *seed,
                    double *grad_psi0, double *grad_sigma_t, double *grad_sigma_s, double abs_tol, int max_its,
                    int precision, int mem);

/* Level-symmetric S_order ordinates reduced to the 2D hemisphere (order in {2,4,6,8,10,12,16}):
 * nq = order*(order+2)/2 directions, weights summing to 4 pi, grouped by quadrant as above.
 * The reference has no level-symmetric set; this plays the role of OrdinateSet's constructors
 * (OrdinateSet.cpp:40-91) and SymmetricReduce (:1171-1190).  Host-only, needs no device.
 * Arrays must hold nq entries, quad_off 5. */
int tsw_ordinates_level_symmetric(int order, double *xi, double *eta, double *w, int *quad_off);

/* The ordinate set a deck names by ordinate_type / ang_order / ordinate_sym_reduce, replacing the
 * OrdinateSet constructor (OrdinateSet.cpp:126-287) for SPACE_DIMS == 2:
 *   "spherical-triangle": T_N, 8 N^2 directions on the sphere (SphericalTriangle.cpp:268-351,
 *     OrdinateSet.cpp:490-520), in the reference's order: the first octant's triangles in generation
 *     order, then the octants made by reflecting mu, eta and xi;
 *   "chebyshev-legendre": 2 N^2 directions (OrdinateSet.cpp:441-488), N even, sorted as the reference
 *     sorts them (descending octant, xi, eta, mu);
 *   "level-symmetric": as tsw_ordinates_level_symmetric (always reduced; no reference counterpart).
 * With symmetric_reduce the pairs (xi, eta, +-mu) are merged as SymmetricReduce does (:1171-1190),
 * halving the count.  Weights sum to 4 pi.  quad_off[k] starts quadrant k (0: xi>0, eta>0; 1: xi>0,
 * eta<0; 2: xi<0, eta>0; 3: xi<0, eta<0), quad_off[4] = nq.  Writes *nq; the arrays are filled when xi
 * is non-NULL and must then hold `capacity` >= nq entries.  Unknown types (the reference's "lebedev"
 * tables are not shipped with it) and orders fail with TSW_E_ARG.  Host-only, needs no device. */
int tsw_ordinates(const char *type, int order, int symmetric_reduce, int capacity, int *nq, double *xi,
                  double *eta, double *w, int *quad_off);

/* n-point Gauss-Legendre rule on [-1, 1]: nodes ascending, exactly antisymmetric, weights summing to 2
 * (the rules Quadrule::ComputeQuadrature returns for NodesType::GaussLegendre, Quadrule.cpp:1004-1035).
 * Host-only. */
int tsw_gauss_legendre(int n, double *nodes, double *weights);

/* One member of a batch of independent problems (mesh, DG degree, ordinate set; see tsw_problem_set). */
typedef struct tsw_problem {
    int nx, ny;
    double ax, bx, ay, by;
    int p, nq;
    const double *xi, *eta, *w; /* host arrays of nq entries */
    int quad_off[5];
} tsw_problem;

/* Installs a batch of `nprob` independent problems of uneven size, DG degree and ordinate set, with
 * their total cross sections (sigma_t[b]: [nx_b+2][ny_b+2], host or device per `mem`) and time step.
 * Replaces a loop of SweepOperator constructions + Apply calls over many problem instances
 * (SweepOperator.cpp:25-49,118-135); the batch keeps the cross sections' material tables.
 * Needs ny_b <= 448 for every member. */
int tsw_batch_set(tsw_ctx *ctx, int nprob, const tsw_problem *probs, const double *const *sigma_t, double dt,
                  int mem);

/* Sweeps every member of the batch in place (psi[b] in the reference layout, halos = inflow), in
 * one set of concurrent launches on four streams: per DG degree the engine tsw_batch_set chose (the
 * staggered kernels, one launch per degree and row count over (member, ordinate) items, through layout
 * conversions or on the caller's arrays; or the ordinate-lane engine, one launch per degree over (member,
 * ordinate block, band) items on the caller's arrays).  The engines give the same bits.  SweepOperator::
 * Apply semantics per member.  Timed as one sweep by tsw_sweep_timing (TSW_MIXED on converted members:
 * each of its FP32 batched sweeps is timed as one).  Device arrays that are not 16-byte aligned send the
 * call's staggered groups through the conversions. */
int tsw_batch_sweep(tsw_ctx *ctx, double *const *psi, int precision, int mem);

/* Timing of the sweeps run on ctx since the last reset, measured with HIP events on the context's
 * stream around each sweep: count, total milliseconds.  Replaces the TMR_AF_Lsv wall timer
 * (SweepOperator.cpp:129,134). */
int tsw_sweep_timing(tsw_ctx *ctx, int reset, long long *n_sweeps, double *total_ms);

#ifdef __cplusplus
}
#endif

#endif
