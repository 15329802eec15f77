// C++ host layer over the libtsw C-ABI, mirroring the reference's operator and solver interface for
// the sweep path (crockatt/idc-transport-solver):
//   RKDG::OrdinateFlux / ScalarFlux / CrossSection   src/objects/RKDG/*.hpp (layouts, operator())
//   Quadrule::OrdinateSet                             src/utils/Quadrule/OrdinateSet.hpp:236-418
//   DomainDecomposition                               src/objects/DomainDecomposition.hpp
//   ParameterList                                     src/utils/ParameterList.cpp:202-290 (key value)
//   SweepOperator::Apply                              src/linear_solvers/Abstract/SweepOperator.hpp:126-131
//   TransportOperator::Pmv / Smv / Lmv                src/operators/RKDG/TransportOperator.hpp:25-43
//   ImplicitSolver + factory key solve_type           src/linear_solvers/Abstract/ImplicitSolver.hpp:117-123,
//                                                     ImplicitSolver.cpp:182-229
//   ParameterList::ReadFromFile / WriteToFile         src/utils/ParameterList.cpp:202-285 (decks)
//   DensityFunction / OrdinateFlux Read/WriteToDisk   src/objects/RKDG/DensityFunction.cpp:1880-2291,
//                                                     OrdinateFlux.cpp:2720-2885, SpatialDomain.cpp:247-407
//   DensityFunction L1Norm / L2Norm / LinfNorm        src/objects/RKDG/DensityFunction.cpp:1082-1400
//   Problem setup from a deck ("problem" key)          src/utils/init.cpp, src/problems/*Problem.cpp
// Arrays live on the host in the reference's layouts; every call hands them to the GPU library.
// Errors follow the reference: std::invalid_argument for bad arguments, std::out_of_range for unknown
// factory keys, std::runtime_error for failed solves and device errors.
#pragma once

#include <array>
#include <csignal>
#include <cstdio>
#include <cstdint>
#include <limits>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "tsw.h"

namespace tsw_host {

// Plain "key value" parameters (ParameterList.cpp:202-290).
class ParameterList {
  public:
    void Set(const std::string &key, const std::string &value) { values_[key] = value; }
    bool Has(const std::string &key) const { return values_.count(key) != 0; }
    std::string GetString(const std::string &key) const;
    double GetDouble(const std::string &key) const { return std::stod(GetString(key)); }
    int64_t GetInt(const std::string &key) const { return std::stoll(GetString(key)); }
    // "true" / "false" as std::boolalpha reads them (ParameterList.cpp:292-335); other text throws
    // std::runtime_error, a missing key std::out_of_range.
    bool GetBool(const std::string &key) const;
    // Reads "key value" lines; '#' starts a comment.
    static ParameterList FromText(const std::string &text);
    // A deck by ReadFromFile's rules (ParameterList.cpp:202-245): text from '#' on is dropped, the line
    // is lower-cased and its first two words are key and value; keys already present are kept.
    ParameterList &ReadFromFile(const std::string &filename);
    // "key value" lines in key order (ParameterList.cpp:254-285).
    const ParameterList &WriteToFile(const std::string &filename) const;
    const std::map<std::string, std::string> &items() const { return values_; }

  private:
    std::map<std::string, std::string> values_;
};

// Uniform Cartesian mesh [ax,bx] x [ay,by] with nx x ny cells.
struct DomainDecomposition {
    int64_t nx = 0, ny = 0;
    double ax = 0, bx = 1, ay = 0, by = 1;
    // reflecting edges (TSW_REFLECT_X_MIN | ...) or TSW_PERIODIC: the reference's Global::reflecting_boundaries
    // and Global::periodic, which Init sets from the deck keys reflect_x_min / _x_max / _y_min / _y_max and
    // periodic (init.cpp:448-510)
    int reflect_mask = 0;
    double dx(int dim) const { return dim == 0 ? (bx - ax) / nx : (by - ay) / ny; }
};

// Ordinates grouped by quadrant (0:++ 1:+- 2:-+ 3:--), weights summing to 4 pi.
class OrdinateSet {
  public:
    static OrdinateSet LevelSymmetric(int order);
    // From the deck keys ordinate_type and ang_order (OrdinateSet constructor, OrdinateSet.cpp:240-280);
    // "level-symmetric" is built, other types throw std::invalid_argument.
    static OrdinateSet FromParameters(const ParameterList &params);
    // The named sets of tsw_ordinates: "spherical-triangle", "chebyshev-legendre", "level-symmetric".
    static OrdinateSet Named(const std::string &type, int order, bool symmetric_reduce = true);
    OrdinateSet(std::vector<double> xi, std::vector<double> eta, std::vector<double> w,
                std::array<int, 5> quadrants, std::string type = "none", int64_t ang_order = 0,
                bool symmetric_reduce = true);
    const std::string &type() const { return type_; }
    int64_t ang_order() const { return order_; }
    bool symmetric_reduce() const { return reduced_; }
    int64_t nq() const { return (int64_t)xi_.size(); }
    double xi(int64_t q) const { return xi_[q]; }
    double eta(int64_t q) const { return eta_[q]; }
    double w(int64_t q) const { return w_[q]; }
    int64_t Quadrants(int k) const { return quad_[k]; }
    const std::vector<double> &xi() const { return xi_; }
    const std::vector<double> &eta() const { return eta_; }
    const std::vector<double> &w() const { return w_; }
    const std::array<int, 5> &quadrant_offsets() const { return quad_; }

  private:
    std::vector<double> xi_, eta_, w_;
    std::array<int, 5> quad_;
    std::string type_;
    int64_t order_ = 0;
    bool reduced_ = true;
};

// Angular flux, density[((i*(ny+2)+j)*nq+q)*(p+1)^2 + d*(p+1) + e] (OrdinateFlux.hpp:639-700).
class OrdinateFlux {
  public:
    OrdinateFlux(const DomainDecomposition &dd, const OrdinateSet &os, int DG_degree);
    double &operator()(int64_t q, int64_t i, int64_t j, int64_t d, int64_t e);
    double operator()(int64_t q, int64_t i, int64_t j, int64_t d, int64_t e) const;
    double *data() { return density_.data(); }
    const double *data() const { return density_.data(); }
    size_t size() const { return density_.size(); }
    int64_t nq() const { return nq_; }
    const DomainDecomposition &domain() const { return dd_; }
    const OrdinateSet &ordinates() const { return os_; }
    // Binary file of RKDG::OrdinateFlux::WriteToDisk (OrdinateFlux.cpp:2720-2885, one rank): the mesh
    // header, the ordinate-set name with "#R" (the 2D sets are symmetry-reduced) as a null-terminated
    // string, the angular order, then DG degree, byte count, value count and every value with halos.
    void WriteToDisk(const std::string &filename) const;
    void WriteToDisk(std::FILE *fp) const;
    // Reads such a file; the ordinate set is rebuilt from its name and order (level-symmetric only).
    static OrdinateFlux ReadFromDisk(const std::string &filename);
    // Reads such a record from fp into this flux; false (values partly read or untouched) when the record's
    // mesh, ordinate set, degree or size differ from this flux's.
    bool ReadValues(std::FILE *fp);
    const int DG_degree;

  private:
    DomainDecomposition dd_;
    OrdinateSet os_;
    int64_t nx_, ny_, nq_;
    std::vector<double> density_;
};

// Scalar flux / density, [i][j][d][e] with halo (DensityFunction.hpp:475-535).
class ScalarFlux {
  public:
    ScalarFlux(const DomainDecomposition &dd, int DG_degree);
    double &operator()(int64_t i, int64_t j, int64_t d, int64_t e);
    double operator()(int64_t i, int64_t j, int64_t d, int64_t e) const;
    double *data() { return density_.data(); }
    const double *data() const { return density_.data(); }
    size_t size() const { return density_.size(); }
    const DomainDecomposition &domain() const { return dd_; }
    // Binary file of RKDG::DensityFunction::WriteToDisk (DensityFunction.cpp:2226-2291, one rank): per
    // dimension int64 nx, double ax, double bx (SpatialDomain.cpp:385-407), then int64 DG degree, int64
    // byte count, int64 value count and the values, halo cells included.
    void WriteToDisk(const std::string &filename) const;
    static ScalarFlux ReadFromDisk(const std::string &filename);  // throws std::runtime_error
    // Norms over the interior, as DensityFunction::L1Norm (Gauss-Legendre with DG degree + 1 points per
    // axis), L2Norm (Legendre orthogonality) and LinfNorm (per cell, uniform samples doubled until the
    // maximum changes by at most tol relative) compute them (DensityFunction.cpp:1082-1400).
    double L1Norm() const;
    double L2Norm() const;
    double LinfNorm(double tol = 1e-3) const;
    double EvaluateAtPoint(double x, double y) const;  // the cell's polynomial at (x, y)
    // first - second, as DensityFunction::ComputeDifference (DensityFunction.cpp:525-720) forms it: on
    // the common mesh the coefficients are subtracted at the higher of the two DG degrees; otherwise both
    // are projected exactly onto the mesh with lcm(nx) x lcm(ny) cells (each old cell split into equal
    // subcells) and subtracted there.  The domains must agree to 1e-14 (std::invalid_argument).  The
    // reference returns the result through an output argument that it reconfigures; here it is returned.
    static ScalarFlux ComputeDifference(const ScalarFlux &first, const ScalarFlux &second);
    const int DG_degree;

  private:
    DomainDecomposition dd_;
    int64_t ny_;
    std::vector<double> density_;
};

// Cross section, [i][j][u][v] with halo (CrossSection.hpp:397-478): DG degree 0, one value per cell, or
// the flux's DG degree p, (p+1)^2 Legendre coefficients per cell (init.cpp:741 with piecewise_cross
// false).  The operators and solvers pass the degree to the device with tsw_set_xs_degree.
class CrossSection {
  public:
    explicit CrossSection(const DomainDecomposition &dd, double value = 0.0);
    // degree DG_degree, cell means `mean`, higher coefficients zero
    CrossSection(const DomainDecomposition &dd, int DG_degree, double mean);
    double &operator()(int64_t i, int64_t j) { return v_[idx(i, j, 0, 0)]; }  // the cell mean
    double operator()(int64_t i, int64_t j) const { return v_[idx(i, j, 0, 0)]; }
    double &operator()(int64_t i, int64_t j, int64_t u, int64_t v) { return v_[idx(i, j, u, v)]; }
    double operator()(int64_t i, int64_t j, int64_t u, int64_t v) const { return v_[idx(i, j, u, v)]; }
    const double *data() const { return v_.data(); }
    double *data() { return v_.data(); }
    size_t size() const { return v_.size(); }
    const int DG_degree;

  private:
    size_t idx(int64_t i, int64_t j, int64_t u, int64_t v) const {
        const int64_t P = DG_degree + 1;
        return (size_t)(v + P * (u + P * (j + (ny_ + 2) * i)));
    }
    int64_t ny_;
    std::vector<double> v_;
};

// Where an operator applies (DomainDecomposition.hpp:35-62).  Single-rank meshes: every halo cell is
// a boundary cell.
enum class OpDomain { Interior = 0x01, Boundary = 0x02, Halo = 0x04, All = 0x07 };

// One device context shared by the operators and solvers built on a mesh and ordinate set.
class Device {
  public:
    Device(const DomainDecomposition &dd, const OrdinateSet &os, int DG_degree, int device = 0);
    ~Device();
    Device(const Device &) = delete;
    Device &operator=(const Device &) = delete;
    tsw_ctx *ctx() const { return ctx_; }
    void Check(int status, const char *what) const;  // throws on status != TSW_OK

  private:
    tsw_ctx *ctx_ = nullptr;
};

// In-place sweep psi <- L^{-1} psi, halos hold the inflow (SweepOperator.cpp:118-135).
// Parameters: sweep_pattern ("hip-staggered" default, "hip-wavefront", ...), sweep_precision
// ("fp64" default, "fp32").
class SweepOperator {
  public:
    SweepOperator(std::shared_ptr<Device> dev, const ParameterList &params,
                  double dt = std::numeric_limits<double>::infinity());
    void Apply(OrdinateFlux &source, const CrossSection &sigma, double dt = std::numeric_limits<double>::infinity(),
               const OrdinateFlux *initial_condition = nullptr) const;

  private:
    std::shared_ptr<Device> dev_;
    int precision_ = TSW_FP64;
};

namespace TransportOperator {
void Pmv(Device &dev, double alpha, double beta, const OrdinateFlux &x, ScalarFlux &y,
         OpDomain op_domain = OpDomain::All);
void Smv(Device &dev, double alpha, double beta, const CrossSection &sigma, const ScalarFlux &x, OrdinateFlux &y,
         OpDomain op_domain = OpDomain::All);
void Lmv(Device &dev, double alpha, double beta, const CrossSection &sigma, const OrdinateFlux &x, OrdinateFlux &y);
}  // namespace TransportOperator

// ImplicitSolver with the reference's Solve contract: source holds Q on entry and psi on exit.
class ImplicitSolver {
  public:
    virtual ~ImplicitSolver() = default;
    virtual void Solve(OrdinateFlux &source, const CrossSection &sigma_t, const CrossSection &sigma_s,
                       double dt = std::numeric_limits<double>::infinity(),
                       const OrdinateFlux *initial = nullptr) = 0;
    virtual std::string GetDescriptor() const = 0;
    int64_t iterations() const { return iterations_; }
    double final_error() const { return final_error_; }

    // Factory on the "solve_type" key (ImplicitSolver.cpp:182-229): "source-iteration",
    // "peierls-gmres" (also accepted under the reference's "petsc-peierls-gmres").
    static std::unique_ptr<ImplicitSolver> Create(std::shared_ptr<Device> dev, const ParameterList &params,
                                                  double dt = std::numeric_limits<double>::infinity());

  protected:
    int64_t iterations_ = 0;
    double final_error_ = 0.0;
};

// Source iteration (SISolver.cpp:190-254); parameters abs_tol (1e-8), max_its (500), sweep_precision.
class SISolver : public ImplicitSolver {
  public:
    SISolver(std::shared_ptr<Device> dev, const ParameterList &params);
    void Solve(OrdinateFlux &source, const CrossSection &sigma_t, const CrossSection &sigma_s, double dt,
               const OrdinateFlux *initial) override;
    std::string GetDescriptor() const override { return "source-iteration"; }

  private:
    std::shared_ptr<Device> dev_;
    double abs_tol_ = 1e-8;
    int64_t max_its_ = 500;
    int precision_ = TSW_FP64;
};

// Sweep-preconditioned GMRES on the Peierls form (PETScPeierlsSolver.cpp:423-572); parameters
// abs_tol (1e-8), max_its (3000), gmres_restart (30), sweep_precision ("fp32" selects FP32 sweeps
// with FP64 refinement), use_dsa (false; PETScPeierlsSolver.cpp:151) and dsa_cycles (2).
class PeierlsGMRESSolver : public ImplicitSolver {
  public:
    PeierlsGMRESSolver(std::shared_ptr<Device> dev, const ParameterList &params);
    void Solve(OrdinateFlux &source, const CrossSection &sigma_t, const CrossSection &sigma_s, double dt,
               const OrdinateFlux *initial) override;
    std::string GetDescriptor() const override { return "peierls-gmres"; }

  private:
    std::shared_ptr<Device> dev_;
    double abs_tol_ = 1e-8;
    int64_t max_its_ = 3000;
    int64_t restart_ = 30;
    int precision_ = TSW_FP64;
};

// IDC time integrator, the non-hybrid IDCIntegrator (time_integrators/RKDG/idc.cpp:846-940) over the same
// device.  Parameters: idc_num_nodes (4), idc_num_corr (3), idc_type ("error" default, "update"),
// idc_nodes_type ("gauss-lobatto" default, "gauss-legendre", "gauss-radau", "chebyshev",
// "chebyshev-radau"), idc_use_collocation (false), abs_tol (1e-10) and max_its
// (3000) of the implicit solves, sweep_precision.  Step advances initial_condition by dt in place
// (halos zero), with the material source and boundary inflow in `source`, constant over the step.
// Checkpoint of a time integration (TimeIntegrator::WriteCheckpoint / ResumeCheckpoint,
// TimeIntegrator.cpp:745-1111): completed steps, time reached, accumulated solve time and the reference's
// 20 timer slots (this build fills the first, the sweep time TMR_AF_Lsv, in seconds).
struct CheckpointState {
    int64_t step = 0;
    double time = 0.0, solve_time = 0.0;
    std::array<double, 20> timers{};
};
// Writes `filename`; throws std::runtime_error when it cannot be written.
void WriteCheckpoint(const std::string &filename, const CheckpointState &st, const OrdinateFlux &psi);
// Reads `filename` into psi (same mesh, ordinates and degree); throws std::runtime_error, leaving psi as it
// was, when the file is missing, short or written for another problem.
CheckpointState ResumeCheckpoint(const std::string &filename, OrdinateFlux &psi);

class IDCIntegrator {
  public:
    IDCIntegrator(std::shared_ptr<Device> dev, const ParameterList &params);
    void Step(OrdinateFlux &initial_condition, const OrdinateFlux &source, const CrossSection &sigma_t,
              const CrossSection &sigma_s, double dt);
    std::string GetDescriptor() const { return "idc"; }

    // TimeIntegrator::Integrate (TimeIntegrator.cpp:325-433): steps of max_step_size up to final_time, the
    // last one shortened to land on it.  With checkpoint_steps or checkpoint_walltime (deck keys of the same
    // names) it first resumes from `<prefix>.chk` when that file fits (a warning otherwise), and after a step
    // writes `<prefix>.<step as %04X>.chk` and returns Checkpointed when: the step count is a multiple of
    // checkpoint_steps (and steps remain), the walltime estimate says the next step would pass
    // checkpoint_walltime seconds of solve time (CheckWalltimeCheckpoint :1141-1199), or *stop is set (the
    // driver's SIGINT/SIGTERM flag, CheckpointSignalHandler :67-103).  `log`, when set, receives a line per
    // step.  The reference's driver exits with status 150 after a checkpoint (global.hpp:179).
    enum class Outcome { Done, Checkpointed };
    Outcome Integrate(OrdinateFlux &psi, const OrdinateFlux &source, const CrossSection &sigma_t,
                      const CrossSection &sigma_s, double final_time, double max_step_size,
                      const std::string &prefix, const volatile std::sig_atomic_t *stop = nullptr,
                      std::FILE *log = nullptr);
    const CheckpointState &state() const { return state_; }
    std::string last_checkpoint() const { return last_checkpoint_; }

  private:
    std::shared_ptr<Device> dev_;
    double abs_tol_ = 1e-10;
    int64_t max_its_ = 3000;
    int precision_ = TSW_FP64;
    int64_t checkpoint_steps_ = 0;
    double checkpoint_walltime_ = 0.0;
    CheckpointState state_;
    std::string last_checkpoint_;
};

// ---------------------------------------------------------------------------------------------------
// Problem definitions selected by the deck key "problem" (Factory<Problem>, src/utils/Problems/*):
// functions of position for the initial condition, isotropic source and cross sections, and the inflow
// written into the source's halo cells.  Built: "lattice" (LatticeProblem.cpp:100-324), "hohlraum"
// (HohlraumProblem.cpp:60-272), "hohlraum2" (Hohlraum2Problem.cpp:14-148), "bump" (BumpProblem.cpp,
// keys sigma_val, bump_radius) and "inhomogeneous-sphere" (InhomogeneousSphereProblem.cpp, keys
// sphere_inner/outer_radius, sphere_inner/outer_trans_radius, sphere_inner/interm/outer_sigma_val);
// other names (Reed's problem is 1D) throw std::out_of_range.  The deck key moll_radius is required, as Problem's constructor requires it
// (Problem.cpp:20-35); mollified evaluation (moll_radius >= 1e-14) is not built and throws
// std::invalid_argument.
class Problem {
  public:
    virtual ~Problem() = default;
    virtual std::string GetDescriptor() const = 0;
    virtual double InitialCondition(double, double) const { return 0.0; }
    virtual double Source(double, double) const { return 0.0; }
    virtual double TotalCross(double, double) const { return 0.0; }
    virtual double ScatterCross(double, double) const { return 0.0; }
    // Writes the boundary inflow density into halo cells of the projected source (Problem.cpp:105-116).
    virtual void SetInflowBoundaries(ScalarFlux &) const {}
    // Replaces deck values the problem fixes (the domain), where the deck names them.
    virtual void OverrideOptions(ParameterList &) const {}
    static std::unique_ptr<Problem> Create(const ParameterList &params);
};
using ProblemEvalFcn = double (Problem::*)(double, double) const;

// L2 projection of a problem function onto the DG space of dest, ComputeProjection of init.cpp:50-240:
// proj_order x proj_order Gauss-Legendre points per interior cell, then, if apply_limiter and the cell
// mean is not negative, the moments of degree > 0 are scaled by the largest theta in [0, 1] (bisection
// to 1e-14) that keeps the polynomial non-negative on a 21 x 21 grid of the cell.
void ComputeProjection(const Problem &prob, ScalarFlux &dest, ProblemEvalFcn f, bool apply_limiter = true,
                       int proj_order = 14);

// Everything Init (init.cpp:422-946) builds from a deck, for a solve on the GPU.
struct ProblemInputs {
    ProblemInputs(const DomainDecomposition &dd, const OrdinateSet &os, int DG_degree, int sigma_degree = 0);
    DomainDecomposition dd;
    OrdinateSet os;
    int DG_degree;
    ScalarFlux sigma_t_density, sigma_s_density;  // the cross-section projections (degree 0 or DG_degree)
    ScalarFlux initial_density, source_density;   // projected initial condition and source (+ inflow)
    CrossSection sigma_t, sigma_s;                // the same cross sections
    OrdinateFlux psi;                             // initial_density / (4 pi) in every ordinate
    OrdinateFlux source;                          // source_density / (4 pi), halos included
    std::string problem;
    double final_time;     // "tfinal", NaN when absent
    double max_step_size;  // "cfl" * dx (TimeIntegrator.cpp:129-130), NaN when absent
};

// Builds the inputs of a deck as Init does: the problem (key "problem") overrides its domain; the mesh
// from nx, ny, ax, bx, ay, by; the DG degree from dg_degree (or dg_degree_x, global.cpp:198-215); the
// ordinate set from ordinate_type (default "chebyshev-legendre"), ang_order and ordinate_sym_reduce
// (default false).  The cross sections are projected at DG degree 0 when piecewise_cross is true and at
// the flux's degree otherwise (init.cpp:741).  Reflecting edges and periodic boundaries go to
// dd.reflect_mask (a Device on that decomposition sets them); both at once throw std::runtime_error.  With output_init true each projection is read from, or else computed and
// written to, the file the reference names (sigma_t_dg<0 or p>_nx0000224.cs2, sigma_s_dg..., 
// initial_dg<p>_..., source_dg<p>_....sd2) in `directory`.
ProblemInputs Init(ParameterList &params, const std::string &directory = ".");

}  // namespace tsw_host
