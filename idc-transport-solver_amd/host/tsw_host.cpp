// C++ host layer over the libtsw C-ABI; see tsw_host.hpp for the reference interfaces mirrored.
#include "tsw_host.hpp"

#include <array>
#include <chrono>
#include <cmath>
#include <map>
#include <sstream>
#include <stdexcept>

namespace tsw_host {

// ---------------------------------------------------------------------------------------------------
std::string ParameterList::GetString(const std::string &key) const {
    auto it = values_.find(key);
    if (it == values_.end()) throw std::out_of_range("parameter '" + key + "' not present");
    return it->second;
}

ParameterList ParameterList::FromText(const std::string &text) {
    ParameterList p;
    std::istringstream in(text);
    std::string line;
    while (std::getline(in, line)) {
        const auto hash = line.find('#');
        if (hash != std::string::npos) line = line.substr(0, hash);
        std::istringstream ls(line);
        std::string key, value;
        if (ls >> key >> value) p.Set(key, value);
    }
    return p;
}

// ---------------------------------------------------------------------------------------------------
OrdinateSet OrdinateSet::LevelSymmetric(int order) {
    if (order <= 0 || order % 2) throw std::invalid_argument("level-symmetric order must be even and positive");
    const int nq = order * (order + 2) / 2;
    std::vector<double> xi(nq), eta(nq), w(nq);
    std::array<int, 5> off{};
    if (tsw_ordinates_level_symmetric(order, xi.data(), eta.data(), w.data(), off.data()) != TSW_OK)
        throw std::invalid_argument("no level-symmetric set of order " + std::to_string(order));
    return OrdinateSet(xi, eta, w, off, "level-symmetric", order, true);
}

OrdinateSet OrdinateSet::Named(const std::string &type, int order, bool symmetric_reduce) {
    int nq = 0;
    if (tsw_ordinates(type.c_str(), order, symmetric_reduce, 0, &nq, nullptr, nullptr, nullptr, nullptr) != TSW_OK)
        throw std::invalid_argument("no ordinate set '" + type + "' of order " + std::to_string(order));
    std::vector<double> xi(nq), eta(nq), w(nq);
    std::array<int, 5> off{};
    if (tsw_ordinates(type.c_str(), order, symmetric_reduce, nq, &nq, xi.data(), eta.data(), w.data(), off.data()) !=
        TSW_OK)
        throw std::invalid_argument("no ordinate set '" + type + "' of order " + std::to_string(order));
    return OrdinateSet(xi, eta, w, off, type, order, type == "level-symmetric" ? true : symmetric_reduce);
}

OrdinateSet::OrdinateSet(std::vector<double> xi, std::vector<double> eta, std::vector<double> w,
                         std::array<int, 5> quadrants, std::string type, int64_t ang_order, bool symmetric_reduce)
    : xi_(std::move(xi)), eta_(std::move(eta)), w_(std::move(w)), quad_(quadrants), type_(std::move(type)),
      order_(ang_order), reduced_(symmetric_reduce) {
    if (xi_.size() != eta_.size() || xi_.size() != w_.size() || quad_[0] != 0 || quad_[4] != (int)xi_.size())
        throw std::invalid_argument("inconsistent ordinate set");
}

// ---------------------------------------------------------------------------------------------------
OrdinateFlux::OrdinateFlux(const DomainDecomposition &dd, const OrdinateSet &os, int degree)
    : DG_degree(degree), dd_(dd), os_(os), nx_(dd.nx), ny_(dd.ny), nq_(os.nq()),
      density_((size_t)(dd.nx + 2) * (dd.ny + 2) * os.nq() * (degree + 1) * (degree + 1), 0.0) {}

double &OrdinateFlux::operator()(int64_t q, int64_t i, int64_t j, int64_t d, int64_t e) {
    const int64_t P = DG_degree + 1;
    return density_[(size_t)(e + P * (d + P * (q + nq_ * (j + (ny_ + 2) * i))))];
}
double OrdinateFlux::operator()(int64_t q, int64_t i, int64_t j, int64_t d, int64_t e) const {
    return const_cast<OrdinateFlux *>(this)->operator()(q, i, j, d, e);
}

ScalarFlux::ScalarFlux(const DomainDecomposition &dd, int degree)
    : DG_degree(degree), dd_(dd), ny_(dd.ny), density_((size_t)(dd.nx + 2) * (dd.ny + 2) * (degree + 1) * (degree + 1), 0.0) {}

double &ScalarFlux::operator()(int64_t i, int64_t j, int64_t d, int64_t e) {
    const int64_t P = DG_degree + 1;
    return density_[(size_t)(e + P * (d + P * (j + (ny_ + 2) * i)))];
}
double ScalarFlux::operator()(int64_t i, int64_t j, int64_t d, int64_t e) const {
    return const_cast<ScalarFlux *>(this)->operator()(i, j, d, e);
}

CrossSection::CrossSection(const DomainDecomposition &dd, double value)
    : DG_degree(0), ny_(dd.ny), v_((size_t)(dd.nx + 2) * (dd.ny + 2), value) {}

CrossSection::CrossSection(const DomainDecomposition &dd, int degree, double mean)
    : DG_degree(degree), ny_(dd.ny), v_((size_t)(dd.nx + 2) * (dd.ny + 2) * (degree + 1) * (degree + 1), 0.0) {
    if (degree < 0) throw std::invalid_argument("negative cross-section DG degree");
    for (int64_t i = 0; i <= dd.nx + 1; ++i)
        for (int64_t j = 0; j <= dd.ny + 1; ++j) (*this)(i, j) = mean;
}

// ---------------------------------------------------------------------------------------------------
Device::Device(const DomainDecomposition &dd, const OrdinateSet &os, int degree, int device) {
    if (tsw_create(device, &ctx_) != TSW_OK) throw std::runtime_error("tsw_create failed: no usable HIP device");
    try {
        const auto &off = os.quadrant_offsets();
        Check(tsw_problem_set(ctx_, (int)dd.nx, (int)dd.ny, dd.ax, dd.bx, dd.ay, dd.by, degree, (int)os.nq(),
                              os.xi().data(), os.eta().data(), os.w().data(), off.data()),
              "tsw_problem_set");
        if (dd.reflect_mask) Check(tsw_set_boundaries(ctx_, dd.reflect_mask, nullptr), "tsw_set_boundaries");
    } catch (...) {
        tsw_destroy(ctx_);
        throw;
    }
}

Device::~Device() { tsw_destroy(ctx_); }

void Device::Check(int status, const char *what) const {
    if (status == TSW_OK) return;
    const std::string msg = std::string(what) + ": " + tsw_last_error(ctx_);
    if (status == TSW_E_ARG) throw std::invalid_argument(msg);
    throw std::runtime_error(msg);
}

namespace {
int precision_from(const ParameterList &p) {
    if (!p.Has("sweep_precision")) return TSW_FP64;
    const std::string v = p.GetString("sweep_precision");
    if (v == "fp64") return TSW_FP64;
    if (v == "fp32") return TSW_FP32;
    if (v == "mixed") return TSW_MIXED;  // FP32 sweeps refined to FP64 accuracy
    throw std::invalid_argument("sweep_precision must be fp64, fp32 or mixed, not '" + v + "'");
}
int all_cells(OpDomain d) { return ((int)d & (int)OpDomain::Boundary) ? 1 : 0; }
// The device takes cross sections of the degree they carry (0 or the flux degree); both of a solve share it.
void set_xs(Device &dev, const CrossSection &a, const CrossSection *b = nullptr) {
    if (b && b->DG_degree != a.DG_degree)
        throw std::invalid_argument("sigma_t and sigma_s must have the same DG degree");
    dev.Check(tsw_set_xs_degree(dev.ctx(), a.DG_degree), "cross-section DG degree");
}
bool flag_from(const ParameterList &p, const std::string &key, bool dflt) {
    if (!p.Has(key)) return dflt;
    const std::string v = p.GetString(key);
    if (v == "true" || v == "1" || v == "yes") return true;
    if (v == "false" || v == "0" || v == "no") return false;
    throw std::invalid_argument(key + " must be true or false, not '" + v + "'");
}
// The reference's sweep_pattern names (SweepPattern*.cpp: "sequential", "wavefront", "wavefront-mpi",
// "kba", "kba2") choose a traversal order of the same upwind solves; on the device they all run the
// staggered wavefront kernel, whose results equal theirs up to rounding.  Device names pass through.
std::string device_pattern(const std::string &name) {
    if (name == "sequential" || name == "wavefront" || name == "wavefront-mpi" || name == "kba" || name == "kba2")
        return "hip-staggered";
    return name;
}
}  // namespace

SweepOperator::SweepOperator(std::shared_ptr<Device> dev, const ParameterList &params, double)
    : dev_(std::move(dev)), precision_(precision_from(params)) {
    const std::string pattern = params.Has("sweep_pattern") ? params.GetString("sweep_pattern") : "hip-staggered";
    const int rc = tsw_set_pattern(dev_->ctx(), device_pattern(pattern).c_str());
    if (rc == TSW_E_ARG) throw std::out_of_range("no sweep_pattern '" + pattern + "'");
    dev_->Check(rc, "tsw_set_pattern");
}

void SweepOperator::Apply(OrdinateFlux &source, const CrossSection &sigma, double dt, const OrdinateFlux *) const {
    set_xs(*dev_, sigma);
    dev_->Check(tsw_sweep_apply(dev_->ctx(), source.data(), sigma.data(), dt, precision_, TSW_MEM_HOST),
                "SweepOperator::Apply");
}

namespace TransportOperator {
void Pmv(Device &dev, double alpha, double beta, const OrdinateFlux &x, ScalarFlux &y, OpDomain d) {
    dev.Check(tsw_pmv(dev.ctx(), alpha, beta, x.data(), y.data(), all_cells(d), TSW_MEM_HOST), "Pmv");
}
void Smv(Device &dev, double alpha, double beta, const CrossSection &sigma, const ScalarFlux &x, OrdinateFlux &y,
         OpDomain d) {
    set_xs(dev, sigma);
    dev.Check(tsw_smv(dev.ctx(), alpha, beta, sigma.data(), x.data(), y.data(), all_cells(d), TSW_MEM_HOST), "Smv");
}
void Lmv(Device &dev, double alpha, double beta, const CrossSection &sigma, const OrdinateFlux &x, OrdinateFlux &y) {
    set_xs(dev, sigma);
    dev.Check(tsw_lmv(dev.ctx(), alpha, beta, sigma.data(), x.data(), y.data(), TSW_MEM_HOST), "Lmv");
}
}  // namespace TransportOperator

// ---------------------------------------------------------------------------------------------------
std::unique_ptr<ImplicitSolver> ImplicitSolver::Create(std::shared_ptr<Device> dev, const ParameterList &params,
                                                       double) {
    const std::string key = params.GetString("solve_type");
    if (params.Has("sweep_pattern")) {
        const std::string pattern = params.GetString("sweep_pattern");
        const int rc = tsw_set_pattern(dev->ctx(), device_pattern(pattern).c_str());
        if (rc == TSW_E_ARG) throw std::out_of_range("no sweep_pattern '" + pattern + "'");
        dev->Check(rc, "tsw_set_pattern");
    }
    if (key == "source-iteration") return std::unique_ptr<ImplicitSolver>(new SISolver(dev, params));
    if (key == "peierls-gmres" || key == "petsc-peierls-gmres")
        return std::unique_ptr<ImplicitSolver>(new PeierlsGMRESSolver(dev, params));
    throw std::out_of_range("Failed to access factory for ImplicitSolver '" + key + "'");
}

SISolver::SISolver(std::shared_ptr<Device> dev, const ParameterList &p)
    : dev_(std::move(dev)), precision_(precision_from(p)) {
    if (p.Has("abs_tol")) abs_tol_ = p.GetDouble("abs_tol");
    if (p.Has("max_its")) max_its_ = p.GetInt("max_its");
}

void SISolver::Solve(OrdinateFlux &source, const CrossSection &sigma_t, const CrossSection &sigma_s, double dt,
                     const OrdinateFlux *) {
    int it = 0;
    double err = 0.0;
    set_xs(*dev_, sigma_t, &sigma_s);
    dev_->Check(tsw_si_solve(dev_->ctx(), source.data(), sigma_t.data(), sigma_s.data(), dt, abs_tol_, (int)max_its_,
                             precision_, TSW_MEM_HOST, &it, &err),
                "SISolver::Solve");
    iterations_ = it;
    final_error_ = err;
}

PeierlsGMRESSolver::PeierlsGMRESSolver(std::shared_ptr<Device> dev, const ParameterList &p)
    : dev_(std::move(dev)), precision_(precision_from(p)) {
    if (p.Has("abs_tol")) abs_tol_ = p.GetDouble("abs_tol");
    if (p.Has("max_its")) max_its_ = p.GetInt("max_its");
    if (p.Has("gmres_restart")) restart_ = p.GetInt("gmres_restart");
    dev_->Check(tsw_set_tuning(dev_->ctx(), "use_dsa", flag_from(p, "use_dsa", false) ? 1 : 0), "use_dsa");
    if (p.Has("dsa_cycles")) dev_->Check(tsw_set_tuning(dev_->ctx(), "dsa_cycles", (int)p.GetInt("dsa_cycles")),
                                         "dsa_cycles");
}

void PeierlsGMRESSolver::Solve(OrdinateFlux &source, const CrossSection &sigma_t, const CrossSection &sigma_s,
                               double dt, const OrdinateFlux *) {
    int it = 0;
    double res = 0.0;
    set_xs(*dev_, sigma_t, &sigma_s);
    const int rc = tsw_gmres_solve(dev_->ctx(), source.data(), sigma_t.data(), sigma_s.data(), dt, abs_tol_,
                                   (int)max_its_, (int)restart_, precision_, TSW_MEM_HOST, nullptr, &it, &res);
    iterations_ = it;
    final_error_ = res;
    if (rc == TSW_E_NOCONV) throw std::runtime_error("GMRES transport solve failed to converge");
    dev_->Check(rc, "PeierlsGMRESSolver::Solve");
}

IDCIntegrator::IDCIntegrator(std::shared_ptr<Device> dev, const ParameterList &p)
    : dev_(std::move(dev)), precision_(precision_from(p)) {
    if (p.Has("abs_tol")) abs_tol_ = p.GetDouble("abs_tol");
    if (p.Has("max_its")) max_its_ = p.GetInt("max_its");
    const int nodes = p.Has("idc_num_nodes") ? (int)p.GetInt("idc_num_nodes") : 4;
    const int corr = p.Has("idc_num_corr") ? (int)p.GetInt("idc_num_corr") : 3;
    // the reference's String_to_NodesType (Quadrule.hpp:62-70) and String_to_IDCType (idc.cpp:44-50)
    static const std::map<std::string, int> kNodes = {
        {"gauss-lobatto", TSW_IDC_LOBATTO}, {"gauss-legendre", TSW_IDC_LEGENDRE}, {"gauss-radau", TSW_IDC_RADAU},
        {"chebyshev", TSW_IDC_CHEBYSHEV}, {"chebyshev-radau", TSW_IDC_CHEBYSHEV_RADAU}};
    static const std::map<std::string, int> kTypes = {{"error", TSW_IDC_ERROR}, {"update", TSW_IDC_UPDATE}};
    int type = TSW_IDC_LOBATTO, form = TSW_IDC_ERROR;
    if (p.Has("idc_nodes_type")) {
        const std::string t = p.GetString("idc_nodes_type");
        auto it = kNodes.find(t);
        if (it == kNodes.end()) throw std::out_of_range("idc_nodes_type '" + t + "' is not built");
        type = it->second;
    }
    if (p.Has("idc_type")) {
        const std::string t = p.GetString("idc_type");
        auto it = kTypes.find(t);
        if (it == kTypes.end()) throw std::out_of_range("idc_type '" + t + "' is not built");
        form = it->second;
    }
    // checkpoint settings (TimeIntegrator.cpp:145-159): negative values are taken as zero
    if (p.Has("checkpoint_steps")) checkpoint_steps_ = std::max<int64_t>(0, p.GetInt("checkpoint_steps"));
    if (p.Has("checkpoint_walltime")) checkpoint_walltime_ = std::max(0.0, p.GetDouble("checkpoint_walltime"));
    // idc_use_collocation is read only for node families holding the right end (idc.cpp:225-229); the
    // others always use collocation
    int rc = tsw_idc_set(dev_->ctx(), nodes, corr, type, flag_from(p, "idc_use_collocation", false) ? 1 : 0);
    if (rc == TSW_OK) rc = tsw_idc_set_type(dev_->ctx(), form);
    if (rc == TSW_E_ARG) throw std::invalid_argument(tsw_last_error(dev_->ctx()));
    dev_->Check(rc, "tsw_idc_set");
}

void IDCIntegrator::Step(OrdinateFlux &initial_condition, const OrdinateFlux &source, const CrossSection &sigma_t,
                         const CrossSection &sigma_s, double dt) {
    set_xs(*dev_, sigma_t, &sigma_s);
    const int rc = tsw_idc_step(dev_->ctx(), initial_condition.data(), source.data(), sigma_t.data(), sigma_s.data(),
                                dt, abs_tol_, (int)max_its_, precision_, TSW_MEM_HOST);
    if (rc == TSW_E_NOCONV) throw std::runtime_error("an implicit solve of the IDC step failed to converge");
    dev_->Check(rc, "IDCIntegrator::Step");
}

IDCIntegrator::Outcome IDCIntegrator::Integrate(OrdinateFlux &psi, const OrdinateFlux &source,
                                                const CrossSection &sigma_t, const CrossSection &sigma_s,
                                                double final_time, double max_step_size, const std::string &prefix,
                                                const volatile std::sig_atomic_t *stop, std::FILE *log) {
    if (!(final_time > 0.0) || !(max_step_size > 0.0))
        throw std::invalid_argument("Integrate needs a positive final time and step size");
    const double TOL = 1.0e-14;
    const int64_t num_steps =
        (int64_t)(final_time / max_step_size) + (std::fmod(final_time, max_step_size) > TOL ? 1 : 0);
    const bool checkpointing = checkpoint_steps_ > 0 || checkpoint_walltime_ > 0.0;
    state_ = CheckpointState{};
    last_checkpoint_.clear();
    double wtime_limit = checkpoint_walltime_, last_wtime = 0.0, est_wtime = 0.0;
    bool first_step = true;
    if (checkpointing) {
        try {
            state_ = ResumeCheckpoint(prefix + ".chk", psi);
            if (log) std::fprintf(log, "resumed from %s.chk at step %lld, t = %.4e\n", prefix.c_str(),
                                  (long long)state_.step, state_.time);
        } catch (const std::exception &) {
            std::fprintf(stderr, "Warning: Failed to resume from checkpoint file: Starting from beginning.\n");
            state_ = CheckpointState{};
        }
        // the walltime limit counts from the resumed solve time (ResumeCheckpoint, :1079-1089)
        if (checkpoint_walltime_ > 0.0) {
            last_wtime = state_.solve_time;
            wtime_limit += last_wtime;
        }
    }
    long long n0 = 0;
    double ms0 = 0.0;
    dev_->Check(tsw_sweep_timing(dev_->ctx(), 0, &n0, &ms0), "tsw_sweep_timing");
    while (state_.step < num_steps) {
        double h = max_step_size;
        if ((state_.step + 1) * max_step_size > final_time + TOL) h = final_time - state_.step * max_step_size;
        const auto t0 = std::chrono::steady_clock::now();
        Step(psi, source, sigma_t, sigma_s, h);
        state_.solve_time += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        state_.time += h;
        ++state_.step;
        long long n1 = 0;
        double ms1 = 0.0;
        dev_->Check(tsw_sweep_timing(dev_->ctx(), 0, &n1, &ms1), "tsw_sweep_timing");
        state_.timers[0] += (ms1 - ms0) * 1e-3;  // TMR_AF_Lsv: time in sweeps
        ms0 = ms1;
        if (log) {
            std::fprintf(log, "% 4d%%    t = % .4e    dt = % .4e\n", (int)(100 * state_.step / num_steps), state_.time, h);
            std::fflush(log);
        }
        if (state_.step >= num_steps) break;
        bool write = checkpoint_steps_ > 0 && state_.step % checkpoint_steps_ == 0;
        if (checkpoint_walltime_ > 0.0) {  // CheckWalltimeCheckpoint: exponential average of the step time
            const double step_wtime = state_.solve_time - last_wtime;
            est_wtime = first_step ? step_wtime : 0.3 * step_wtime + 0.7 * est_wtime;
            first_step = false;
            last_wtime = state_.solve_time;
            const double left = wtime_limit - last_wtime;
            if (left <= 0.0 || (int64_t)(left / est_wtime) <= 0) write = true;
        }
        if (checkpointing && stop && *stop) write = true;
        if (write) {
            char tag[32];
            std::snprintf(tag, sizeof tag, "%04llX", (unsigned long long)state_.step);
            last_checkpoint_ = prefix + "." + tag + ".chk";
            WriteCheckpoint(last_checkpoint_, state_, psi);
            return Outcome::Checkpointed;
        }
    }
    if (state_.step == num_steps) state_.time = final_time;
    return Outcome::Done;
}

}  // namespace tsw_host
