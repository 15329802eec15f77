// tsw_solve DECK: runs a deck on the GPU the way the reference's solver driver does
// (main_solver.cpp:140-420), writing <deck name without extension>.sd2, the scalar flux Pmv(psi) at the
// end (TimeIntegrator.cpp:595-635), and <prefix>.af2, the angular flux, when output_af is true.
//   time_integrator idc: IDC steps of cfl * dx up to tfinal, the last one shortened to land on tfinal
//     (TimeIntegrator.cpp:245-300), from the projected initial condition; with checkpoint_steps or
//     checkpoint_walltime it resumes from <prefix>.chk when that file fits, writes <prefix>.<step>.chk and
//     exits with status 150 when a checkpoint is due or on SIGINT / SIGTERM (TimeIntegrator.cpp:67-103,
//     :325-433, global.hpp:179), as the reference's driver does;
//   no time_integrator key: one steady solve (dt = infinity) with the deck's solve_type.
// Other time integrators of the reference (dirk, erk, stdg, hybrid methods) are not available.
#include <algorithm>
#include <cmath>
#include <csignal>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <stdexcept>

#include "tsw_host.hpp"

using namespace tsw_host;

namespace {
volatile std::sig_atomic_t g_stop = 0;
extern "C" void on_signal(int) { g_stop = 1; }  // checkpoint after the step in progress
constexpr int kExitCheckpoint = 150;            // ExitCode::Checkpoint (global.hpp:179)
}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s DECK\n", argv[0]);
        return 2;
    }
    try {
        std::string prefix = argv[1];
        const size_t dot = prefix.rfind('.');
        if (dot != std::string::npos && prefix.find('/', dot) == std::string::npos) prefix.erase(dot);
        ParameterList params;
        params.ReadFromFile(argv[1]);
        ProblemInputs in = Init(params, ".");
        auto dev = std::make_shared<Device>(in.dd, in.os, in.DG_degree);
        OrdinateFlux psi = in.psi;
        const std::string ti = params.Has("time_integrator") ? params.GetString("time_integrator") : "";
        if (ti == "idc") {
            if (!(in.final_time > 0) || !(in.max_step_size > 0))
                throw std::invalid_argument("idc needs tfinal and cfl");
            IDCIntegrator idc(dev, params);
            if ((params.Has("checkpoint_steps") && params.GetInt("checkpoint_steps") > 0) ||
                (params.Has("checkpoint_walltime") && params.GetDouble("checkpoint_walltime") > 0.0)) {
                std::signal(SIGINT, on_signal);
                std::signal(SIGTERM, on_signal);
            }
            // a line per step, as the reference's PRINT_LOG in Integrate
            if (idc.Integrate(psi, in.source, in.sigma_t, in.sigma_s, in.final_time, in.max_step_size, prefix, &g_stop,
                              stdout) == IDCIntegrator::Outcome::Checkpointed) {
                std::printf("%s: checkpoint %s after %lld IDC steps, t = %.6e\n", in.problem.c_str(),
                            idc.last_checkpoint().c_str(), (long long)idc.state().step, idc.state().time);
                std::fflush(stdout);
                return kExitCheckpoint;
            }
            std::printf("%s: %lld IDC steps to t = %.6e\n", in.problem.c_str(), (long long)idc.state().step,
                        in.final_time);
        } else if (ti.empty()) {
            auto solver = ImplicitSolver::Create(dev, params);
            std::copy(in.source.data(), in.source.data() + in.source.size(), psi.data());
            solver->Solve(psi, in.sigma_t, in.sigma_s);
            std::printf("%s: steady %s, %lld iterations, error %.3e\n", in.problem.c_str(),
                        solver->GetDescriptor().c_str(), (long long)solver->iterations(), solver->final_error());
        } else {
            throw std::out_of_range("time integrator '" + ti + "' is not available");
        }
        ScalarFlux phi(in.dd, in.DG_degree);
        TransportOperator::Pmv(*dev, 1.0, 0.0, psi, phi);
        phi.WriteToDisk(prefix + ".sd2");
        bool af = false;
        try {
            af = params.GetBool("output_af");
        } catch (...) {
        }
        if (af) psi.WriteToDisk(prefix + ".af2");
    } catch (const std::exception &e) {
        std::fprintf(stderr, "tsw_solve: %s\n", e.what());
        return 1;
    }
    return 0;
}
