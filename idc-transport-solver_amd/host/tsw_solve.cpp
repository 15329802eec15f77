// tsw_solve DECK: runs a deck on the GPU the way the reference's solver driver does
// (main_solver.cpp:140-420), writing <deck name without extension>.sd2, the scalar flux Pmv(psi) at the
// end (TimeIntegrator.cpp:595-635), and <prefix>.af2, the angular flux, when output_af is true.
//   time_integrator idc: IDC steps of cfl * dx up to tfinal, the last one shortened to land on tfinal
//     (TimeIntegrator.cpp:245-300), from the projected initial condition;
//   no time_integrator key: one steady solve (dt = infinity) with the deck's solve_type.
// Other time integrators of the reference (dirk, erk, stdg, hybrid methods) are not available.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <stdexcept>

#include "tsw_host.hpp"

using namespace tsw_host;

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
            const double TOL = 1.0e-14, T = in.final_time, h = in.max_step_size;
            const int64_t steps = (int64_t)(T / h) + (std::fmod(T, h) > TOL ? 1 : 0);
            IDCIntegrator idc(dev, params);
            for (int64_t s = 0; s < steps; ++s) {
                const double dt = (s + 1) * h > T + TOL ? T - s * h : h;
                idc.Step(psi, in.source, in.sigma_t, in.sigma_s, dt);
            }
            std::printf("%s: %lld IDC steps to t = %.6e\n", in.problem.c_str(), (long long)steps, T);
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
