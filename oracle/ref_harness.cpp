/*
 * ref_harness.cpp -- a C entry layer over the reference's own sources (crockatt/idc-transport-solver,
 * compiled unmodified from their place in the reference tree by oracle/build_ref.sh into
 * oracle/_ref/libref.so).
 *
 * TEST INFRASTRUCTURE ONLY: it lets tests/golden/make_ref_fixtures.py drive the reference's own Init,
 * OrdinateSet, SweepOperator and SISolver to write fixtures, and lets tests compare the oracle with the
 * reference directly where the library is present.  The product library never loads it.
 *
 * The functions stand in for the reference's driver (main_solver.cpp:263-331): a deck is read into
 * Global::input_list, Init builds psi, source and the cross sections, and single operator applications
 * and solves are run on those objects.  Arrays cross this layer in the reference's host layouts
 * (psi [i][j][q][d][e], cross sections [i][j][u][v] with (sigma degree + 1)^2 coefficients per cell).
 */
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <memory>
#include <string>

#include "linear_solvers/Abstract/ImplicitSolver.hpp"
#include "linear_solvers/Abstract/SweepOperator.hpp"
#include "objects/RKDG/CrossSection.hpp"
#include "objects/RKDG/OrdinateFlux.hpp"
#include "time_integrators/RKDG/idc.hpp"
#include "utils/CLog.hpp"
#include "utils/global.hpp"
#include "utils/init.hpp"
#include "utils/ParameterList.hpp"
#include "utils/Quadrule/OrdinateSet.hpp"

namespace {

struct State {
    DomainDecomposition dd;
    RKDG::OrdinateFlux *psi = nullptr, *source = nullptr;
    RKDG::CrossSection sigma_t, sigma_s;
};
State *g = nullptr;
std::string g_err;

template <typename F>
int guarded(F &&f) {
    try {
        f();
        return 0;
    } catch (const std::exception &e) {
        g_err = e.what();
    } catch (...) {
        g_err = "unknown exception";
    }
    return -1;
}

RKDG::OrdinateFlux &flux(int which) { return which == 0 ? *g->psi : *g->source; }
RKDG::CrossSection &xs(int which) { return which == 0 ? g->sigma_t : g->sigma_s; }

}  // namespace

extern "C" {

const char *ref_last_error(void) { return g_err.c_str(); }

/* Reads `deck` (ParameterList::ReadFromFile) and runs Init (init.cpp:422-946), as main_solver.cpp does
 * before integrating.  Replaces the state of an earlier call, boundary conditions included. */
int ref_init(const char *deck) {
    return guarded([&] {
        delete g;
        g = new State();
        Global::input_list = ParameterList();
        Global::input_list.ReadFromFile(deck);
        // Init only adds the deck's boundary conditions to these globals (init.cpp:448-477)
        Global::reflecting_boundaries = BoundaryEdge::None;
        Global::periodic = false;
        Global::DG_degree = GetDGDegreeX(Global::input_list);
        g->psi = new RKDG::OrdinateFlux();
        g->source = new RKDG::OrdinateFlux();
        Init(*g->psi, *g->source, g->sigma_t, g->sigma_s, g->dd);
    });
}

/* dims[0..4] = nx, ny, DG degree, nq, cross-section DG degree; box[0..3] = ax, bx, ay, by */
int ref_dims(int64_t *dims, double *box) {
    return guarded([&] {
        const RKDG::OrdinateFlux &p = *g->psi;
        dims[0] = p.nx(0);
        dims[1] = p.nx(1);
        dims[2] = p.DG_degree;
        dims[3] = p.nq();
        dims[4] = g->sigma_t.DG_degree;
        box[0] = g->dd.ax(0);
        box[1] = g->dd.bx(0);
        box[2] = g->dd.ax(1);
        box[3] = g->dd.bx(1);
    });
}

/* The ordinate set of the state's psi: xi, eta, w (nq each) and the quadrant starts quad[0..4]
 * (OrdinateSet accessors, OrdinateSet.hpp:236-418). */
int ref_ordinates(double *xi, double *eta, double *w, int64_t *quad) {
    return guarded([&] {
        const RKDG::OrdinateFlux &p = *g->psi;
        for (int64_t q = 0; q < p.nq(); ++q) {
            xi[q] = p.xi(q);
            eta[q] = p.eta(q);
            w[q] = p.w(q);
        }
        for (int k = 0; k < 5; ++k) quad[k] = p.Quadrants(k);
    });
}

/* A stand-alone OrdinateSet (OrdinateSet.cpp:126-287) of `type`, order and reduction: nq, and the
 * arrays when xi is not null (capacity entries at most). */
int ref_ordinate_set(const char *type, int64_t order, int reduce, int64_t capacity, int64_t *nq, double *xi,
                     double *eta, double *w, int64_t *quad) {
    return guarded([&] {
        const Quadrule::OrdinateSet os(order, reduce != 0, Quadrule::OrdinateSet::String_to_OrdinateType.at(type));
        *nq = os.nq();
        if (!xi) return;
        if (os.nq() > capacity) throw std::length_error("ordinate arrays too small");
        for (int64_t q = 0; q < os.nq(); ++q) {
            xi[q] = os.xi(q);
            eta[q] = os.eta(q);
            w[q] = os.w(q);
        }
        for (int k = 0; k < 5; ++k) quad[k] = os.Quadrants(k);
    });
}

/* Angular flux `which` (0 psi, 1 source) to / from a host array in the reference layout. */
int ref_get_flux(int which, double *out) {
    return guarded([&] {
        RKDG::OrdinateFlux &f = flux(which);
        const int64_t P = f.DG_degree + 1;
        size_t k = 0;
        for (int64_t i = 0; i <= f.nx(0) + 1; ++i)
            for (int64_t j = 0; j <= f.nx(1) + 1; ++j)
                for (int64_t q = 0; q < f.nq(); ++q)
                    for (int64_t d = 0; d < P; ++d)
                        for (int64_t e = 0; e < P; ++e) out[k++] = f(q, i, j, d, e);
    });
}
int ref_set_flux(int which, const double *in) {
    return guarded([&] {
        RKDG::OrdinateFlux &f = flux(which);
        const int64_t P = f.DG_degree + 1;
        size_t k = 0;
        for (int64_t i = 0; i <= f.nx(0) + 1; ++i)
            for (int64_t j = 0; j <= f.nx(1) + 1; ++j)
                for (int64_t q = 0; q < f.nq(); ++q)
                    for (int64_t d = 0; d < P; ++d)
                        for (int64_t e = 0; e < P; ++e) f(q, i, j, d, e) = in[k++];
    });
}

/* Cross section `which` (0 sigma_t, 1 sigma_s): its Legendre coefficients [i][j][u][v]; ref_set_xs also
 * recomputes the triple-product tensor (CrossSection::ComputeTensor), as Init does after projecting. */
int ref_get_xs(int which, double *out) {
    return guarded([&] {
        RKDG::CrossSection &s = xs(which);
        const int64_t P = s.DG_degree + 1;
        size_t k = 0;
        for (int64_t i = 0; i <= s.nx(0) + 1; ++i)
            for (int64_t j = 0; j <= s.nx(1) + 1; ++j)
                for (int64_t u = 0; u < P; ++u)
                    for (int64_t v = 0; v < P; ++v) out[k++] = s(i, j, u, v);
    });
}
int ref_set_xs(int which, const double *in) {
    return guarded([&] {
        RKDG::CrossSection &s = xs(which);
        const int64_t P = s.DG_degree + 1;
        size_t k = 0;
        for (int64_t i = 0; i <= s.nx(0) + 1; ++i)
            for (int64_t j = 0; j <= s.nx(1) + 1; ++j)
                for (int64_t u = 0; u < P; ++u)
                    for (int64_t v = 0; v < P; ++v) s(i, j, u, v) = in[k++];
        s.ComputeTensor();
    });
}

/* source <- L^{-1} source in place: SweepOperator constructed from the deck (sweep_pattern, sweep_solve)
 * and applied with sigma_t and dt (SweepOperator.cpp:118-135). */
int ref_sweep(double dt) {
    return guarded([&] {
        SweepOperator<RKDG::OrdinateFlux> op(g->dd, *g->source, Global::input_list, dt);
        op.Apply(*g->source, g->sigma_t, dt);
    });
}

/* ImplicitSolver of the deck's solve_type on (source, sigma_t, sigma_s, dt) (ImplicitSolver.cpp:182-229,
 * e.g. SISolver::Solve SISolver.cpp:190-254): source holds Q on entry and psi on exit.  log_path (may be
 * null) receives the solver's log, whose "Iterations:" line gives the iteration count and last change. */
int ref_solve(double dt, const char *log_path) {
    return guarded([&] {
        FILE *old = Global::logfp;
        FILE *lf = log_path ? std::fopen(log_path, "w") : nullptr;
        if (lf) Global::logfp = lf;
        try {
            std::unique_ptr<ImplicitSolver<RKDG::OrdinateFlux>> solver(
                ImplicitSolver<RKDG::OrdinateFlux>::ImplicitSolverFactory::CreateSolver(g->dd, *g->source,
                                                                                         Global::input_list, dt));
            solver->Solve(*g->source, g->sigma_t, g->sigma_s, dt);
        } catch (...) {
            Global::logfp = old;
            if (lf) std::fclose(lf);
            throw;
        }
        Global::logfp = old;
        if (lf) std::fclose(lf);
    });
}

/* One step of the deck's IDC integrator (IDCIntegrator::Create and Step, idc.cpp:98-110,521-600 over
 * Step_Nonhybrid :846-940: idc_type, idc_nodes_type, idc_num_nodes, idc_num_corr, idc_use_collocation, and
 * the deck's solve_type for the implicit solves): psi (ghost cells zero, as the step assumes) advanced by
 * dt in place with the source held constant. */
int ref_idc_step(double dt) {
    return guarded([&] {
        std::unique_ptr<IDCIntegrator> idc(IDCIntegrator::Create(Global::input_list));
        idc->Step(*g->psi, *g->source, g->sigma_t, g->sigma_s, dt);
    });
}

}  // extern "C"
