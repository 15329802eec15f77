// Drives the C++ host layer (tsw_host.hpp) the way the reference's driver drives its operators:
// builds a mesh, ordinate set, cross sections and source, then runs SweepOperator::Apply and the
// "source-iteration" and "petsc-peierls-gmres" solvers from the factory.  Writes inputs and results
// as raw float64 files into the directory given as argv[1] for comparison with the oracle, and
// checks the factory's error behaviour.  Exit status 0 when every check passes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "tsw_host.hpp"

using namespace tsw_host;

static void dump(const std::string &path, const double *p, size_t n) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(p, sizeof(double), n, f) != n) throw std::runtime_error("cannot write " + path);
    std::fclose(f);
}

static double lcg(uint64_t &s) {  // deterministic inputs
    s = s * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(s >> 11) / 9007199254740992.0;
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: host_api_test OUTDIR\n");
        return 2;
    }
    const std::string out = argv[1];
    int failures = 0;
    auto check = [&](bool ok, const char *what) {
        std::printf("%s: %s\n", ok ? "ok" : "FAIL", what);
        if (!ok) ++failures;
    };
    DomainDecomposition dd;
    dd.nx = 12; dd.ny = 10; dd.ax = 0.0; dd.bx = 1.5; dd.ay = 0.0; dd.by = 1.0;
    const int p = 2;
    const OrdinateSet os = OrdinateSet::LevelSymmetric(6);
    CrossSection st(dd), ss(dd);
    OrdinateFlux q(dd, os, p);
    uint64_t seed = 12345;
    for (int64_t i = 0; i <= dd.nx + 1; ++i)
        for (int64_t j = 0; j <= dd.ny + 1; ++j) {
            st(i, j) = 0.5 + 3.0 * lcg(seed);
            ss(i, j) = 0.6 * st(i, j);
        }
    for (size_t k = 0; k < q.size(); ++k) q.data()[k] = lcg(seed) - 0.3;
    dump(out + "/sigma_t.bin", st.data(), (size_t)(dd.nx + 2) * (dd.ny + 2));
    dump(out + "/sigma_s.bin", ss.data(), (size_t)(dd.nx + 2) * (dd.ny + 2));
    dump(out + "/source.bin", q.data(), q.size());

    auto dev = std::make_shared<Device>(dd, os, p);
    ParameterList params = ParameterList::FromText("sweep_pattern hip-staggered   # traversal\nabs_tol 1e-13\n"
                                                   "max_its 1000\ngmres_restart 30\n");
    SweepOperator sweep(dev, params);
    OrdinateFlux psi = q;
    sweep.Apply(psi, st);
    dump(out + "/sweep.bin", psi.data(), psi.size());

    // L (L^{-1} q) = q on the interior
    OrdinateFlux back(dd, os, p);
    TransportOperator::Lmv(*dev, 1.0, 0.0, st, psi, back);
    double num = 0, den = 0;
    for (int64_t i = 1; i <= dd.nx; ++i)
        for (int64_t j = 1; j <= dd.ny; ++j)
            for (int64_t k = 0; k < os.nq(); ++k)
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e) {
                        const double a = back(k, i, j, d, e), b = q(k, i, j, d, e);
                        num += (a - b) * (a - b);
                        den += b * b;
                    }
    check(std::sqrt(num / den) < 1e-12, "Lmv(Apply(q)) == q");

    params.Set("solve_type", "source-iteration");
    auto si = ImplicitSolver::Create(dev, params);
    OrdinateFlux psi_si = q;
    si->Solve(psi_si, st, ss);
    check(si->GetDescriptor() == "source-iteration" && si->final_error() <= 1e-13, "source iteration converged");
    dump(out + "/si.bin", psi_si.data(), psi_si.size());

    params.Set("solve_type", "petsc-peierls-gmres");
    auto gm = ImplicitSolver::Create(dev, params);
    OrdinateFlux psi_gm = q;
    gm->Solve(psi_gm, st, ss);
    check(gm->final_error() <= 1e-13, "Peierls GMRES converged");
    dump(out + "/gmres.bin", psi_gm.data(), psi_gm.size());

    ScalarFlux phi_si(dd, p), phi_gm(dd, p);
    TransportOperator::Pmv(*dev, 1.0, 0.0, psi_si, phi_si, OpDomain::Interior);
    TransportOperator::Pmv(*dev, 1.0, 0.0, psi_gm, phi_gm, OpDomain::Interior);
    num = den = 0;
    for (size_t k = 0; k < phi_si.size(); ++k) {
        num += (phi_si.data()[k] - phi_gm.data()[k]) * (phi_si.data()[k] - phi_gm.data()[k]);
        den += phi_si.data()[k] * phi_si.data()[k];
    }
    check(std::sqrt(num / den) < 1e-10, "GMRES and SI reach the same scalar flux");

    // DSA-preconditioned GMRES (use_dsa) reaches the same solution
    params.Set("use_dsa", "true");
    auto gd = ImplicitSolver::Create(dev, params);
    OrdinateFlux psi_gd = q;
    gd->Solve(psi_gd, st, ss);
    check(gd->final_error() <= 1e-13, "Peierls GMRES with DSA converged");
    dump(out + "/gmres_dsa.bin", psi_gd.data(), psi_gd.size());
    params.Set("use_dsa", "false");

    // two IDC steps (4 Gauss-Lobatto nodes, 3 corrections) from the interior of q, vacuum inflow
    ParameterList ip = ParameterList::FromText("idc_num_nodes 4\nidc_num_corr 3\nidc_nodes_type gauss-lobatto\n"
                                               "abs_tol 1e-13\n");
    IDCIntegrator idc(dev, ip);
    OrdinateFlux psi_t = q, src0 = q;
    for (int64_t i = 0; i < dd.nx + 2; ++i)
        for (int64_t j = 0; j < dd.ny + 2; ++j)
            if (i == 0 || j == 0 || i == dd.nx + 1 || j == dd.ny + 1)
                for (int64_t k = 0; k < os.nq() * (p + 1) * (p + 1); ++k) {
                    const size_t at = ((size_t)i * (dd.ny + 2) + j) * os.nq() * (p + 1) * (p + 1) + k;
                    psi_t.data()[at] = 0.0;
                    src0.data()[at] = 0.0;
                }
    idc.Step(psi_t, src0, st, ss, 0.05);
    idc.Step(psi_t, src0, st, ss, 0.05);
    dump(out + "/idc.bin", psi_t.data(), psi_t.size());
    // the reference's SetNodes rejects the equispaced families (idc.cpp:2344-2352), and idc_type has two values
    bool threw_idc = false;
    try {
        ParameterList badi;
        badi.Set("idc_nodes_type", "equispaced-both");
        IDCIntegrator bad_idc(dev, badi);
    } catch (const std::out_of_range &) {
        threw_idc = true;
    }
    check(threw_idc, "unbuilt idc_nodes_type throws std::out_of_range");
    bool threw_type = false;
    try {
        ParameterList badt;
        badt.Set("idc_type", "none");
        IDCIntegrator bad_idc(dev, badt);
    } catch (const std::out_of_range &) {
        threw_type = true;
    }
    check(threw_type, "unbuilt idc_type throws std::out_of_range");
    {
        ParameterList ok;
        ok.Set("idc_type", "update");
        ok.Set("idc_nodes_type", "gauss-radau");
        ok.Set("idc_num_nodes", "2");
        ok.Set("idc_num_corr", "2");
        IDCIntegrator radau(dev, ok);  // the reference input deck's IDC settings construct
        check(radau.GetDescriptor() == "idc", "gauss-radau update-form integrator constructs");
    }

    bool threw = false;
    try {
        params.Set("solve_type", "no-such-solver");
        ImplicitSolver::Create(dev, params);
    } catch (const std::out_of_range &) {
        threw = true;
    }
    check(threw, "unknown solve_type throws std::out_of_range");
    threw = false;
    try {
        ParameterList bad;
        bad.Set("sweep_pattern", "no-such-pattern");
        SweepOperator s2(dev, bad);
    } catch (const std::out_of_range &) {
        threw = true;
    }
    check(threw, "unknown sweep_pattern throws std::out_of_range");
    return failures ? 1 : 0;
}
