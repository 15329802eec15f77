// Checks the host layer's field files, norms, differences, decks and projections without a GPU:
// byte layout of the reference's binary format, round trips, norms and projections against closed
// forms, the deck reader's rules, and Init on lattice and hohlraum decks.  Scratch files go to the
// directory given as argv[1].  Exit status 0 when every check passes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "tsw_host.hpp"

using namespace tsw_host;

static int failures = 0;
static void check(bool ok, const std::string &what) {
    std::printf("%s: %s\n", ok ? "ok" : "FAIL", what.c_str());
    if (!ok) ++failures;
}
static bool near(double a, double b, double tol) { return std::fabs(a - b) <= tol * std::max(1.0, std::fabs(b)); }

static std::vector<unsigned char> slurp(const std::string &path) {
    std::ifstream f(path, std::ios::binary);
    return std::vector<unsigned char>((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
}
template <class T>
static T at(const std::vector<unsigned char> &b, size_t off) {
    T v;
    std::memcpy(&v, b.data() + off, sizeof(T));
    return v;
}

static DomainDecomposition mesh(int64_t nx, int64_t ny, double ax, double bx, double ay, double by) {
    DomainDecomposition d;
    d.nx = nx;
    d.ny = ny;
    d.ax = ax;
    d.bx = bx;
    d.ay = ay;
    d.by = by;
    return d;
}

// Legendre coefficients of f(x, y) = a + b x + c y + g x y on each cell of dd, where x, y are global
// coordinates: on a cell, x = xc + (dx/2) xb.
static ScalarFlux bilinear(const DomainDecomposition &dd, int p, double a, double b, double c, double g) {
    ScalarFlux f(dd, p);
    const double dx = dd.dx(0), dy = dd.dx(1);
    for (int64_t i = 1; i <= dd.nx; ++i)
        for (int64_t j = 1; j <= dd.ny; ++j) {
            const double xc = dd.ax + (i - 0.5) * dx, yc = dd.ay + (j - 0.5) * dy, hx = dx / 2, hy = dy / 2;
            f(i, j, 0, 0) = a + b * xc + c * yc + g * xc * yc;
            if (p >= 1) {
                f(i, j, 1, 0) = (b + g * yc) * hx;
                f(i, j, 0, 1) = (c + g * xc) * hy;
                f(i, j, 1, 1) = g * hx * hy;
            }
        }
    return f;
}

static void test_density_format(const std::string &dir) {
    const DomainDecomposition dd = mesh(3, 2, -1.0, 2.0, 0.5, 1.5);
    ScalarFlux f(dd, 1);
    for (size_t k = 0; k < f.size(); ++k) f.data()[k] = 0.25 * (double)k - 3.0;
    const std::string path = dir + "/phi.sd2";
    f.WriteToDisk(path);
    const std::vector<unsigned char> b = slurp(path);
    const int64_t count = (3 + 2) * (2 + 2) * 4;
    check(b.size() == 6 * 8 + 3 * 8 + (size_t)count * 8, "DensityFunction file size: header + (nx+2)(ny+2)(p+1)^2 values");
    check(at<int64_t>(b, 0) == 3 && at<double>(b, 8) == -1.0 && at<double>(b, 16) == 2.0 && at<int64_t>(b, 24) == 2 &&
              at<double>(b, 32) == 0.5 && at<double>(b, 40) == 1.5,
          "mesh header: nx ax bx ny ay by");
    check(at<int64_t>(b, 48) == 1 && at<int64_t>(b, 56) == count * 8 && at<int64_t>(b, 64) == count,
          "density header: DG degree, bytes, count");
    check(at<double>(b, 72) == -3.0 && at<double>(b, 72 + 8 * (count - 1)) == 0.25 * (count - 1) - 3.0,
          "values in memory order, halos included");
    const ScalarFlux g = ScalarFlux::ReadFromDisk(path);
    check(g.DG_degree == 1 && g.domain().nx == 3 && g.domain().ny == 2 && g.domain().ay == 0.5 &&
              std::memcmp(g.data(), f.data(), f.size() * 8) == 0,
          "DensityFunction round trip is bitwise");
    // a corrupted count is refused
    {
        std::vector<unsigned char> c = b;
        const int64_t bad = count + 1;
        std::memcpy(c.data() + 64, &bad, 8);
        std::ofstream(dir + "/bad.sd2", std::ios::binary).write((const char *)c.data(), (std::streamsize)c.size());
        bool threw = false;
        try {
            ScalarFlux::ReadFromDisk(dir + "/bad.sd2");
        } catch (const std::runtime_error &) {
            threw = true;
        }
        check(threw, "inconsistent value count throws std::runtime_error");
        bool threw2 = false;
        try {
            ScalarFlux::ReadFromDisk(dir + "/missing.sd2");
        } catch (const std::runtime_error &) {
            threw2 = true;
        }
        check(threw2, "missing file throws std::runtime_error");
    }
}

static void test_ordinate_format(const std::string &dir) {
    const DomainDecomposition dd = mesh(2, 3, 0.0, 1.0, 0.0, 2.0);
    const OrdinateSet os = OrdinateSet::Named("spherical-triangle", 2, true);
    OrdinateFlux psi(dd, os, 2);
    for (size_t k = 0; k < psi.size(); ++k) psi.data()[k] = std::sin(0.1 * (double)k);
    const std::string path = dir + "/psi.af2";
    psi.WriteToDisk(path);
    const std::vector<unsigned char> b = slurp(path);
    const std::string name = "spherical-triangle#R";
    check(std::memcmp(b.data() + 48, name.c_str(), name.size() + 1) == 0, "ordinate name 'spherical-triangle#R' null-terminated");
    const size_t o = 48 + name.size() + 1;
    const int64_t count = 4 * 5 * 16 * 9;
    check(at<int64_t>(b, o) == 2 && at<int64_t>(b, o + 8) == 2 && at<int64_t>(b, o + 16) == count * 8 &&
              at<int64_t>(b, o + 24) == count && b.size() == o + 32 + (size_t)count * 8,
          "angular order, DG degree, bytes, count; nq = 16 for reduced T_2");
    const OrdinateFlux r = OrdinateFlux::ReadFromDisk(path);
    check(r.nq() == 16 && r.DG_degree == 2 && r.ordinates().type() == "spherical-triangle" &&
              r.ordinates().symmetric_reduce() && std::memcmp(r.data(), psi.data(), psi.size() * 8) == 0,
          "OrdinateFlux round trip rebuilds the ordinate set and is bitwise");
    bool same = true;
    for (int64_t q = 0; q < 16; ++q)
        same &= r.ordinates().xi(q) == os.xi(q) && r.ordinates().eta(q) == os.eta(q) && r.ordinates().w(q) == os.w(q);
    check(same, "rebuilt ordinate set equals the written one");
    const OrdinateSet full = OrdinateSet::Named("chebyshev-legendre", 4, false);
    OrdinateFlux psi2(dd, full, 0);
    psi2.WriteToDisk(dir + "/psi2.af2");
    const std::vector<unsigned char> b2 = slurp(dir + "/psi2.af2");
    check(std::strcmp((const char *)b2.data() + 48, "chebyshev-legendre") == 0, "unreduced set has no '#R'");
    check(OrdinateFlux::ReadFromDisk(dir + "/psi2.af2").nq() == 32, "unreduced Chebyshev-Legendre N=4 has 32 ordinates");
}

// Checkpoint files (TimeIntegrator::WriteCheckpoint, TimeIntegrator.cpp:745-870): int64 step, double time,
// double solve time, the OrdinateFlux record, 20 timer doubles; ResumeCheckpoint reads them back and refuses
// files of another problem or cut short, leaving psi unchanged.
static void test_checkpoint_format(const std::string &dir) {
    const DomainDecomposition dd = mesh(3, 2, 0.0, 1.5, 0.0, 1.0);
    const OrdinateSet os = OrdinateSet::Named("chebyshev-legendre", 4, true);
    OrdinateFlux psi(dd, os, 1);
    for (size_t k = 0; k < psi.size(); ++k) psi.data()[k] = std::cos(0.3 * (double)k);
    CheckpointState st;
    st.step = 0x2A;
    st.time = 4.2;
    st.solve_time = 1.25;
    for (int k = 0; k < 20; ++k) st.timers[k] = 0.5 * k;
    const std::string path = dir + "/run.002A.chk";
    WriteCheckpoint(path, st, psi);
    const std::vector<unsigned char> b = slurp(path);
    const size_t rec = 48 + std::strlen("chebyshev-legendre#R") + 1 + 32 + psi.size() * 8;
    check(at<int64_t>(b, 0) == 0x2A && at<double>(b, 8) == 4.2 && at<double>(b, 16) == 1.25 &&
              b.size() == 24 + rec + 20 * 8 && at<double>(b, 24 + rec + 8) == 0.5,
          "checkpoint layout: step, time, solve time, flux record, 20 timers");
    OrdinateFlux back(dd, os, 1);
    const CheckpointState r = ResumeCheckpoint(path, back);
    check(r.step == 0x2A && r.time == 4.2 && r.solve_time == 1.25 && r.timers[19] == 9.5 &&
              std::memcmp(back.data(), psi.data(), psi.size() * 8) == 0,
          "checkpoint round trip is bitwise");
    OrdinateFlux other(mesh(3, 3, 0.0, 1.5, 0.0, 1.0), os, 1);
    bool threw = false;
    try {
        ResumeCheckpoint(path, other);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    bool untouched = true;
    for (size_t k = 0; k < other.size(); ++k) untouched &= other.data()[k] == 0.0;
    check(threw && untouched, "a checkpoint of another mesh is refused and psi left as it was");
    {
        std::FILE *f = std::fopen((dir + "/short.chk").c_str(), "wb");
        std::fwrite(b.data(), 1, b.size() - 9, f);
        std::fclose(f);
    }
    threw = false;
    try {
        ResumeCheckpoint(dir + "/short.chk", back);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    check(threw, "a checkpoint cut short is refused");
    threw = false;
    try {
        ResumeCheckpoint(dir + "/missing.chk", back);
    } catch (const std::runtime_error &) {
        threw = true;
    }
    check(threw, "a missing checkpoint throws std::runtime_error");
}

static void test_norms() {
    // constant c on [0,2] x [0,3]: L1 = 6|c|, L2 = sqrt(6)|c|, Linf = |c|
    const DomainDecomposition d1 = mesh(4, 5, 0.0, 2.0, 0.0, 3.0);
    const ScalarFlux c = bilinear(d1, 2, -1.5, 0, 0, 0);
    check(near(c.L1Norm(), 9.0, 1e-14) && near(c.L2Norm(), std::sqrt(6.0) * 1.5, 1e-14) && c.LinfNorm() == 1.5,
          "norms of a constant");
    // non-negative bilinear f = 1 + x + 2y + xy on [0,1]^2: L1 = integral = 1 + 1/2 + 1 + 1/4
    const DomainDecomposition d2 = mesh(3, 4, 0.0, 1.0, 0.0, 1.0);
    const ScalarFlux f = bilinear(d2, 1, 1, 1, 2, 1);
    check(near(f.L1Norm(), 2.75, 1e-14), "L1 of a non-negative bilinear function is its integral");
    // L2^2 = integral over [0,1]^2 of f^2 = 1 + x^2 + 4y^2 + x^2y^2 + 2x + 4y + 2xy + 4xy + 2x^2y + 4xy^2,
    // term by term
    const double l2sq = 1 + 1.0 / 3 + 4.0 / 3 + 1.0 / 9 + 1 + 2 + 0.5 + 1 + 1.0 / 3 + 2.0 / 3;
    check(near(f.L2Norm(), std::sqrt(l2sq), 1e-14), "L2 of a bilinear function (Parseval)");
    // Linf of f: maximum 5 at (1,1), a sample point of every grid
    check(near(f.LinfNorm(), 5.0, 1e-14), "Linf of a bilinear function, maximum at a corner");
    // |xb| on one cell with p = 1: the 2-point rule gives 4/sqrt(3), not the exact 2
    const DomainDecomposition d3 = mesh(1, 1, 0.0, 2.0, 0.0, 2.0);
    ScalarFlux s(d3, 1);
    s(1, 1, 1, 0) = 1.0;
    check(near(s.L1Norm(), 4.0 / std::sqrt(3.0), 1e-14), "L1 uses DG degree + 1 Gauss points per axis");
    // 1 - xb^2 = 2/3 - (2/3) P2: uniform grids of 2, 4, ..., 64 points; 64 points give 1 - 1/63^2
    ScalarFlux q(d3, 2);
    q(1, 1, 0, 0) = 2.0 / 3.0;
    q(1, 1, 2, 0) = -2.0 / 3.0;
    check(near(q.LinfNorm(), 1.0 - 1.0 / (63.0 * 63.0), 1e-13), "Linf sampling stops at 64 points (tol 1e-3)");
    check(near(q.LinfNorm(1e-2), 1.0 - 1.0 / (31.0 * 31.0), 1e-13), "Linf sampling stops at 32 points (tol 1e-2)");
    check(near(f.EvaluateAtPoint(0.3, 0.7), 1 + 0.3 + 1.4 + 0.21, 1e-14) && near(f.EvaluateAtPoint(1.0, 1.0), 5.0, 1e-14),
          "EvaluateAtPoint, including the upper corner");
}

static void test_difference() {
    const DomainDecomposition c = mesh(2, 3, 0.0, 1.0, 0.0, 1.5), fmesh = mesh(6, 6, 0.0, 1.0, 0.0, 1.5);
    const ScalarFlux a = bilinear(c, 1, 0.5, 1, -2, 3), b = bilinear(fmesh, 2, 0.5, 1, -2, 3);
    const ScalarFlux d = ScalarFlux::ComputeDifference(a, b);
    check(d.domain().nx == 6 && d.domain().ny == 6 && d.DG_degree == 2, "difference on lcm mesh at the higher degree");
    check(d.L2Norm() < 1e-14 && d.LinfNorm() < 1e-14, "same bilinear function on two meshes: zero difference");
    const ScalarFlux e = ScalarFlux::ComputeDifference(b, bilinear(fmesh, 1, 0.0, 1, -2, 3));
    check(near(e.L1Norm(), 0.5 * 1.5, 1e-14) && near(e.LinfNorm(), 0.5, 1e-14), "same mesh: coefficient difference");
    // a piecewise-linear coarse function against a finer one: difference equals the analytic L2 gap
    const DomainDecomposition m3 = mesh(3, 1, 0.0, 3.0, 0.0, 1.0), m2 = mesh(2, 1, 0.0, 3.0, 0.0, 1.0);
    ScalarFlux u(m3, 0), v(m2, 0);
    u(1, 1, 0, 0) = 1;
    u(2, 1, 0, 0) = 2;
    u(3, 1, 0, 0) = 3;
    v(1, 1, 0, 0) = 1;
    v(2, 1, 0, 0) = 3;
    // on [0,1] 0, [1,1.5] -1, [1.5,2] 1, [2,3] 0: L1 = 1, L2 = 1
    const ScalarFlux w = ScalarFlux::ComputeDifference(u, v);
    check(w.domain().nx == 6 && near(w.L1Norm(), 1.0, 1e-14) && near(w.L2Norm(), 1.0, 1e-14) && near(w.LinfNorm(), 1.0, 1e-14),
          "piecewise constants on 3 and 2 cells: difference on 6 cells");
    bool threw = false;
    try {
        ScalarFlux::ComputeDifference(a, bilinear(mesh(2, 3, 0.0, 1.1, 0.0, 1.5), 1, 0, 0, 0, 0));
    } catch (const std::invalid_argument &) {
        threw = true;
    }
    check(threw, "different domains throw std::invalid_argument");
}

namespace {
struct Poly : Problem {
    std::string GetDescriptor() const override { return "poly"; }
    double Source(double x, double y) const override { return 1 + 2 * x - y + 0.5 * x * x * y; }
    double TotalCross(double x, double y) const override { return (x < 0.5 ? 1.0 : 0.0) * (y > -10 ? 1.0 : 0.0); }
};
}  // namespace

static void test_projection() {
    const DomainDecomposition dd = mesh(2, 2, 0.0, 1.0, 0.0, 2.0);
    Poly P;
    ScalarFlux s(dd, 2);
    ComputeProjection(P, s, &Problem::Source);
    // cell (1,1): x = 0.25 + 0.25 xb, y = 0.5 + 0.5 yb;  f = 1 + 2x - y + x^2 y / 2
    // exact Legendre coefficients of that polynomial:
    // x^2 = 1/16 (1 + xb)^2 = 1/16 (1 + 2 xb + xb^2), xb^2 = 1/3 + 2/3 P2
    const double c00 = 1 + 0.5 - 0.5 + 0.5 * (1.0 / 16) * (1 + 1.0 / 3) * 0.5;
    check(near(s(1, 1, 0, 0), c00, 1e-14), "projection of a polynomial: cell mean");
    check(near(s(1, 1, 2, 1), 0.5 * (1.0 / 16) * (2.0 / 3) * 0.5, 1e-14), "projection of a polynomial: (2,1) moment");
    check(std::fabs(s(0, 0, 0, 0)) == 0.0, "halo cells stay zero");
    // a step x < 0.5 on the cell [0, 1] (p = 2): mean 1/2, the limiter keeps it and makes samples >= 0
    const DomainDecomposition d1 = mesh(1, 1, 0.0, 1.0, 0.0, 1.0);
    ScalarFlux raw(d1, 2), lim(d1, 2);
    ComputeProjection(P, raw, &Problem::TotalCross, false);
    ComputeProjection(P, lim, &Problem::TotalCross, true);
    check(near(raw(1, 1, 0, 0), 0.5, 1e-14) && near(lim(1, 1, 0, 0), 0.5, 1e-14), "step: mean 1/2 with and without limiter");
    // raw P1 moment: the 14-point rule applied to the step, (3/2) sum over negative nodes of w x, close
    // to the exact -3/4; the P2 moment vanishes by symmetry of the rule
    double gx[14], gw[14], c1 = 0.0;
    tsw_gauss_legendre(14, gx, gw);
    for (int l = 0; l < 14; ++l)
        if (gx[l] < 0) c1 += 1.5 * gw[l] * gx[l];
    check(near(raw(1, 1, 1, 0), c1, 1e-14) && std::fabs(c1 + 0.75) < 5e-3 && std::fabs(raw(1, 1, 2, 0)) < 1e-14,
          "step: P1 moment from the 14-point rule, P2 moment zero");
    // limited: the largest theta keeping c0 + theta c1 P1 >= 0 on the samples; xb = 1 is one of them
    const double theta = lim(1, 1, 1, 0) / raw(1, 1, 1, 0);
    check(std::fabs(theta - 0.5 / -c1) < 1e-13 && theta <= 0.5 / -c1, "limiter theta makes the value at xb = 1 zero");
}

static void write_text(const std::string &path, const std::string &text) { std::ofstream(path) << text; }

static void test_decks(const std::string &dir) {
    write_text(dir + "/a.deck",
               "# comment line\n"
               "NX  12   # trailing comment\n"
               "Problem Lattice\n"
               "\n"
               "lonely\n"
               "nx 99\n"
               "ordinate_type Spherical-Triangle\n");
    ParameterList p;
    p.Set("keep", "1");
    p.ReadFromFile(dir + "/a.deck");
    check(p.GetString("nx") == "12", "keys are lower-cased, comments dropped, first value kept");
    check(p.GetString("problem") == "lattice" && p.GetString("ordinate_type") == "spherical-triangle",
          "values are lower-cased");
    check(p.GetString("lonely") == "lattice", "a key without value takes the previous line's value");
    check(p.Has("") && p.GetString("") == "", "a leading comment line inserts the empty key");
    check(p.GetString("keep") == "1", "keys already present are kept");
    p.WriteToFile(dir + "/b.deck");
    ParameterList q;
    q.ReadFromFile(dir + "/b.deck");
    check(q.items() == p.items(), "WriteToFile / ReadFromFile round trip");
    bool threw = false;
    try {
        ParameterList().ReadFromFile(dir + "/none.deck");
    } catch (const std::runtime_error &) {
        threw = true;
    }
    check(threw, "missing deck throws std::runtime_error");

    // lattice at 4 cells per unit, the reference deck's keys
    write_text(dir + "/lattice.deck",
               "ax 0.0\nbx 5.0\nay 0.0\nby 7.0\nnx 28\nny 28\ndg_degree 2\npiecewise_cross true\n"
               "problem lattice\ntfinal 3.2\ncfl 25.6\nmoll_radius 0.0\nordinate_type spherical-triangle\n"
               "ang_order 4\nordinate_sym_reduce true\nsolve_type petsc-peierls-gmres\n");
    ParameterList lp;
    lp.ReadFromFile(dir + "/lattice.deck");
    const ProblemInputs in = Init(lp);
    check(lp.GetString("bx") == "7.0" && in.dd.bx == 7.0, "lattice overrides bx to 7");
    check(in.os.nq() == 64 && in.os.type() == "spherical-triangle", "T_4 reduced: 64 ordinates");
    check(near(in.max_step_size, 25.6 * 0.25, 1e-15) && in.final_time == 3.2, "cfl * dx and tfinal");
    static const int absorbers[11][2] = {{1, 1}, {1, 3}, {1, 5}, {2, 2}, {2, 4}, {3, 5}, {4, 2}, {4, 4}, {5, 1}, {5, 3}, {5, 5}};
    bool sig_ok = true, src_ok = true, hi_ok = true;
    for (int64_t i = 1; i <= 28; ++i)
        for (int64_t j = 1; j <= 28; ++j) {
            const int bi = (int)((i - 1) / 4), bj = (int)((j - 1) / 4);
            bool a = false;
            for (auto &ab : absorbers) a |= ab[0] == bi && ab[1] == bj;
            sig_ok &= in.sigma_t(i, j) == (a ? 10.0 : 1.0) || near(in.sigma_t(i, j), a ? 10.0 : 1.0, 1e-14);
            sig_ok &= near(in.sigma_s(i, j), a ? 0.0 : 1.0, 1e-14);
            const double q = (bi == 3 && bj == 3) ? 1.0 : 0.0;
            src_ok &= near(in.source_density(i, j, 0, 0), q, 1e-14) && near(in.source(7, i, j, 0, 0), q / (4 * M_PI), 1e-14);
            for (int d = 0; d <= 2; ++d)
                for (int e = 0; e <= 2; ++e)
                    if (d + e) hi_ok &= std::fabs(in.source_density(i, j, d, e)) < 1e-14;
        }
    check(sig_ok, "lattice cross sections on aligned cells");
    check(src_ok && hi_ok, "lattice source: unit density on (3,4)^2, no higher moments");
    bool vac = true;
    for (int64_t j = 0; j <= 29; ++j)
        for (int64_t q = 0; q < 64; ++q) vac &= in.source(q, 0, j, 0, 0) == 0.0 && in.source(q, 29, j, 0, 0) == 0.0;
    check(vac, "lattice inflow is vacuum");

    write_text(dir + "/hohl.deck",
               "nx 26\nny 13\nax 0\nbx 1.3\nay 0\nby 1.3\ndg_degree 1\npiecewise_cross true\nproblem hohlraum\n"
               "moll_radius 0\nang_order 4\nordinate_type chebyshev-legendre\n");
    ParameterList hp;
    hp.ReadFromFile(dir + "/hohl.deck");
    const ProblemInputs h = Init(hp);
    check(h.os.nq() == 32 && !h.os.symmetric_reduce(), "ordinate_sym_reduce defaults to false");
    bool inflow = true;
    for (int64_t j = 0; j <= 14; ++j) inflow &= near(h.source(5, 0, j, 0, 0), 1.0 / (4 * M_PI), 1e-15);
    inflow &= h.source(5, 27, 3, 0, 0) == 0.0 && h.source(5, 3, 0, 0, 0) == 0.0;
    check(inflow, "hohlraum inflow on the x = 0 halo only");
    // cell (1, 7): x in [0, 0.05], y in [0.6, 0.7]: a wall cell; cell (2, 7): x in [0.05, 0.1]: interior
    check(near(h.sigma_t(1, 7), 100.0, 1e-13) && near(h.sigma_t(2, 7), 0.1, 1e-14), "hohlraum walls");
    // cell (10, 7): x in [0.45, 0.5], wall with total 100
    check(near(h.sigma_t(10, 7), 100.0, 1e-13), "hohlraum inner block");

    auto throws = [&](const std::string &text, const char *what) {
        write_text(dir + "/x.deck", text);
        ParameterList x;
        x.ReadFromFile(dir + "/x.deck");
        bool t = false;
        try {
            Init(x);
        } catch (const std::exception &) {
            t = true;
        }
        check(t, what);
    };
    const std::string base = "nx 4\nny 4\nax 0\nbx 7\nay 0\nby 7\ndg_degree 1\nproblem lattice\nmoll_radius 0\nang_order 2\n";
    {
        // piecewise_cross absent (false): cross sections at the flux degree (init.cpp:741); the 4 x 4 mesh of
        // the 7 x 7 lattice cuts the blocks, so the projections carry x and y moments
        write_text(dir + "/ho.deck", base);
        ParameterList x;
        x.ReadFromFile(dir + "/ho.deck");
        const ProblemInputs ho = Init(x);
        double hi = 0.0;
        for (int64_t i = 1; i <= 4; ++i)
            for (int64_t j = 1; j <= 4; ++j) hi = std::max(hi, std::fabs(ho.sigma_t(i, j, 1, 0)));
        check(ho.sigma_t.DG_degree == 1 && ho.sigma_s.DG_degree == 1 && hi > 0.1 &&
                  ho.sigma_t(2, 2) == ho.sigma_t_density(2, 2, 0, 0) &&
                  ho.sigma_t(2, 2, 1, 1) == ho.sigma_t_density(2, 2, 1, 1),
              "piecewise_cross false: degree-1 cross sections with higher moments");
    }
    throws(base + "piecewise_cross true\nperiodic true\nreflect_x_min true\n",
           "periodic boundaries combined with reflecting ones are refused");
    {  // reflecting edges: the mask Init derives from the reflect_* keys (init.cpp:448-477)
        write_text(dir + "/r.deck", base + "piecewise_cross true\nreflect_y_min true\nreflect_x_max true\n");
        ParameterList r;
        r.ReadFromFile(dir + "/r.deck");
        ProblemInputs in = Init(r);
        check(in.dd.reflect_mask == (TSW_REFLECT_Y_MIN | TSW_REFLECT_X_MAX), "reflect_* keys give the edge mask");
    }
    throws("nx 4\nny 4\nax 0\nbx 7\nay 0\nby 7\ndg_degree 1\nproblem reeds\nmoll_radius 0\nang_order 2\npiecewise_cross true\n",
           "unknown problem is refused");
    throws(base + "piecewise_cross true\nordinate_type lebedev\n", "unavailable ordinate type is refused");
    throws("nx 4\nny 4\nax 0\nbx 7\nay 0\nby 7\ndg_degree 1\nproblem lattice\nang_order 2\npiecewise_cross true\n",
           "moll_radius is required");
}

static void test_init_files(const std::string &dir) {
    write_text(dir + "/files.deck",
               "nx 14\nny 14\nax 0\nbx 7\nay 0\nby 7\ndg_degree 1\npiecewise_cross true\nproblem lattice\n"
               "moll_radius 0\nang_order 2\nordinate_type spherical-triangle\noutput_init true\n");
    ParameterList p;
    p.ReadFromFile(dir + "/files.deck");
    const ProblemInputs a = Init(p, dir);
    const ScalarFlux st = ScalarFlux::ReadFromDisk(dir + "/sigma_t_dg0_nx0000014.cs2");
    const ScalarFlux sq = ScalarFlux::ReadFromDisk(dir + "/source_dg1_nx0000014.sd2");
    check(st.DG_degree == 0 && std::memcmp(st.data(), a.sigma_t_density.data(), st.size() * 8) == 0,
          "output_init writes sigma_t_dg0_nx0000014.cs2");
    check(sq.DG_degree == 1 && std::memcmp(sq.data(), a.source_density.data(), sq.size() * 8) == 0,
          "output_init writes source_dg1_nx0000014.sd2");
    // a second Init reads the files back: change the file and see the change
    ScalarFlux mod = st;
    mod(3, 3, 0, 0) = 42.0;
    mod.WriteToDisk(dir + "/sigma_t_dg0_nx0000014.cs2");
    ParameterList p2;
    p2.ReadFromFile(dir + "/files.deck");
    const ProblemInputs b = Init(p2, dir);
    check(b.sigma_t(3, 3) == 42.0, "output_init reads existing projection files");
}

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: fields_test OUTDIR\n");
        return 2;
    }
    const std::string dir = argv[1];
    try {
        test_density_format(dir);
        test_ordinate_format(dir);
        test_checkpoint_format(dir);
        test_norms();
        test_difference();
        test_projection();
        test_decks(dir);
        test_init_files(dir);
    } catch (const std::exception &e) {
        std::printf("FAIL: exception %s\n", e.what());
        return 1;
    }
    std::printf("%d failures\n", failures);
    return failures == 0 ? 0 : 1;
}
