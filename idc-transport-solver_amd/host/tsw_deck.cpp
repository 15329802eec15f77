// Deck input path of the host layer: "key value" decks, the named ordinate sets, problem definitions
// and their L2 projection into the DG space, and Init, which turns a deck into sweep inputs.
#include <cctype>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <limits>
#include <sstream>
#include <stdexcept>

#include "tsw_host.hpp"

namespace tsw_host {

// ---------------------------------------------------------------------------------------------------
// ReadFromFile keeps the reference's reading rules (ParameterList.cpp:202-245): key and value are read
// into the same two strings line after line, so a line with a key and no value takes the previous
// line's value and a blank line re-inserts the previous pair; insert never replaces an existing key.
ParameterList &ParameterList::ReadFromFile(const std::string &filename) {
    std::ifstream file(filename);
    if (!file.is_open()) throw std::runtime_error("Failed to open file '" + filename + "' in 'ReadFromFile'.");
    std::string line, key, value;
    while (std::getline(file, line)) {
        line = line.substr(0, line.find('#'));
        for (char &c : line) c = (char)std::tolower((unsigned char)c);
        std::stringstream(line) >> key >> value;
        values_.insert(std::make_pair(key, value));
    }
    return *this;
}

const ParameterList &ParameterList::WriteToFile(const std::string &filename) const {
    std::ofstream file(filename);
    if (!file.is_open()) throw std::runtime_error("Failed to open file '" + filename + "' in 'WriteToFile'.");
    for (const auto &kv : values_) file << kv.first << " " << kv.second << std::endl;
    return *this;
}

bool ParameterList::GetBool(const std::string &key) const {
    auto it = values_.find(key);
    if (it == values_.end()) throw std::out_of_range("Key '" + key + "' not found in input list.");
    bool v = false;
    if (!(std::stringstream(it->second) >> std::boolalpha >> v))
        throw std::runtime_error("Error converting value '" + it->second + "' for key '" + key + "' to type 'bool'.");
    return v;
}

OrdinateSet OrdinateSet::FromParameters(const ParameterList &params) {
    const int64_t order = params.GetInt("ang_order");
    const std::string type = params.Has("ordinate_type") ? params.GetString("ordinate_type") : "chebyshev-legendre";
    bool reduce = false;
    try {
        reduce = params.GetBool("ordinate_sym_reduce");
    } catch (...) {
        reduce = false;
    }
    return Named(type, (int)order, reduce);
}

// ---------------------------------------------------------------------------------------------------
namespace {

// Overrides a domain bound the deck gives with another value (LatticeProblem.cpp:100-140).
void override_bound(ParameterList &params, const char *key, double value, const char *text) {
    try {
        if (params.GetDouble(key) != value) params.Set(key, text);
    } catch (...) {
    }
}

// Checkerboard lattice on [0, 7]^2 (LatticeProblem.cpp:100-324).
class LatticeProblem : public Problem {
  public:
    std::string GetDescriptor() const override { return "lattice"; }
    void OverrideOptions(ParameterList &params) const override {
        override_bound(params, "ax", 0.0, "0.0");
        override_bound(params, "bx", 7.0, "7.0");
        override_bound(params, "ay", 0.0, "0.0");
        override_bound(params, "by", 7.0, "7.0");
    }
    double Source(double x, double y) const override { return (x > 3.0 && x < 4.0 && y > 3.0 && y < 4.0) ? 1.0 : 0.0; }
    double TotalCross(double x, double y) const override { return absorber(x, y) ? 10.0 : 1.0; }
    double ScatterCross(double x, double y) const override { return absorber(x, y) ? 0.0 : 1.0; }

  private:
    // unit blocks (a, a+1] x (b, b+1] with lower corners (1,1) (1,3) (1,5) (2,2) (2,4) (3,5) (4,2) (4,4)
    // (5,1) (5,3) (5,5)
    static bool in(double v, double a) { return v > a && v <= a + 1.0; }
    static bool absorber(double x, double y) {
        if ((in(x, 1.0) || in(x, 5.0)) && (in(y, 1.0) || in(y, 3.0) || in(y, 5.0))) return true;
        if ((in(x, 2.0) || in(x, 4.0)) && (in(y, 2.0) || in(y, 4.0))) return true;
        return in(x, 3.0) && in(y, 5.0);
    }
};

// Hohlraum on [0, 1.3]^2 with unit isotropic inflow on the x = 0 edge (HohlraumProblem.cpp:60-272).
class HohlraumProblem : public Problem {
  public:
    std::string GetDescriptor() const override { return "hohlraum"; }
    void SetInflowBoundaries(ScalarFlux &source) const override {
        for (int64_t j = 0; j <= source.domain().ny + 1; ++j) source(0, j, 0, 0) = 1.0;
    }
    void OverrideOptions(ParameterList &params) const override {
        override_bound(params, "ax", 0.0, "0.0");
        override_bound(params, "bx", 1.3, "1.3");
        override_bound(params, "ay", 0.0, "0.0");
        override_bound(params, "by", 1.3, "1.3");
        try {
            if (params.GetBool("periodic")) params.Set("periodic", "false");
        } catch (...) {
        }
        try {
            if (params.GetBool("reflect_x_min")) params.Set("reflect_x_min", "false");
        } catch (...) {
        }
    }
    double TotalCross(double x, double y) const override { return wall(x, y) ? 100.0 : 0.1; }
    double ScatterCross(double x, double y) const override { return wall(x, y) ? 1.0 : 0.1; }

  protected:
    static bool wall(double x, double y) {
        if (x > 1.25 || y <= 0.05 || y > 1.25) return true;
        return (y > 0.25 && y <= 1.05) && (x <= 0.05 || (x > 0.45 && x <= 0.85));
    }
};

// Hohlraum with a graded scattering cross section (Hohlraum2Problem.cpp:50-75).
class Hohlraum2Problem : public HohlraumProblem {
  public:
    std::string GetDescriptor() const override { return "hohlraum2"; }
    double ScatterCross(double x, double y) const override {
        if (x > 1.25 || y > 1.25 || y < 0.05) return 50.0;
        if (x > 0.50 && x <= 0.85 && y > 0.30 && y <= 1.00) return 0.0;
        if (x > 0.45 && x <= 0.50 && y > 0.25 && y <= 1.05) return 90.0;
        if (x > 0.50 && x <= 0.85 && y > 1.00 && y <= 1.05) return 90.0;
        if (x > 0.50 && x <= 0.85 && y > 0.25 && y <= 0.30) return 90.0;
        if (x <= 0.05 && y > 0.25 && y <= 1.05) return 95.0;
        return 0.1;
    }
};

// Isotropic initial bump exp(1 / (r^2/R^2 - 1)) for r < R, normalised to unit integral, in a medium of
// uniform sigma_t = sigma_s = sigma_val with no source (BumpProblem.cpp, Mollifier.cpp:200-300).  The
// integral over the disc is pi R^2 int_0^1 exp(-1/t) dt (substituting t = 1 - r^2/R^2); the reference
// evaluates it by adaptive Gauss-Kronrod to 1e-15, here by Gauss-Legendre panels to the same accuracy.
class BumpProblem : public Problem {
  public:
    explicit BumpProblem(const ParameterList &params)
        : sigma_(params.GetDouble("sigma_val")), R_(params.GetDouble("bump_radius")) {
        if (!(R_ > 0.0)) throw std::invalid_argument("bump_radius must be positive");
        double g[20], w[20];
        tsw_gauss_legendre(20, g, w);
        double I = 0.0;
        const int panels = 64;
        for (int k = 0; k < panels; ++k)
            for (int l = 0; l < 20; ++l) {
                const double t = (k + 0.5 * (g[l] + 1.0)) / panels;
                I += 0.5 / panels * w[l] * std::exp(-1.0 / t);
            }
        norm_ = 1.0 / (M_PI * R_ * R_ * I);
    }
    std::string GetDescriptor() const override { return "bump"; }
    double InitialCondition(double x, double y) const override {
        const double u = (x * x + y * y) / (R_ * R_);
        return u < 1.0 ? std::exp(1.0 / (u - 1.0)) * norm_ : 0.0;
    }
    double TotalCross(double, double) const override { return sigma_; }
    double ScatterCross(double, double) const override { return sigma_; }

  private:
    double sigma_, R_, norm_;
};

// Sphere (disc in 2D) of a pure scatterer with smooth transitions between an inner, an intermediate and
// an outer cross section, and a unit source smoothed over the inner transition
// (InhomogeneousSphereProblem.cpp:24-275).  Each transition of width w around radius R is
// 1 / (1 + exp(2 rho / (R + rho - r) - 2 rho / (r - R + rho))), rho = w / 2.
class InhomogeneousSphereProblem : public Problem {
  public:
    explicit InhomogeneousSphereProblem(const ParameterList &params)
        : ri_(params.GetDouble("sphere_inner_radius")), ro_(params.GetDouble("sphere_outer_radius")),
          wi_(params.GetDouble("sphere_inner_trans_radius")), wo_(params.GetDouble("sphere_outer_trans_radius")),
          si_(params.GetDouble("sphere_inner_sigma_val")), sm_(params.GetDouble("sphere_interm_sigma_val")),
          so_(params.GetDouble("sphere_outer_sigma_val")) {}
    std::string GetDescriptor() const override { return "inhomogeneous-sphere"; }
    double Source(double x, double y) const override {
        const double r = std::sqrt(x * x + y * y);
        if (r <= ri_ - 0.5 * wi_) return 1.0;
        if (r < ri_ + 0.5 * wi_) return step(r, ri_, wi_);
        return 0.0;
    }
    double TotalCross(double x, double y) const override {
        const double r = std::sqrt(x * x + y * y);
        if (r <= ri_ - 0.5 * wi_) return si_;
        if (r < ri_ + 0.5 * wi_) return sm_ + (si_ - sm_) * step(r, ri_, wi_);
        if (r <= ro_ - 0.5 * wo_) return sm_;
        if (r < ro_ + 0.5 * wo_) return so_ + (sm_ - so_) * step(r, ro_, wo_);
        return so_;
    }
    double ScatterCross(double x, double y) const override { return TotalCross(x, y); }

  private:
    static double step(double r, double R, double w) {
        const double rho = 0.5 * w;
        return 1.0 / (1.0 + std::exp(2 * rho / ((R + rho) - r) - 2 * rho / (r - (R - rho))));
    }
    double ri_, ro_, wi_, wo_, si_, sm_, so_;
};

double legendre(int n, double x) {
    double p0 = 1.0, p1 = x;
    if (n == 0) return 1.0;
    for (int k = 2; k <= n; ++k) {
        const double pk = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    return p1;
}

}  // namespace

std::unique_ptr<Problem> Problem::Create(const ParameterList &params) {
    const std::string name = params.GetString("problem");
    const double moll = params.GetDouble("moll_radius");
    std::unique_ptr<Problem> p;
    if (name == "lattice")
        p.reset(new LatticeProblem);
    else if (name == "hohlraum")
        p.reset(new HohlraumProblem);
    else if (name == "hohlraum2")
        p.reset(new Hohlraum2Problem);
    else if (name == "bump")
        p.reset(new BumpProblem(params));
    else if (name == "inhomogeneous-sphere")
        p.reset(new InhomogeneousSphereProblem(params));
    else
        throw std::out_of_range("problem '" + name + "' is not available");
    if (moll >= 1e-14) throw std::invalid_argument("mollified problem functions (moll_radius > 0) are not available");
    return p;
}

// ---------------------------------------------------------------------------------------------------
void ComputeProjection(const Problem &prob, ScalarFlux &dest, ProblemEvalFcn f, bool apply_limiter, int proj_order) {
    const double TOL = 1.0e-14;
    const int sample_pts = 20;
    std::fill(dest.data(), dest.data() + dest.size(), 0.0);
    std::vector<double> gx(proj_order), gw(proj_order);
    if (tsw_gauss_legendre(proj_order, gx.data(), gw.data()) != TSW_OK)
        throw std::invalid_argument("bad projection order");
    const DomainDecomposition &dd = dest.domain();
    const int p = dest.DG_degree;
    const double dx = dd.dx(0), dy = dd.dx(1);
    std::vector<double> lg((size_t)proj_order * (p + 1)), ls((size_t)(sample_pts + 1) * (p + 1));
    for (int l = 0; l < proj_order; ++l)
        for (int d = 0; d <= p; ++d) lg[(size_t)l * (p + 1) + d] = legendre(d, gx[l]);
    for (int l = 0; l <= sample_pts; ++l)
        for (int d = 0; d <= p; ++d) ls[(size_t)l * (p + 1) + d] = legendre(d, -1.0 + (2.0 / sample_pts) * l);
    // the polynomial on the 21 x 21 sample grid with the moments of degree > 0 scaled by theta
    auto negative = [&](int64_t i, int64_t j, double theta) {
        bool neg = false;
        for (int l = 0; l <= sample_pts; ++l)
            for (int m = 0; m <= sample_pts; ++m) {
                double v = 0.0;
                for (int d = 0; d <= p; ++d)
                    for (int e = 0; e <= p; ++e)
                        v += (d + e == 0 ? 1.0 : theta) * dest(i, j, d, e) * ls[(size_t)l * (p + 1) + d] *
                             ls[(size_t)m * (p + 1) + e];
                neg |= v < 0.0;
            }
        return neg;
    };
    for (int64_t i = 1; i <= dd.nx; ++i)
        for (int64_t j = 1; j <= dd.ny; ++j) {
            for (int l = 0; l < proj_order; ++l)
                for (int m = 0; m < proj_order; ++m) {
                    const double xl = dd.ax + (i - 1) * dx + (gx[l] + 1.0) * dx / 2.0;
                    const double ym = dd.ay + (j - 1) * dy + (gx[m] + 1.0) * dy / 2.0;
                    const double fv = (prob.*f)(xl, ym);
                    for (int d = 0; d <= p; ++d)
                        for (int e = 0; e <= p; ++e)
                            dest(i, j, d, e) += gw[l] * gw[m] * lg[(size_t)l * (p + 1) + d] * lg[(size_t)m * (p + 1) + e] *
                                                (2 * d + 1) * (2 * e + 1) / 4.0 * fv;
                }
            if (!apply_limiter || dest(i, j, 0, 0) < 0.0 || !negative(i, j, 1.0)) continue;
            double lo = 0.0, hi = 1.0;
            while (hi - lo > TOL) {
                const double mid = 0.5 * (hi + lo);
                if (negative(i, j, mid))
                    hi = mid;
                else
                    lo = mid;
            }
            for (int d = 0; d <= p; ++d)
                for (int e = 0; e <= p; ++e)
                    if (d + e > 0) dest(i, j, d, e) *= lo;
        }
}

// ---------------------------------------------------------------------------------------------------
ProblemInputs::ProblemInputs(const DomainDecomposition &dd_, const OrdinateSet &os_, int degree, int sdeg)
    : dd(dd_), os(os_), DG_degree(degree), sigma_t_density(dd_, sdeg), sigma_s_density(dd_, sdeg),
      initial_density(dd_, degree), source_density(dd_, degree), sigma_t(dd_, sdeg, 0.0), sigma_s(dd_, sdeg, 0.0),
      psi(dd_, os_, degree),
      source(dd_, os_, degree), final_time(std::numeric_limits<double>::quiet_NaN()),
      max_step_size(std::numeric_limits<double>::quiet_NaN()) {}

namespace {

bool optional_bool(const ParameterList &params, const char *key, bool fallback) {
    try {
        return params.GetBool(key);
    } catch (...) {
        return fallback;
    }
}

// The projection in `out`, read from `path` when `use_file` and the file holds the same mesh and degree,
// else computed and then written when `use_file` (init.cpp:741-860).
void projection_or_file(const Problem &prob, ScalarFlux &out, ProblemEvalFcn f, bool use_file, const std::string &path) {
    if (use_file) {
        try {
            const ScalarFlux in = ScalarFlux::ReadFromDisk(path);
            const DomainDecomposition &a = in.domain(), &b = out.domain();
            if (in.DG_degree == out.DG_degree && a.nx == b.nx && a.ny == b.ny && a.ax == b.ax && a.bx == b.bx &&
                a.ay == b.ay && a.by == b.by) {
                std::copy(in.data(), in.data() + in.size(), out.data());
                return;
            }
        } catch (const std::runtime_error &) {
        }
    }
    ComputeProjection(prob, out, f);
    if (use_file) out.WriteToDisk(path);
}

}  // namespace

ProblemInputs Init(ParameterList &params, const std::string &directory) {
    std::unique_ptr<Problem> prob = Problem::Create(params);
    prob->OverrideOptions(params);
    int reflect_mask = 0;  // init.cpp:448-510
    const char *reflect_keys[4] = {"reflect_x_min", "reflect_x_max", "reflect_y_min", "reflect_y_max"};
    for (int b = 0; b < 4; ++b)
        if (optional_bool(params, reflect_keys[b], false)) reflect_mask |= 1 << b;
    if (optional_bool(params, "periodic", false)) {
        if (reflect_mask) throw std::runtime_error("Combined reflecting and periodic boundary conditions not supported.");
        reflect_mask = TSW_PERIODIC;
    }
    const bool pwc_cross_sections = optional_bool(params, "piecewise_cross", false);

    DomainDecomposition dd;
    dd.reflect_mask = reflect_mask;
    dd.nx = params.GetInt("nx");
    dd.ax = params.GetDouble("ax");
    dd.bx = params.GetDouble("bx");
    dd.ny = params.GetInt("ny");
    dd.ay = params.GetDouble("ay");
    dd.by = params.GetDouble("by");
    if (dd.nx <= 0 || dd.ny <= 0 || !(dd.bx > dd.ax) || !(dd.by > dd.ay)) throw std::invalid_argument("bad mesh");
    int64_t degree = -1;
    try {
        degree = params.GetInt("dg_degree");
    } catch (...) {
    }
    try {
        degree = params.GetInt("dg_degree_x");
    } catch (...) {
    }
    if (degree < 0) throw std::runtime_error("Failed to determine degree of DG spatial discretization.");
    const OrdinateSet os = OrdinateSet::FromParameters(params);

    const int sigma_degree = pwc_cross_sections ? 0 : (int)degree;  // init.cpp:741
    ProblemInputs in(dd, os, (int)degree, sigma_degree);
    in.problem = prob->GetDescriptor();
    if (params.Has("tfinal")) in.final_time = params.GetDouble("tfinal");
    if (params.Has("cfl")) in.max_step_size = params.GetDouble("cfl") * dd.dx(0);

    const bool files = optional_bool(params, "output_init", false);
    char nxs[32];
    std::snprintf(nxs, sizeof nxs, "%07" PRId64, dd.nx);
    const std::string dir = directory.empty() ? std::string(".") : directory;
    const std::string ps = std::to_string(degree);
    const std::string ss = std::to_string(sigma_degree);
    projection_or_file(*prob, in.sigma_t_density, &Problem::TotalCross, files,
                       dir + "/sigma_t_dg" + ss + "_nx" + nxs + ".cs2");
    projection_or_file(*prob, in.sigma_s_density, &Problem::ScatterCross, files,
                       dir + "/sigma_s_dg" + ss + "_nx" + nxs + ".cs2");
    projection_or_file(*prob, in.initial_density, &Problem::InitialCondition, files,
                       dir + "/initial_dg" + ps + "_nx" + nxs + ".sd2");
    projection_or_file(*prob, in.source_density, &Problem::Source, files, dir + "/source_dg" + ps + "_nx" + nxs + ".sd2");
    prob->SetInflowBoundaries(in.source_density);

    for (int64_t i = 0; i <= dd.nx + 1; ++i)
        for (int64_t j = 0; j <= dd.ny + 1; ++j)
            for (int u = 0; u <= sigma_degree; ++u)
                for (int v = 0; v <= sigma_degree; ++v) {
                    in.sigma_t(i, j, u, v) = in.sigma_t_density(i, j, u, v);
                    in.sigma_s(i, j, u, v) = in.sigma_s_density(i, j, u, v);
                }
    const double fourpi = 4.0 * M_PI;
    for (int64_t i = 0; i <= dd.nx + 1; ++i)
        for (int64_t j = 0; j <= dd.ny + 1; ++j)
            for (int64_t q = 0; q < os.nq(); ++q)
                for (int d = 0; d <= degree; ++d)
                    for (int e = 0; e <= degree; ++e) {
                        const bool interior = i >= 1 && i <= dd.nx && j >= 1 && j <= dd.ny;
                        in.psi(q, i, j, d, e) = interior ? in.initial_density(i, j, d, e) / fourpi : 0.0;
                        in.source(q, i, j, d, e) = in.source_density(i, j, d, e) / fourpi;
                    }
    return in;
}

}  // namespace tsw_host
