// Field files, norms and differences of the host layer: the reference's single-rank binary format for
// DensityFunction and OrdinateFlux, and the norms and difference its post-processor reports
// (main_postproc.cpp:394-436).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <limits>
#include <memory>
#include <numeric>
#include <stdexcept>

#include "tsw_host.hpp"

namespace tsw_host {
namespace {

struct File {
    std::FILE *fp;
    File(const std::string &name, const char *mode) : fp(std::fopen(name.c_str(), mode)) {}
    ~File() {
        if (fp) std::fclose(fp);
    }
};

template <class T>
void put(std::FILE *fp, const T &v) {
    if (std::fwrite(&v, sizeof(T), 1, fp) != 1) throw std::runtime_error("write failed");
}
template <class T>
bool get(std::FILE *fp, T &v) {
    return std::fread(&v, sizeof(T), 1, fp) == 1;
}

// Mesh header, per dimension int64 nx, double ax, double bx (SpatialDomain.cpp:385-407, read :247-320).
void put_mesh(std::FILE *fp, const DomainDecomposition &dd) {
    put<int64_t>(fp, dd.nx);
    put<double>(fp, dd.ax);
    put<double>(fp, dd.bx);
    put<int64_t>(fp, dd.ny);
    put<double>(fp, dd.ay);
    put<double>(fp, dd.by);
}
bool get_mesh(std::FILE *fp, DomainDecomposition &dd) {
    return get(fp, dd.nx) && get(fp, dd.ax) && get(fp, dd.bx) && get(fp, dd.ny) && get(fp, dd.ay) && get(fp, dd.by) &&
           dd.nx > 0 && dd.ny > 0 && dd.nx < (int64_t(1) << 31) && dd.ny < (int64_t(1) << 31);
}

// DG degree, byte count, value count, values (DensityFunction.cpp:2350-2358 and OrdinateFlux.cpp:2866-2874).
void put_density(std::FILE *fp, int64_t degree, const std::vector<double> &v) {
    put<int64_t>(fp, degree);
    put<int64_t>(fp, (int64_t)(v.size() * sizeof(double)));
    put<int64_t>(fp, (int64_t)v.size());
    if (!v.empty() && std::fwrite(v.data(), sizeof(double), v.size(), fp) != v.size())
        throw std::runtime_error("write failed");
}

// Null-terminated string (WriteStringToFile / ReadStringFromFile, global.cpp:129-190).
void put_string(std::FILE *fp, const std::string &s) {
    if (std::fwrite(s.c_str(), 1, s.size() + 1, fp) != s.size() + 1) throw std::runtime_error("write failed");
}
bool get_string(std::FILE *fp, std::string &s) {
    s.clear();
    for (int c; (c = std::fgetc(fp)) != EOF;) {
        if (c == '\0') return true;
        s.push_back((char)c);
        if (s.size() > 4096) return false;
    }
    return false;
}

// Legendre P_0..P_p at x.
void legendre_all(int p, double x, double *out) {
    out[0] = 1.0;
    if (p >= 1) out[1] = x;
    for (int k = 2; k <= p; ++k) out[k] = ((2 * k - 1) * x * out[k - 1] - (k - 1) * out[k - 2]) / k;
}

std::vector<double> gl_nodes(int n, std::vector<double> *weights) {
    std::vector<double> x(n), w(n);
    if (tsw_gauss_legendre(n, x.data(), w.data()) != TSW_OK) throw std::invalid_argument("bad quadrature size");
    if (weights) *weights = w;
    return x;
}

}  // namespace

// ---------------------------------------------------------------------------------------------------
void ScalarFlux::WriteToDisk(const std::string &filename) const {
    File f(filename, "wb");
    if (!f.fp) throw std::runtime_error("Failed to open file '" + filename + "' in 'WriteToDisk'.");
    put_mesh(f.fp, dd_);
    put_density(f.fp, DG_degree, density_);
}

ScalarFlux ScalarFlux::ReadFromDisk(const std::string &filename) {
    File f(filename, "rb");
    const std::string err = "Failed to read RKDG::DensityFunction from file '" + filename + "'.";
    if (!f.fp) throw std::runtime_error(err);
    DomainDecomposition dd;
    int64_t degree, bytes, count;
    if (!get_mesh(f.fp, dd) || !get(f.fp, degree) || !get(f.fp, bytes) || !get(f.fp, count) || degree < 0 ||
        degree > 64)
        throw std::runtime_error(err);
    ScalarFlux out(dd, (int)degree);
    // the sizes must be those the header's mesh and degree imply (DensityFunction.cpp:2080-2087)
    if (count != (int64_t)out.density_.size() || bytes != count * (int64_t)sizeof(double)) throw std::runtime_error(err);
    if (std::fread(out.density_.data(), sizeof(double), out.density_.size(), f.fp) != out.density_.size())
        throw std::runtime_error(err);
    return out;
}

void OrdinateFlux::WriteToDisk(const std::string &filename) const {
    File f(filename, "wb");
    if (!f.fp) throw std::runtime_error("Failed to open file '" + filename + "' in 'WriteToDisk'.");
    WriteToDisk(f.fp);
}

void OrdinateFlux::WriteToDisk(std::FILE *fp) const {
    put_mesh(fp, dd_);
    put_string(fp, os_.type() + (os_.symmetric_reduce() ? "#R" : ""));
    put<int64_t>(fp, os_.ang_order());
    put_density(fp, DG_degree, density_);
}

bool OrdinateFlux::ReadValues(std::FILE *fp) {
    DomainDecomposition dd;
    std::string name;
    int64_t order, degree, bytes, count;
    if (!get_mesh(fp, dd) || !get_string(fp, name) || !get(fp, order) || !get(fp, degree) || !get(fp, bytes) ||
        !get(fp, count))
        return false;
    const std::string mine = os_.type() + (os_.symmetric_reduce() ? "#R" : "");
    if (dd.nx != dd_.nx || dd.ny != dd_.ny || dd.ax != dd_.ax || dd.bx != dd_.bx || dd.ay != dd_.ay ||
        dd.by != dd_.by || name != mine || order != os_.ang_order() || degree != DG_degree ||
        count != (int64_t)density_.size() || bytes != count * (int64_t)sizeof(double))
        return false;
    return std::fread(density_.data(), sizeof(double), density_.size(), fp) == density_.size();
}

// ---------------------------------------------------------------------------------------------------
// Checkpoint files (TimeIntegrator::WriteCheckpoint / ResumeCheckpoint, TimeIntegrator.cpp:745-1111,
// single rank): int64 step, double time, double solve time, the angular flux as OrdinateFlux::WriteToDisk
// writes it, then the 20 timer values the reference keeps (TMR_AF_Lsv .. TMR_io).
void WriteCheckpoint(const std::string &filename, const CheckpointState &st, const OrdinateFlux &psi) {
    File f(filename, "wb");
    if (!f.fp) throw std::runtime_error("Failed to open file '" + filename + "' in 'WriteCheckpoint'.");
    put<int64_t>(f.fp, st.step);
    put<double>(f.fp, st.time);
    put<double>(f.fp, st.solve_time);
    psi.WriteToDisk(f.fp);
    for (double v : st.timers) put<double>(f.fp, v);
}

CheckpointState ResumeCheckpoint(const std::string &filename, OrdinateFlux &psi) {
    File f(filename, "rb");
    const std::string err = "Failed to resume from checkpoint file '" + filename + "'.";
    if (!f.fp) throw std::runtime_error("Failed to open file '" + filename + "' in 'ResumeCheckpoint'.");
    CheckpointState st;
    OrdinateFlux tmp = psi;  // psi is left as it was when the file does not fit it
    if (!get(f.fp, st.step) || !get(f.fp, st.time) || !get(f.fp, st.solve_time) || !tmp.ReadValues(f.fp))
        throw std::runtime_error(err);
    for (double &v : st.timers)
        if (!get(f.fp, v)) throw std::runtime_error(err);
    if (st.step < 0 || !(st.time >= 0.0)) throw std::runtime_error(err);
    std::copy(tmp.data(), tmp.data() + tmp.size(), psi.data());
    return st;
}

OrdinateFlux OrdinateFlux::ReadFromDisk(const std::string &filename) {
    File f(filename, "rb");
    const std::string err = "Failed to read RKDG::OrdinateFlux from file '" + filename + "'.";
    if (!f.fp) throw std::runtime_error(err);
    DomainDecomposition dd;
    std::string name;
    int64_t order, degree, bytes, count;
    if (!get_mesh(f.fp, dd) || !get_string(f.fp, name) || !get(f.fp, order)) throw std::runtime_error(err);
    // "type#R" marks a symmetry-reduced set (OrdinateFlux.cpp:2577-2585)
    bool reduced = false;
    if (name.find('#') != std::string::npos) {
        reduced = name.back() == 'R';
        name.erase(name.size() - 2, 2);
    }
    if (name != "spherical-triangle" && name != "chebyshev-legendre" && name != "level-symmetric")
        throw std::out_of_range("ordinate type '" + name + "' is not available");
    const OrdinateSet os = OrdinateSet::Named(name, (int)order, reduced);
    if (!get(f.fp, degree) || !get(f.fp, bytes) || !get(f.fp, count) || degree < 0 || degree > 64)
        throw std::runtime_error(err);
    OrdinateFlux out(dd, os, (int)degree);
    if (count != (int64_t)out.density_.size() || bytes != count * (int64_t)sizeof(double)) throw std::runtime_error(err);
    if (std::fread(out.density_.data(), sizeof(double), out.density_.size(), f.fp) != out.density_.size())
        throw std::runtime_error(err);
    return out;
}

// ---------------------------------------------------------------------------------------------------
// Value of the cell polynomial at (x, y): the cell is found as DomainDecomposition::PointToCell does
// (truncation of (x - ax)/dx, DomainDecomposition.cpp:553-570), clamped so that x = bx falls in the last
// cell, and the local coordinate is 2 (x - ax)/dx - (2i + 1) (DensityFunction.cpp:418-470).
double ScalarFlux::EvaluateAtPoint(double x, double y) const {
    if (x < dd_.ax || x > dd_.bx || y < dd_.ay || y > dd_.by)
        throw std::range_error("evaluation point outside the domain");
    const double dx = dd_.dx(0), dy = dd_.dx(1);
    const int64_t i = std::min<int64_t>((int64_t)((x - dd_.ax) / dx), dd_.nx - 1);
    const int64_t j = std::min<int64_t>((int64_t)((y - dd_.ay) / dy), dd_.ny - 1);
    const double xb = 2.0 * (x - dd_.ax) / dx - (2 * i + 1), yb = 2.0 * (y - dd_.ay) / dy - (2 * j + 1);
    const int p = DG_degree;
    std::vector<double> lx(p + 1), ly(p + 1);
    legendre_all(p, xb, lx.data());
    legendre_all(p, yb, ly.data());
    double v = 0.0;
    for (int d = 0; d <= p; ++d)
        for (int e = 0; e <= p; ++e) v += (*this)(i + 1, j + 1, d, e) * lx[d] * ly[e];
    return v;
}

// Gauss-Legendre rule with DG degree + 1 points per axis on each interior cell, |f| at the points
// (DensityFunction.cpp:1082-1170).
double ScalarFlux::L1Norm() const {
    std::vector<double> w;
    const std::vector<double> x = gl_nodes(DG_degree + 1, &w);
    const int n = DG_degree + 1;
    const double dx = dd_.dx(0), dy = dd_.dx(1);
    double norm = 0.0;
    for (int64_t i = 1; i <= dd_.nx; ++i)
        for (int64_t j = 1; j <= dd_.ny; ++j)
            for (int l = 0; l < n; ++l)
                for (int m = 0; m < n; ++m) {
                    const double xl = dd_.ax + dx * (i - 0.5) + dx * x[l] / 2.0;
                    const double ym = dd_.ay + dy * (j - 0.5) + dy * x[m] / 2.0;
                    norm += w[l] * w[m] * dx * dy * std::fabs(EvaluateAtPoint(xl, ym)) / 4.0;
                }
    return norm;
}

// Orthogonality of the Legendre basis: ||P_d P_e||^2 over a cell is dx dy / ((2d+1)(2e+1))
// (DensityFunction.cpp:1175-1232).
double ScalarFlux::L2Norm() const {
    const double dx = dd_.dx(0), dy = dd_.dx(1);
    double norm = 0.0;
    for (int64_t i = 1; i <= dd_.nx; ++i)
        for (int64_t j = 1; j <= dd_.ny; ++j)
            for (int d = 0; d <= DG_degree; ++d)
                for (int e = 0; e <= DG_degree; ++e) {
                    const double c = (*this)(i, j, d, e);
                    norm += c * c * dx * dy / ((2 * d + 1) * (2 * e + 1));
                }
    return std::sqrt(norm);
}

// Per cell, the maximum of |f| on a uniform grid of 2, 4, 8, ... points per axis (end points included)
// until two successive maxima differ by at most tol relative; the largest over the cells
// (DensityFunction.cpp:1239-1400).  The grid stops growing at 4096 points per axis.
double ScalarFlux::LinfNorm(double tol) const {
    const int p = DG_degree;
    double norm = -1.0;
    std::vector<double> lx, ly;
    for (int64_t i = 1; i <= dd_.nx; ++i)
        for (int64_t j = 1; j <= dd_.ny; ++j) {
            int64_t part = 1;
            double f = std::numeric_limits<double>::max(), c;
            do {
                c = f;
                f = 0.0;
                part *= 2;
                if (p == 0) {
                    f = std::fabs((*this)(i, j, 0, 0));
                } else {
                    const double h = 2.0 / (part - 1.0);
                    lx.assign((size_t)part * (p + 1), 0.0);
                    for (int64_t l = 0; l < part; ++l) legendre_all(p, -1.0 + h * l, &lx[(size_t)l * (p + 1)]);
                    for (int64_t l = 0; l < part; ++l)
                        for (int64_t m = 0; m < part; ++m) {
                            double v = 0.0;
                            for (int d = 0; d <= p; ++d)
                                for (int e = 0; e <= p; ++e)
                                    v += (*this)(i, j, d, e) * lx[(size_t)l * (p + 1) + d] * lx[(size_t)m * (p + 1) + e];
                            f = std::max(f, std::fabs(v));
                        }
                }
            } while (std::fabs(f - c) / f > tol && part < 4096);
            norm = std::max(norm, f);
        }
    return norm;
}

// ---------------------------------------------------------------------------------------------------
ScalarFlux ScalarFlux::ComputeDifference(const ScalarFlux &first, const ScalarFlux &second) {
    const double TOL = 1e-14;
    const DomainDecomposition &A = first.dd_, &B = second.dd_;
    if (std::fabs(A.ax - B.ax) > TOL || std::fabs(A.ay - B.ay) > TOL || std::fabs(A.bx - B.bx) > TOL ||
        std::fabs(A.by - B.by) > TOL)
        throw std::invalid_argument("Domains of RKDG::DensityFunction objects in 'ComputeDifference' are not "
                                    "compatible for computing difference.");
    const int R = std::max(first.DG_degree, second.DG_degree);
    if (A.nx == B.nx && A.ny == B.ny) {
        ScalarFlux out(A, R);
        for (int64_t i = 1; i <= A.nx; ++i)
            for (int64_t j = 1; j <= A.ny; ++j) {
                for (int d = 0; d <= first.DG_degree; ++d)
                    for (int e = 0; e <= first.DG_degree; ++e) out(i, j, d, e) += first(i, j, d, e);
                for (int d = 0; d <= second.DG_degree; ++d)
                    for (int e = 0; e <= second.DG_degree; ++e) out(i, j, d, e) -= second(i, j, d, e);
            }
        return out;
    }
    DomainDecomposition rd = A;
    rd.nx = std::lcm(A.nx, B.nx);
    rd.ny = std::lcm(A.ny, B.ny);
    ScalarFlux out(rd, R);
    std::vector<double> gw;
    const std::vector<double> gx = gl_nodes(R + 1, &gw);
    const int G = R + 1, P = R + 1;
    // projection of the old basis P_d(x) on subcell s of k equal subcells onto the new basis P_a:
    // (2a+1)/2 * sum_l w_l P_a(g_l) P_d((g_l + 2s + 1)/k - 1), exact at R + 1 points
    auto project = [&](const ScalarFlux &in, int64_t k) {
        const int D = in.DG_degree + 1;
        std::vector<double> T((size_t)k * P * D, 0.0), la(P), ld(D);
        for (int64_t s = 0; s < k; ++s)
            for (int l = 0; l < G; ++l) {
                legendre_all(R, gx[l], la.data());
                legendre_all(in.DG_degree, (gx[l] + 2 * s + 1) / (double)k - 1.0, ld.data());
                for (int a = 0; a < P; ++a)
                    for (int d = 0; d < D; ++d) T[((size_t)s * P + a) * D + d] += gw[l] * (2 * a + 1) / 2.0 * la[a] * ld[d];
            }
        return T;
    };
    auto accumulate = [&](const ScalarFlux &in, double sign) {
        const int64_t kx = rd.nx / in.dd_.nx, ky = rd.ny / in.dd_.ny;
        const int D = in.DG_degree + 1;
        const std::vector<double> Tx = project(in, kx), Ty = project(in, ky);
        for (int64_t i = 0; i < rd.nx; ++i)
            for (int64_t j = 0; j < rd.ny; ++j) {
                const int64_t si = i % kx, sj = j % ky, ci = i / kx + 1, cj = j / ky + 1;
                for (int a = 0; a < P; ++a)
                    for (int b = 0; b < P; ++b) {
                        double v = 0.0;
                        for (int d = 0; d < D; ++d)
                            for (int e = 0; e < D; ++e)
                                v += Tx[((size_t)si * P + a) * D + d] * Ty[((size_t)sj * P + b) * D + e] * in(ci, cj, d, e);
                        out(i + 1, j + 1, a, b) += sign * v;
                    }
            }
    };
    accumulate(first, 1.0);
    accumulate(second, -1.0);
    return out;
}

}  // namespace tsw_host
