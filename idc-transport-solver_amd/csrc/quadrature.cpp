// Level-symmetric (LQ_N) ordinate sets reduced to the 2D hemisphere.
//
// The reference builds Chebyshev-Legendre, T_N and Lebedev sets (OrdinateSet.cpp:40-91) and reduces
// them to 2D by merging +-mu_z pairs (SymmetricReduce, :1171-1190); it has no level-symmetric set.
// Construction here (Lewis & Miller, Computational Methods of Neutron Transport, 1984, ch. 4):
//   mu_i^2 = mu_1^2 + (i-1) * 2(1 - 3 mu_1^2)/(N-2),  i = 1..N/2, mu_1 from the published table;
//   octant points (mu_a, mu_b, mu_c) with a+b+c = N/2+2; points whose index triples are
//   permutations of each other share one weight ("weight class");
//   class weights solve the even moment conditions sum_w mu^(2k) = 1/(2k+1), k = 0,2,..,N/2-1
//   (k = 1 follows from mu_a^2+mu_b^2+mu_c^2 = 1).  For N <= 12 the system is square.  For N = 16 it
//   leaves one free parameter; the weights are those closest (l2) to the published LQ16 table that
//   satisfy the conditions exactly.
// The 2D set keeps the mu_c > 0 half of the sphere with doubled weights: nq = N(N+2)/2 directions,
// weights summing to 4 pi, grouped by quadrant 0:(+,+) 1:(+,-) 2:(-,+) 3:(-,-) as in
// OrdinateSet::SetSigns (OrdinateSet.cpp:658-700).
#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "tsw.h"

namespace {

double mu1_for(int order) {
    switch (order) {
        case 2: return 0.5773503;
        case 4: return 0.3500212;
        case 6: return 0.2666355;
        case 8: return 0.2182179;
        case 10: return 0.1893213;
        case 12: return 0.1672126;
        case 16: return 0.1389568;
        default: return -1.0;
    }
}

// Dense solve with partial pivoting (host, tiny systems).
std::vector<double> solve_dense(std::vector<std::vector<double>> A, std::vector<double> b) {
    const int n = (int)b.size();
    for (int k = 0; k < n; ++k) {
        int piv = k;
        for (int r = k + 1; r < n; ++r)
            if (std::fabs(A[r][k]) > std::fabs(A[piv][k])) piv = r;
        std::swap(A[k], A[piv]);
        std::swap(b[k], b[piv]);
        if (A[k][k] == 0.0) throw std::runtime_error("singular moment system");
        for (int r = k + 1; r < n; ++r) {
            const double f = A[r][k] / A[k][k];
            for (int c = k; c < n; ++c) A[r][c] -= f * A[k][c];
            b[r] -= f * b[k];
        }
    }
    std::vector<double> x(n);
    for (int k = n - 1; k >= 0; --k) {
        double s = b[k];
        for (int c = k + 1; c < n; ++c) s -= A[k][c] * x[c];
        x[k] = s / A[k][k];
    }
    return x;
}

using Triple = std::array<int, 3>;

Triple sorted_desc(Triple t) {
    std::sort(t.begin(), t.end(), [](int a, int b) { return a > b; });
    return t;
}

}  // namespace

extern "C" int tsw_ordinates_level_symmetric(int order, double *xi, double *eta, double *w, int *quad_off) {
    const double m1 = mu1_for(order);
    if (m1 < 0 || !xi || !eta || !w || !quad_off) return TSW_E_ARG;
    try {
        const int n = order / 2;
        std::vector<double> mu(n);
        if (order == 2) {
            mu[0] = m1;
        } else {
            const double delta = 2.0 * (1.0 - 3.0 * m1 * m1) / (order - 2);
            for (int i = 0; i < n; ++i) mu[i] = std::sqrt(m1 * m1 + i * delta);
        }
        // octant points and their weight classes
        std::vector<Triple> pts;
        for (int a = 1; a <= n; ++a)
            for (int b = 1; b <= n; ++b) {
                const int c = n + 2 - a - b;
                if (c >= 1 && c <= n) pts.push_back({a, b, c});
            }
        std::map<Triple, int> cls;
        for (auto &t : pts) cls.emplace(sorted_desc(t), 0);
        int nc = 0;
        for (auto &kv : cls) kv.second = nc++;
        std::vector<int> mult(nc, 0);
        for (auto &t : pts) ++mult[cls[sorted_desc(t)]];

        // moment rows: k = 0, 2, 3, .., n-1
        std::vector<std::vector<double>> A;
        std::vector<double> rhs;
        for (int k = 0; k < n; ++k) {
            if (k == 1) continue;
            std::vector<double> row(nc, 0.0);
            for (auto &t : pts) row[cls[sorted_desc(t)]] += std::pow(mu[t[0] - 1], 2 * k);
            A.push_back(row);
            rhs.push_back(1.0 / (2 * k + 1));
        }
        std::vector<double> wc;
        if ((int)A.size() == nc) {
            wc = solve_dense(A, rhs);
        } else if (order == 16) {
            // published LQ16 point weights by class (a>=b>=c index triples)
            const std::map<Triple, double> tab = {{{8, 1, 1}, 0.0489872}, {{7, 2, 1}, 0.0413296},
                                                  {{6, 3, 1}, 0.0212326}, {{6, 2, 2}, 0.0360486},
                                                  {{5, 4, 1}, 0.0256207}, {{5, 3, 2}, 0.0144627},
                                                  {{4, 4, 2}, 0.0344001}, {{4, 3, 3}, 0.0085448}};
            std::vector<double> w0(nc);
            for (auto &kv : cls) w0[kv.second] = tab.at(kv.first);
            // w = w0 + A^T y,  (A A^T) y = rhs - A w0
            const int m = (int)A.size();
            std::vector<std::vector<double>> AAt(m, std::vector<double>(m, 0.0));
            std::vector<double> res(m);
            for (int r = 0; r < m; ++r) {
                double s = 0.0;
                for (int c = 0; c < nc; ++c) s += A[r][c] * w0[c];
                res[r] = rhs[r] - s;
                for (int r2 = 0; r2 < m; ++r2) {
                    double t = 0.0;
                    for (int c = 0; c < nc; ++c) t += A[r][c] * A[r2][c];
                    AAt[r][r2] = t;
                }
            }
            const std::vector<double> y = solve_dense(AAt, res);
            wc = w0;
            for (int c = 0; c < nc; ++c)
                for (int r = 0; r < m; ++r) wc[c] += A[r][c] * y[r];
        } else {
            return TSW_E_ARG;
        }
        for (double v : wc)
            if (!(v > 0.0)) return TSW_E_ARG;

        // 2D hemisphere: quadrants in reference order, points ordered by (a,b)
        const int sx[4] = {+1, +1, -1, -1}, sy[4] = {+1, -1, +1, -1};
        int q = 0;
        for (int quad = 0; quad < 4; ++quad) {
            quad_off[quad] = q;
            for (auto &t : pts) {
                xi[q] = sx[quad] * mu[t[0] - 1];
                eta[q] = sy[quad] * mu[t[1] - 1];
                w[q] = M_PI * wc[cls[sorted_desc(t)]];  // octant sums to 1 -> sphere 4 pi, hemisphere doubled
                ++q;
            }
        }
        quad_off[4] = q;
        return TSW_OK;
    } catch (...) {
        return TSW_E_ARG;
    }
}

// ---------------------------------------------------------------------------------------------------
// Gauss-Legendre rules and the reference's named ordinate sets.
namespace {

// Legendre P_n and P_n' at x by the three-term recurrence, in quadruple precision (__float128,
// software arithmetic on the host).
using quad_t = __float128;
void legendre_q(int n, quad_t x, quad_t &p, quad_t &dp) {
    quad_t p0 = 1, p1 = x;
    if (n == 0) {
        p = 1;
        dp = 0;
        return;
    }
    for (int k = 2; k <= n; ++k) {
        const quad_t pk = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
        p0 = p1;
        p1 = pk;
    }
    p = p1;
    dp = n * (x * p1 - p0) / (x * x - 1);
}

// Nodes ascending, exactly antisymmetric (x[n-1-i] = -x[i], middle node 0), and weights; Newton in
// quadruple precision from the Tricomi guess, rounded once to double.  The reference tabulates these
// rules to 25 digits (Tabulated.cpp:29-1092); rounding the quadruple-precision root gives the same
// doubles, the correctly rounded roots.
void gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n / 2; ++i) {
        quad_t r = std::cos(M_PI * (i + 0.75) / (n + 0.5)), p, dp;
        for (int it = 0; it < 100; ++it) {
            legendre_q(n, r, p, dp);
            const quad_t dx = p / dp;
            r -= dx;
            if (std::fabs((double)dx) < 1e-30) break;
        }
        legendre_q(n, r, p, dp);
        const double xr = (double)r, wr = (double)(2 / ((1 - r * r) * dp * dp));
        x[n - 1 - i] = xr;
        x[i] = -xr;
        w[n - 1 - i] = w[i] = wr;
    }
    if (n % 2) {
        quad_t p, dp;
        legendre_q(n, 0, p, dp);
        x[n / 2] = 0.0;
        w[n / 2] = (double)(2 / (dp * dp));
    }
}

struct Dir {
    double xi, eta, mu, w;
};

// Octant code (xi > 0, eta > 0, mu > 0) as bits 2..0, the reference's labelling of octants
// (OrdinateSet.cpp:878-939) for directions with no zero component.
unsigned octant_of(const Dir &d) { return (d.xi > 0) << 2 | (d.eta > 0) << 1 | (d.mu > 0); }

// Chebyshev-Legendre product set (OrdinateSet.cpp:441-488): GL nodes in mu, 2N equally spaced azimuths
// phi = (k + 1/2) pi / N, weight pi/N * w_GL; ordered by descending (octant, xi, eta, mu), which is what
// the reference's heap sort with a min-heap produces (:1065-1090).
std::vector<Dir> chebyshev_legendre(int n) {
    std::vector<double> gx(n), gw(n);
    gauss_legendre(n, gx.data(), gw.data());
    std::vector<Dir> d;
    for (int a = 0; a < n; ++a)
        for (int k = 0; k < 2 * n; ++k) {
            const double phi = (k + 0.5) * M_PI / n, mu = gx[a], s = std::sqrt(1.0 - mu * mu);
            d.push_back({s * std::cos(phi), s * std::sin(phi), mu, M_PI / n * gw[a]});
        }
    std::sort(d.begin(), d.end(), [](const Dir &a, const Dir &b) {
        const unsigned oa = octant_of(a), ob = octant_of(b);
        if (oa != ob) return oa > ob;
        if (a.xi != b.xi) return a.xi > b.xi;
        if (a.eta != b.eta) return a.eta > b.eta;
        return a.mu > b.mu;
    });
    return d;
}

// T_N spherical-triangle set (SphericalTriangle.cpp:268-351): the octant face x + y + z = 1, written in
// coordinates (u, v) with x = 1/2 - u - v/2, y = 1/2 + u - v/2, z = v, is cut into N^2 triangles; each
// gives the direction of its centroid and the solid angle of the spherical triangle on its vertices.
struct Tri {
    double u[3], v[3];
};

void xyz_of(double u, double v, double out[3]) {
    const double x = 0.5 - u - 0.5 * v, y = 0.5 + u - 0.5 * v, z = v;
    const double r = std::sqrt(x * x + y * y + z * z);
    out[0] = x / r;
    out[1] = y / r;
    out[2] = z / r;
}

// Solid angle as the spherical excess A + B + C - pi, the angles from the half-angle formulas of
// spherical trigonometry (tan(A/2) = tan(r) / sin(s - a), tan(r) the tangent of the inradius).
double solid_angle(const Tri &t) {
    double p[3][3];
    for (int k = 0; k < 3; ++k) xyz_of(t.u[k], t.v[k], p[k]);
    auto arc = [&](int i, int j) { return std::acos(p[i][0] * p[j][0] + p[i][1] * p[j][1] + p[i][2] * p[j][2]); };
    const double a = arc(1, 2), b = arc(0, 2), c = arc(0, 1), s = 0.5 * (a + b + c);
    const double tr = std::sqrt(std::sin(s - a) * std::sin(s - b) * std::sin(s - c) / std::sin(s));
    return 2.0 * std::atan(tr / std::sin(s - a)) + 2.0 * std::atan(tr / std::sin(s - b)) +
           2.0 * std::atan(tr / std::sin(s - c)) - M_PI;
}

// Power-of-two N: the base triangle split into four by its edge midpoints, recursively, children in
// the order (v1, m12, m13), (m12, v2, m23), (m23, m12, m13), (m13, m23, v3) (:180-230).
void subdivide(const Tri &t, int level, std::vector<Tri> &out) {
    if (level == 0) {
        out.push_back(t);
        return;
    }
    const double m12u = 0.5 * (t.u[0] + t.u[1]), m12v = 0.5 * (t.v[0] + t.v[1]);
    const double m13u = 0.5 * (t.u[0] + t.u[2]), m13v = 0.5 * (t.v[0] + t.v[2]);
    const double m23u = 0.5 * (t.u[1] + t.u[2]), m23v = 0.5 * (t.v[1] + t.v[2]);
    subdivide({{t.u[0], m12u, m13u}, {t.v[0], m12v, m13v}}, level >> 1, out);
    subdivide({{m12u, t.u[1], m23u}, {m12v, t.v[1], m23v}}, level >> 1, out);
    subdivide({{m23u, m12u, m13u}, {m23v, m12v, m13v}}, level >> 1, out);
    subdivide({{m13u, m23u, t.u[2]}, {m13v, m23v, t.v[2]}}, level >> 1, out);
}

std::vector<Dir> spherical_triangle(int n) {
    std::vector<Tri> tris;
    if ((n & (n - 1)) == 0) {
        subdivide({{0.0, -0.5, 0.5}, {1.0, 0.0, 0.0}}, n >> 1, tris);
    } else {
        // rows r = 0..N-1 from the apex, 2r + 1 triangles each, alternately pointing up and down
        const double h = 1.0 / n;
        for (int r = 0; r < n; ++r)
            for (int k = 0; k < 2 * r + 1; ++k) {
                const double top = 1.0 - r * h, bot = 1.0 - (r + 1) * h;
                if (k % 2)
                    tris.push_back({{(double)(k - r + 1) / (2 * n), (double)(k - r - 1) / (2 * n),
                                     (double)(k - r) / (2 * n)},
                                    {top, top, bot}});
                else
                    tris.push_back({{(double)(k - r) / (2 * n), (double)(k - r - 1) / (2 * n),
                                     (double)(k - r + 1) / (2 * n)},
                                    {top, bot, bot}});
            }
    }
    const size_t m = tris.size();
    std::vector<Dir> d(8 * m);
    for (size_t q = 0; q < m; ++q) {
        const Tri &t = tris[q];
        double c[3];
        xyz_of((t.u[0] + t.u[1] + t.u[2]) / 3.0, (t.v[0] + t.v[1] + t.v[2]) / 3.0, c);
        d[q] = {c[0], c[1], c[2], solid_angle(t)};
    }
    // the other octants by reflection: mu, then eta, then xi (FillByOctahedralReflection, :790-815)
    for (size_t q = 0; q < m; ++q) d[m + q] = {d[q].xi, d[q].eta, -d[q].mu, d[q].w};
    for (size_t q = 0; q < 2 * m; ++q) d[2 * m + q] = {d[q].xi, -d[q].eta, d[q].mu, d[q].w};
    for (size_t q = 0; q < 4 * m; ++q) d[4 * m + q] = {-d[q].xi, d[q].eta, d[q].mu, d[q].w};
    return d;
}

// 2D reduction (SymmetricReduce, OrdinateSet.cpp:1171-1190): each direction absorbs the weight of the
// first later direction with the same (xi, eta) and opposite mu, which is then removed.
void symmetric_reduce(std::vector<Dir> &d) {
    for (size_t i = 0; i < d.size(); ++i)
        for (size_t j = i + 1; j < d.size(); ++j)
            if (d[i].xi == d[j].xi && d[i].eta == d[j].eta && d[i].mu == -d[j].mu) {
                d[i].w += d[j].w;
                d.erase(d.begin() + (ptrdiff_t)j);
                break;
            }
}

}  // namespace

extern "C" int tsw_gauss_legendre(int n, double *nodes, double *weights) {
    if (n < 1 || n > 4096 || !nodes || !weights) return TSW_E_ARG;
    gauss_legendre(n, nodes, weights);
    return TSW_OK;
}

extern "C" int tsw_ordinates(const char *type, int order, int symmetric_reduce, int capacity, int *nq, double *xi,
                             double *eta, double *w, int *quad_off) {
    if (!type || !nq || order < 1 || order > 256) return TSW_E_ARG;
    try {
        const std::string t(type);
        std::vector<Dir> d;
        if (t == "level-symmetric") {
            const int n = order * (order + 2) / 2;
            std::vector<double> a(n), b(n), c(n);
            int off[5];
            if (order % 2 || tsw_ordinates_level_symmetric(order, a.data(), b.data(), c.data(), off) != TSW_OK)
                return TSW_E_ARG;
            *nq = n;
            if (xi) {
                if (!eta || !w || !quad_off || capacity < n) return TSW_E_ARG;
                std::copy(a.begin(), a.end(), xi);
                std::copy(b.begin(), b.end(), eta);
                std::copy(c.begin(), c.end(), w);
                std::copy(off, off + 5, quad_off);
            }
            return TSW_OK;
        } else if (t == "chebyshev-legendre") {
            // odd N has a node at mu = 0, whose octant the reference takes from an azimuth shifted by pi
            // (OrdinateSet.cpp:466,893-905), so only even N is accepted
            if (order % 2) return TSW_E_ARG;
            d = chebyshev_legendre(order);
        } else if (t == "spherical-triangle") {
            d = spherical_triangle(order);
        } else {
            return TSW_E_ARG;
        }
        if (symmetric_reduce) ::symmetric_reduce(d);
        // octants must come in the reference's order 111, 110, ..., 000; quadrant k of the 2D sweep is the
        // pair of octants with the same (xi, eta) signs
        int off[5] = {0, 0, 0, 0, 0};
        unsigned prev = 8;
        for (size_t q = 0; q < d.size(); ++q) {
            const unsigned o = octant_of(d[q]);
            if (o > prev) return TSW_E_ARG;
            prev = o;
        }
        for (int quad = 0; quad < 4; ++quad) {
            int q = 0;
            while (q < (int)d.size() && (int)(7 - octant_of(d[q])) / 2 < quad) ++q;
            off[quad] = q;
        }
        off[4] = (int)d.size();
        *nq = (int)d.size();
        if (xi) {
            if (!eta || !w || !quad_off || capacity < *nq) return TSW_E_ARG;
            for (size_t q = 0; q < d.size(); ++q) {
                xi[q] = d[q].xi;
                eta[q] = d[q].eta;
                w[q] = d[q].w;
            }
            std::copy(off, off + 5, quad_off);
        }
        return TSW_OK;
    } catch (...) {
        return TSW_E_ARG;
    }
}

// ---------------------------------------------------------------------------------------------------
// IDC quadrature nodes and integration matrices (IDCIntegrator::SetNodes, idc.cpp:2305-2376, over
// Quadrule::ComputeQuadratureNodes, Quadrule.cpp:1077-1145, and ComputeQuadratureWeights).
namespace {

using qreal = __float128;

// Newton's method in quadruple precision: f(t, v, dv) evaluates the function and its derivative.
template <typename F>
qreal newton_q(F &&f, qreal r) {
    for (int it = 0; it < 100; ++it) {
        qreal v, dv;
        f(r, v, dv);
        const qreal dx = v / dv;
        r -= dx;
        if (std::fabs((double)dx) < 1e-30) break;
    }
    return r;
}

// The m nodes of a family on [-1, 1], ascending, in quadruple precision:
//   Gauss-Lobatto: -1, the roots of P'_{m-1}, 1 (P'' from the Legendre equation);
//   Gauss-Legendre: the roots of P_m;
//   Gauss-Radau IIA: the roots of P_m - P_{m-1} (Quadrule::Radau, Quadrule.cpp:412-446, is this
//     polynomial up to a factor), the last of which is 1; Newton from the Chebyshev-Radau points;
//   Chebyshev cos((2i+1) pi / 2m) and Chebyshev-Radau cos(2 i pi / (2m - 1)), i = m-1..0, evaluated in
//     double as the reference does.
std::vector<qreal> idc_family(int kind, int m) {
    std::vector<qreal> x(m);
    switch (kind) {
    case TSW_IDC_LOBATTO: {
        const int n = m - 1;
        x[0] = -1;
        x[m - 1] = 1;
        for (int k = 1; k < m - 1; ++k)
            x[k] = newton_q([n](qreal t, qreal &v, qreal &dv) {
                qreal p, dp;
                legendre_q(n, t, p, dp);
                v = dp;
                dv = (2 * t * dp - qreal(n) * (n + 1) * p) / (1 - t * t);
            }, -std::cos(M_PI * k / n));
        break;
    }
    case TSW_IDC_LEGENDRE:
        for (int k = 0; k < m; ++k)
            x[k] = newton_q([m](qreal t, qreal &v, qreal &dv) { legendre_q(m, t, v, dv); },
                            -std::cos(M_PI * (k + 0.75) / (m + 0.5)));
        break;
    case TSW_IDC_RADAU:
        x[m - 1] = 1;
        for (int k = 0; k < m - 1; ++k)
            x[k] = newton_q([m](qreal t, qreal &v, qreal &dv) {
                qreal p1, dp1, p0, dp0;
                legendre_q(m, t, p1, dp1);
                legendre_q(m - 1, t, p0, dp0);
                v = p1 - p0;
                dv = dp1 - dp0;
            }, std::cos(2.0 * M_PI * (m - 1 - k) / (2 * m - 1)));
        break;
    case TSW_IDC_CHEBYSHEV:
        for (int i = 0; i < m; ++i) x[m - 1 - i] = std::cos((double)(2 * i + 1) * M_PI / (2 * m));
        break;
    case TSW_IDC_CHEBYSHEV_RADAU:
        for (int i = 0; i < m; ++i) x[m - 1 - i] = std::cos(2 * i * M_PI / (2 * m - 1));
        break;
    default:
        throw std::invalid_argument("node type");
    }
    return x;
}

// integrals over [a, b] of the Lagrange basis polynomials of `t` (degree m-1), by an (m+1)-point
// Gauss-Legendre rule (exact for degree 2m+1), in quadruple precision
std::vector<qreal> lagrange_integrals_q(const std::vector<qreal> &t, qreal a, qreal b) {
    const int m = (int)t.size(), ng = m + 1;
    const std::vector<qreal> xg = idc_family(TSW_IDC_LEGENDRE, ng);
    std::vector<qreal> out(m, 0);
    for (int k = 0; k < ng; ++k) {
        qreal p, dp;
        legendre_q(ng, xg[k], p, dp);
        const qreal wg = 2 / ((1 - xg[k] * xg[k]) * dp * dp);
        const qreal s = (b - a) / 2 * (xg[k] + 1) + a;
        for (int l = 0; l < m; ++l) {
            qreal v = 1;
            for (int j = 0; j < m; ++j)
                if (j != l) v *= (s - t[j]) / (t[l] - t[j]);
            out[l] += (b - a) / 2 * wg * v;
        }
    }
    return out;
}

}  // namespace

extern "C" int tsw_idc_nodes(int nodes_type, int num_nodes, int capacity, int *num_stages, double *nodes, double *W) {
    const bool left = nodes_type == TSW_IDC_LOBATTO;
    if (nodes_type < TSW_IDC_LOBATTO || nodes_type > TSW_IDC_CHEBYSHEV_RADAU || !num_stages) return TSW_E_ARG;
    if (num_nodes < (left ? 2 : 1) || num_nodes > 12) return TSW_E_ARG;
    const int S = num_nodes + (left ? 0 : 1), off = left ? 0 : 1;
    *num_stages = S;
    if (!nodes) return TSW_OK;
    if (!W || capacity < S) return TSW_E_ARG;
    try {
        std::vector<qreal> t = idc_family(nodes_type, num_nodes);
        for (qreal &v : t) v = (v + 1) / 2;  // to [0, 1]
        std::vector<qreal> st(S, 0);
        std::copy(t.begin(), t.end(), st.begin() + off);  // stage 0 is t = 0 where no node lies there
        for (int i = 0; i < S; ++i) nodes[i] = (double)st[i];
        std::fill(W, W + (size_t)S * S, 0.0);
        for (int n = 0; n < S; ++n) {
            const std::vector<qreal> row = lagrange_integrals_q(t, n == 0 ? qreal(0) : st[n - 1], n == 0 ? qreal(1) : st[n]);
            for (int l = 0; l < num_nodes; ++l) W[(size_t)n * S + off + l] = (double)row[l];
        }
        return TSW_OK;
    } catch (...) {
        return TSW_E_ARG;
    }
}
