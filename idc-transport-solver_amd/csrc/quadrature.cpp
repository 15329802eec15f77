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
