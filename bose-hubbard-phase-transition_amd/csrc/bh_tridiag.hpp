// bh_tridiag.hpp -- FP64 analysis of the Lanczos tridiagonal matrix on the host.
//
// T_j: diagonal a[0..j-1], off-diagonal b[1..j-1] (b[t] couples t-1 and t).  Lowest eigenvalues by
// Sturm-sequence bisection, eigenvectors by inverse iteration with a partially pivoted tridiagonal LU.
#pragma once

#include <cmath>
#include <vector>

namespace bhtri {

// number of eigenvalues of T strictly below x (negative pivots of the LDL^T of T - x I)
inline int sturm_count(int n, const double* a, const double* b, double x, double tiny) {
    int c = 0;
    double q = a[0] - x;
    if (q == 0.0) q = -tiny;
    if (q < 0) ++c;
    for (int t = 1; t < n; ++t) {
        q = (a[t] - x) - b[t] * b[t] / q;
        if (q == 0.0) q = -tiny;
        if (q < 0) ++c;
    }
    return c;
}

inline void gershgorin(int n, const double* a, const double* b, double* lo, double* hi) {
    double l = a[0], h = a[0];
    for (int t = 0; t < n; ++t) {
        const double r = (t > 0 ? std::fabs(b[t]) : 0.0) + (t + 1 < n ? std::fabs(b[t + 1]) : 0.0);
        l = std::fmin(l, a[t] - r);
        h = std::fmax(h, a[t] + r);
    }
    *lo = l;
    *hi = h;
}

// eigenvalue of index idx (0 = smallest), to about 2 eps |x|
inline double bisect(int n, const double* a, const double* b, int idx, double lo, double hi) {
    const double tiny = 1e-300 + 1e-18 * std::fmax(std::fabs(lo), std::fabs(hi));
    for (int it = 0; it < 200; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (hi - lo <= 4.0e-16 * std::fmax(std::fabs(lo), std::fabs(hi))) break;
        if (sturm_count(n, a, b, mid, tiny) > idx) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}

// inverse iteration: unit eigenvector s of T for the eigenvalue approximation theta
inline void inverse_iteration(int n, const double* a, const double* b, double theta, double tnorm, double* s) {
    std::vector<double> dl(n), d(n), du(n), du2(n), x(n);
    std::vector<int> piv(n);
    const double pert = 2.2e-16 * std::fmax(tnorm, 1e-300);
    for (int t = 0; t < n; ++t) {
        d[t] = a[t] - theta;
        if (t + 1 < n) {
            dl[t] = b[t + 1];
            du[t] = b[t + 1];
        }
        du2[t] = 0.0;
    }
    // LU with partial pivoting (row interchanges between t and t+1)
    for (int t = 0; t + 1 < n; ++t) {
        if (std::fabs(d[t]) >= std::fabs(dl[t])) {
            piv[t] = 0;
            if (d[t] == 0.0) d[t] = pert;
            const double f = dl[t] / d[t];
            dl[t] = f;
            d[t + 1] -= f * du[t];
        } else {
            piv[t] = 1;
            const double f = d[t] / dl[t];
            d[t] = dl[t];
            dl[t] = f;
            const double tmp = du[t];
            du[t] = d[t + 1];
            d[t + 1] = tmp - f * d[t + 1];
            if (t + 2 < n) {
                du2[t] = du[t + 1];
                du[t + 1] = -f * du[t + 1];
            }
        }
    }
    if (d[n - 1] == 0.0) d[n - 1] = pert;
    for (int t = 0; t < n; ++t) x[t] = 1.0 / std::sqrt(static_cast<double>(n)) * (1.0 + 0.01 * ((t * 7919) % 13));
    for (int rep = 0; rep < 3; ++rep) {
        // forward: apply L^-1 with the recorded interchanges
        for (int t = 0; t + 1 < n; ++t) {
            if (piv[t]) std::swap(x[t], x[t + 1]);
            x[t + 1] -= dl[t] * x[t];
        }
        // back substitution with U (diag d, super du, second super du2)
        for (int t = n - 1; t >= 0; --t) {
            double v = x[t];
            if (t + 1 < n) v -= du[t] * x[t + 1];
            if (t + 2 < n) v -= du2[t] * x[t + 2];
            x[t] = v / d[t];
        }
        double nrm = 0.0;
        for (int t = 0; t < n; ++t) nrm += x[t] * x[t];
        nrm = std::sqrt(nrm);
        if (!(nrm > 0.0) || !std::isfinite(nrm)) {
            for (int t = 0; t < n; ++t) x[t] = (t == 0) ? 1.0 : 0.0;
            break;
        }
        for (int t = 0; t < n; ++t) x[t] /= nrm;
    }
    // sign convention: first nonzero component positive
    double sg = 1.0;
    for (int t = 0; t < n; ++t)
        if (x[t] != 0.0) {
            sg = x[t] > 0 ? 1.0 : -1.0;
            break;
        }
    for (int t = 0; t < n; ++t) s[t] = sg * x[t];
}

// all eigenvalues of a symmetric tridiagonal matrix, ascending (implicit QL, Wilkinson shifts)
inline int ql_eigvals(int n, const double* d_in, const double* e_in, double* w) {
    std::vector<double> d(d_in, d_in + n), e(static_cast<size_t>(n) + 1, 0.0);
    for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int mm = l;
            for (; mm < n - 1; ++mm) {
                const double dd = std::fabs(d[mm]) + std::fabs(d[mm + 1]);
                if (std::fabs(e[mm]) <= 2.2e-16 * dd + 1e-300) break;
            }
            if (mm == l) break;
            if (++iter > 300) return -1;
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = std::hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            bool early = false;
            for (int i = mm - 1; i >= l; --i) {
                const double f = s * e[i], bb = c * e[i];
                r = std::hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    early = true;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * bb;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - bb;
            }
            if (early) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    for (int i = 0; i < n; ++i) w[i] = d[i];
    for (int i = 1; i < n; ++i) {
        const double x = w[i];
        int j = i - 1;
        while (j >= 0 && w[j] > x) {
            w[j + 1] = w[j];
            --j;
        }
        w[j + 1] = x;
    }
    return 0;
}

}  // namespace bhtri
