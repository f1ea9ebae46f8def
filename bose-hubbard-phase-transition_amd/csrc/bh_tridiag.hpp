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

// Thick restart: the kept Ritz pairs (theta_i, y_i) of a Lanczos run couple to the next Lanczos vector
// only through sigma_i (H y_i = theta_i y_i + sigma_i v).  This builds an orthonormal p x p matrix Q
// (column-major, Q[i + p * c]) such that Q^T diag(theta) Q is tridiagonal (diagonal a[0..p-1],
// off-diagonal b[1..p-1], b[c] coupling c-1 and c) and the last column is sigma / ||sigma||; in the basis
// u_c = sum_i Q[i, c] y_i only u_{p-1} couples to v, with coefficient ||sigma||, so the restarted run
// continues with the ordinary three-term recurrence.  Lanczos with full reorthogonalisation on
// diag(theta) started from sigma, then reversed; after a breakdown (a kept pair that does not couple to
// v) it continues from the unit vector least represented so far, with coupling 0.
inline void arrow_to_tridiag(int p, const double* theta, const double* sigma, double* Q, double* a, double* b) {
    std::vector<double> q(static_cast<size_t>(p) * p, 0.0), al(p, 0.0), be(p + 1, 0.0), w(p);
    double ns = 0.0, tmax = 0.0;
    for (int i = 0; i < p; ++i) {
        ns += sigma[i] * sigma[i];
        tmax = std::fmax(tmax, std::fabs(theta[i]));
    }
    ns = std::sqrt(ns);
    auto fresh = [&](int col) {
        // unit vector e_m with the smallest weight in columns [0, col), orthogonalised against them
        int best = 0;
        double bw = 1e300;
        for (int m = 0; m < p; ++m) {
            double s = 0.0;
            for (int c = 0; c < col; ++c) s += q[m + static_cast<size_t>(p) * c] * q[m + static_cast<size_t>(p) * c];
            if (s < bw) {
                bw = s;
                best = m;
            }
        }
        for (int i = 0; i < p; ++i) w[i] = (i == best) ? 1.0 : 0.0;
    };
    auto orth_into = [&](int col) {
        for (int rep = 0; rep < 2; ++rep)
            for (int c = 0; c < col; ++c) {
                double d = 0.0;
                for (int i = 0; i < p; ++i) d += q[i + static_cast<size_t>(p) * c] * w[i];
                for (int i = 0; i < p; ++i) w[i] -= d * q[i + static_cast<size_t>(p) * c];
            }
        double n2 = 0.0;
        for (int i = 0; i < p; ++i) n2 += w[i] * w[i];
        n2 = std::sqrt(n2);
        for (int i = 0; i < p; ++i) q[i + static_cast<size_t>(p) * col] = n2 > 0 ? w[i] / n2 : 0.0;
        return n2;
    };
    if (ns > 0.0) {
        for (int i = 0; i < p; ++i) w[i] = sigma[i] / ns;
    } else {
        fresh(0);
    }
    orth_into(0);
    const double brk = 1e-13 * std::fmax(tmax, 1e-300);
    for (int c = 0; c < p; ++c) {
        const double* qc = &q[static_cast<size_t>(p) * c];
        for (int i = 0; i < p; ++i) w[i] = theta[i] * qc[i];
        double alpha = 0.0;
        for (int i = 0; i < p; ++i) alpha += qc[i] * w[i];
        al[c] = alpha;
        if (c + 1 == p) break;
        const double nb = orth_into(c + 1);  // w minus its components on columns 0..c, normalised
        if (nb > brk) {
            be[c + 1] = nb;
        } else {
            be[c + 1] = 0.0;
            fresh(c + 1);
            orth_into(c + 1);
        }
    }
    // reversed order: new column c is forward column p-1-c
    for (int c = 0; c < p; ++c) {
        a[c] = al[p - 1 - c];
        for (int i = 0; i < p; ++i) Q[i + static_cast<size_t>(p) * c] = q[i + static_cast<size_t>(p) * (p - 1 - c)];
    }
    b[0] = 0.0;
    for (int c = 0; c + 1 < p; ++c) b[c + 1] = be[p - 1 - c];
}

// eigen-decomposition of a small dense symmetric matrix (cyclic Jacobi): A (n x n, row-major, destroyed),
// eigenvalues w ascending, eigenvectors as the columns of V (row-major V[i * n + c])
inline void jacobi_eig(int n, double* A, double* w, double* V) {
    for (int i = 0; i < n; ++i)
        for (int c = 0; c < n; ++c) V[i * n + c] = (i == c) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < n; ++c) {
                tot += A[i * n + c] * A[i * n + c];
                if (i != c) off += A[i * n + c] * A[i * n + c];
            }
        if (off <= 1e-30 * tot || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
                const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int r = 0; r < n; ++r) {
                    const double arp = A[r * n + p], arq = A[r * n + q];
                    A[r * n + p] = c * arp - s * arq;
                    A[r * n + q] = s * arp + c * arq;
                }
                for (int r = 0; r < n; ++r) {
                    const double apr = A[p * n + r], aqr = A[q * n + r];
                    A[p * n + r] = c * apr - s * aqr;
                    A[q * n + r] = s * apr + c * aqr;
                }
                for (int r = 0; r < n; ++r) {
                    const double vrp = V[r * n + p], vrq = V[r * n + q];
                    V[r * n + p] = c * vrp - s * vrq;
                    V[r * n + q] = s * vrp + c * vrq;
                }
            }
    }
    std::vector<int> ord(n);
    for (int i = 0; i < n; ++i) ord[i] = i;
    std::vector<double> d(n);
    for (int i = 0; i < n; ++i) d[i] = A[i * n + i];
    for (int i = 1; i < n; ++i)
        for (int j = i; j > 0 && d[ord[j]] < d[ord[j - 1]]; --j) std::swap(ord[j], ord[j - 1]);
    std::vector<double> Vs(static_cast<size_t>(n) * n);
    for (int c = 0; c < n; ++c) {
        w[c] = d[ord[c]];
        for (int r = 0; r < n; ++r) Vs[r * n + c] = V[r * n + ord[c]];
    }
    for (int i = 0; i < n * n; ++i) V[i] = Vs[i];
}

// all eigenvalues of a symmetric tridiagonal matrix, ascending: the implicit QL algorithm with Wilkinson
// shifts of Bowdler, Martin, Reinsch and Wilkinson (Numer. Math. 11, 293-306, 1968), the procedure tql1 of
// the Handbook for Automatic Computation vol. II and of EISPACK, whose step sequence (shift g, rotation
// parameters r, s, c, p, the search for a negligible off-diagonal element) the loop below follows
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
