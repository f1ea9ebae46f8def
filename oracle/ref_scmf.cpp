// ref_scmf.cpp -- C entry points over the reference's mean-field functions, for the checker.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ as part of oracle/_ref/libbhref.so.
//
// Analysis::SCMF, SF_density, gap_ratios and coherence are declared `static` at namespace scope
// (include/analysis.hpp:13-16, 36, 57), so they have internal linkage: this translation unit compiles
// src/analysis.cpp itself, included from where it lies in the reference tree, and exports them.
#include "analysis.cpp"

#include <cmath>

extern "C" {

// Analysis::SCMF (src/analysis.cpp:111-178); NaN when it throws
double ref_scmf(double mu, double J, int q, double psi0, int precision) {
    try {
        return Analysis::SCMF(mu, J, q, psi0, precision);
    } catch (...) {
        return std::nan("");
    }
}

// BH::h_MF (src/hamiltonian.cpp:290-309) into out[(2p+1)^2], row-major
void ref_h_mf(double psi, int p, double mu, double J, int q, double* out) {
    const int N = 2 * p + 1;
    Eigen::MatrixXd h = Eigen::MatrixXd::Zero(N, N);
    BH::h_MF(psi, p, mu, J, q, h);
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) out[i * N + j] = h(i, j);
}

// Analysis::SF_density (src/analysis.cpp:95-105) of phi[0..2p]
double ref_sf_density(const double* phi, int p) {
    Eigen::VectorXd v(2 * p + 1);
    for (int i = 0; i < 2 * p + 1; ++i) v[i] = phi[i];
    return Analysis::SF_density(v, p);
}

// Analysis::gap_ratios (src/analysis.cpp:433-454) of ev[0..nb_eigen-1] (real); out receives nb_eigen - 2 ratios
void ref_gap_ratios(const double* ev, int nb_eigen, double* out) {
    Eigen::VectorXcd e(nb_eigen);
    for (int i = 0; i < nb_eigen; ++i) e[i] = std::complex<double>(ev[i], 0.0);
    const Eigen::VectorXd r = Analysis::gap_ratios(e, nb_eigen);
    for (Eigen::Index i = 0; i < r.size(); ++i) out[i] = r[i];
}

// Analysis::coherence (src/analysis.cpp:541-556) of a real m x m matrix (row-major)
double ref_coherence(const double* rho, int m) {
    Eigen::MatrixXcd s(m, m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) s(i, j) = std::complex<double>(rho[i * m + j], 0.0);
    return Analysis::coherence(s);
}

}  // extern "C"
