// ref_driver.cpp -- C entry points over the reference's own functions, compiled together with the
// reference sources where they lie (src/hamiltonian.cpp, src/neighbours.cpp, src/operator.cpp and
// the vendored Eigen 3.4.0 / Spectra 1.1.0 headers) into oracle/_ref/libbhref.so.
//
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and the CPU-baseline leg of bench.py as a checker.
//
// Only deterministic reference functions are exposed.  BH::fill_hopping is never called: it reads
// past the end of a one-column buffer (src/hamiltonian.cpp:180 -> :94), so its output is undefined.
// The reference's own solvers (Op::IRLM_eigen = Spectra GenEigs, Op::exact_eigen = Eigen's dense
// SelfAdjointEigenSolver) are run on a Hamiltonian handed over as CSR arrays.
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <exception>
#include <vector>

#include "hamiltonian.hpp"
#include "neighbours.hpp"
#include "operator.hpp"

namespace {

Eigen::SparseMatrix<double> to_sparse(int64_t D, const int64_t* row_ptr, const int32_t* col, const double* val,
                                      const double* diag, double J, double U) {
    std::vector<Eigen::Triplet<double>> trip;
    trip.reserve(static_cast<size_t>(row_ptr[D] + D));
    for (int64_t r = 0; r < D; ++r) {
        for (int64_t a = row_ptr[r]; a < row_ptr[r + 1]; ++a)
            trip.emplace_back(static_cast<int>(r), col[a], J * val[a]);
        if (diag && U != 0.0) trip.emplace_back(static_cast<int>(r), static_cast<int>(r), U * diag[r]);
    }
    Eigen::SparseMatrix<double> H(D, D);
    H.setFromTriplets(trip.begin(), trip.end());
    return H;
}

int flatten(const std::vector<std::vector<int>>& nb, int* off, int* sites) {
    int t = 0;
    for (size_t i = 0; i < nb.size(); ++i) {
        off[i] = t;
        for (int s : nb[i]) sites[t++] = s;
    }
    off[nb.size()] = t;
    return t;
}

}  // namespace

extern "C" {

// Neighbours::chain_neighbours (src/neighbours.cpp:21-34)
int ref_chain_neighbours(int m, int closed, int* off, int* sites) {
    Neighbours nb(m);
    nb.chain_neighbours(closed != 0);
    return flatten(nb.getNeighbours(), off, sites);
}

// Neighbours::square_neighbours (src/neighbours.cpp:40-71); -1 when it throws
int ref_square_neighbours(int m, int closed, int* off, int* sites) {
    try {
        Neighbours nb(m);
        nb.square_neighbours(closed != 0);
        return flatten(nb.getNeighbours(), off, sites);
    } catch (const std::exception&) {
        return -1;
    }
}

// BH::max_set_basis (src/hamiltonian.cpp:151-166): tags (D) and states (D x m) of every boson number 1..n;
// returns D
int ref_max_set_basis(int m, int n, double* tags, double* states) {
    auto tb = BH::max_set_basis(m, n);
    const Eigen::VectorXd& t = tb.first;
    const Eigen::MatrixXd& b = tb.second;
    for (Eigen::Index k = 0; k < b.cols(); ++k) {
        tags[k] = t[k];
        for (int i = 0; i < m; ++i) states[k * m + i] = b(i, k);
    }
    return static_cast<int>(b.cols());
}

// BH::fixed_set_basis (src/hamiltonian.cpp:143-149): tags (D) and states (D x m, row k = column k)
int ref_fixed_set_basis(int m, int n, double* tags, double* states) {
    auto tb = BH::fixed_set_basis(m, n);
    const Eigen::VectorXd& t = tb.first;
    const Eigen::MatrixXd& b = tb.second;
    for (Eigen::Index k = 0; k < b.cols(); ++k) {
        tags[k] = t[k];
        for (int i = 0; i < m; ++i) states[k * m + i] = b(i, k);
    }
    return static_cast<int>(b.cols());
}

// Diagonal of BH::fixed_bosons_hamiltonian with (J,U,mu) = (0,1,0) [which=1, fill_interaction,
// src/hamiltonian.cpp:194-211] or (0,0,1) [which=2, fill_chemical, :214-232].
int ref_diagonal_term(int m, int n, int which, double* diag) {
    Neighbours nb(m);
    nb.chain_neighbours(true);
    auto tb = BH::fixed_set_basis(m, n);
    double U = which == 1 ? 1.0 : 0.0, mu = which == 2 ? 1.0 : 0.0;
    Eigen::SparseMatrix<double> H =
        BH::fixed_bosons_hamiltonian(nb.getNeighbours(), tb.second, tb.first, m, n, 0.0, U, mu);
    for (Eigen::Index k = 0; k < H.rows(); ++k) diag[k] = H.coeff(k, k);
    return static_cast<int>(H.rows());
}

// Op::IRLM_eigen (src/operator.cpp:22-33): Spectra GenEigs, nev = nb_eigen, ncv = 2 nev + 1,
// SmallestReal, tol 1e-10.  Real parts returned ascending.  0 on success, 1 when it throws.
int ref_irlm_eigen(int64_t D, const int64_t* row_ptr, const int32_t* col, const double* val, const double* diag,
                   double J, double U, int nb_eigen, double* evals) {
    try {
        Eigen::SparseMatrix<double> H = to_sparse(D, row_ptr, col, val, diag, J, U);
        Eigen::MatrixXcd vecs;
        Eigen::VectorXcd ev = Op::IRLM_eigen(H, nb_eigen, vecs);
        std::vector<double> re(ev.size());
        for (Eigen::Index i = 0; i < ev.size(); ++i) re[i] = ev[i].real();
        std::sort(re.begin(), re.end());
        for (size_t i = 0; i < re.size(); ++i) evals[i] = re[i];
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

// The reference's exact sweep over a list of points (the grid loop of Analysis::calculate_and_save,
// src/analysis.cpp:302-314): `#pragma omp parallel for schedule(dynamic)` over the points, each assembling
// its own H = JH * J + UH * U (:311) and calling Op::IRLM_eigen(H, nb_eigen) (:314).  Outputs per point:
// rc (0 converged, 1 the solver threw: Spectra's "Eigenvalue computation failed" after its 1000 restarts),
// seconds of the solve, the nb_eigen real parts ascending.  With progress != 0 a line "point p rc seconds" goes
// to stderr as each point ends.  Returns the OpenMP threads used.
int ref_irlm_sweep(int64_t D, const int64_t* row_ptr, const int32_t* col, const double* val, const double* diag,
                   int npts, const double* J, const double* U, int nb_eigen, int* rc, double* secs, double* evals,
                   int progress) {
    const Eigen::SparseMatrix<double> JH = to_sparse(D, row_ptr, col, val, diag, 1.0, 0.0);
    const Eigen::SparseMatrix<double> UH = to_sparse(D, row_ptr, col, val, diag, 0.0, 1.0);
    int threads = 1;
#pragma omp parallel
    {
#pragma omp single
        threads = omp_get_num_threads();
    }
#pragma omp parallel for schedule(dynamic)
    for (int p = 0; p < npts; ++p) {
        const auto t0 = std::chrono::steady_clock::now();
        try {
            Eigen::SparseMatrix<double> H = JH * J[p] + UH * U[p];
            Eigen::MatrixXcd vecs;
            Eigen::VectorXcd ev = Op::IRLM_eigen(H, nb_eigen, vecs);
            std::vector<double> re(ev.size());
            for (Eigen::Index i = 0; i < ev.size(); ++i) re[i] = ev[i].real();
            std::sort(re.begin(), re.end());
            for (int i = 0; i < nb_eigen; ++i) evals[static_cast<size_t>(p) * nb_eigen + i] = re[i];
            rc[p] = 0;
        } catch (const std::exception&) {
            rc[p] = 1;
        }
        secs[p] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (progress) {
            std::fprintf(stderr, "point %d %d %.3f\n", p, rc[p], secs[p]);
            std::fflush(stderr);
        }
    }
    return threads;
}

// Op::exact_eigen (src/operator.cpp:92-101): dense SelfAdjointEigenSolver, all D eigenvalues ascending
int ref_exact_eigen(int64_t D, const int64_t* row_ptr, const int32_t* col, const double* val, const double* diag,
                    double J, double U, double* evals) {
    try {
        Eigen::SparseMatrix<double> H = to_sparse(D, row_ptr, col, val, diag, J, U);
        Eigen::MatrixXd vecs;
        Eigen::VectorXd ev = Op::exact_eigen(H, vecs);
        for (Eigen::Index i = 0; i < ev.size(); ++i) evals[i] = ev[i];
        return 0;
    } catch (const std::exception&) {
        return 1;
    }
}

}  // extern "C"
