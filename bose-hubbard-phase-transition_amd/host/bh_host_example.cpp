// bh_host_example.cpp -- the reference's call sequence through the C++ host surface, as a caller writes it:
//   Neighbours -> BH::fixed_set_basis -> BH::fixed_bosons_hamiltonian (one term each) -> H = JH*J + UH*U
//   + uH*mu -> Op::IRLM_eigen (and Op::exact_eigen for small sectors) -> Analysis::SPDM, condensate fraction,
//   coherence.
// Prints one JSON object: dimension, the first tags, the CSR size of H, the Lanczos levels, the ground
// state's density-matrix observables and, when the sector is small, the lowest exact eigenvalues.
//   bhx-host-example M N J U MU [K]
// With --max, the variable-boson-number calls instead: BH::max_set_basis(M, NMAX) (every tag and state) and
// Op::IRLM_eigen of BH::max_bosons_hamiltonian(neighbours, M, NMIN, NMAX, J, U, MU) (its size and levels).
//   bhx-host-example --max M NMIN NMAX J U MU [K]
// With --hmf, BH::h_MF(psi, p, mu, J, q) as a dense (2p+1)^2 matrix and Analysis::SF_density of the vector
// v_i = 1 / (i + 1) (host only, no device); with --scmf, Analysis::SCMF(mu, J, q, psi0, precision).
//   bhx-host-example --hmf PSI P MU J Q
//   bhx-host-example --scmf MU J Q PSI0 PRECISION
// With --folm, Op::FOLM_eigen(H, NITER) of the same H: its values, the Lanczos steps taken and the residual
// norm |H v - e v| of each returned vector.
//   bhx-host-example --folm M N J U MU NITER
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "bh_host.hpp"

static int run_max(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: bhx-host-example --max M NMIN NMAX J U MU [K]\n");
        return 1;
    }
    const int m = std::atoi(argv[2]), nmin = std::atoi(argv[3]), nmax = std::atoi(argv[4]);
    const double J = std::atof(argv[5]), U = std::atof(argv[6]), mu = std::atof(argv[7]);
    const int k = argc > 8 ? std::atoi(argv[8]) : 4;
    try {
        Neighbours nb(m);
        nb.chain_neighbours();
        auto [tags, basis] = BH::max_set_basis(m, nmax);
        std::printf("{\"basis_D\": %lld, \"tags\": [", static_cast<long long>(basis.cols()));
        for (size_t i = 0; i < tags.size(); ++i) std::printf("%s%.17g", i ? ", " : "", tags[i]);
        std::printf("], \"states\": [");
        for (int64_t c = 0; c < basis.cols(); ++c)
            for (int i = 0; i < m; ++i) std::printf("%s%g", c || i ? ", " : "", basis(i, c));
        const BH::SparseMatrix H = BH::max_bosons_hamiltonian(nb.getNeighbours(), m, nmin, nmax, J, U, mu);
        std::vector<int64_t> rp;
        std::vector<int32_t> col;
        std::vector<double> val;
        H.to_csr(rp, col, val);
        bhx::MatrixXcd vecs;
        const auto ev = Op::IRLM_eigen(H, k, vecs);
        std::printf("], \"D\": %lld, \"blocks\": %zu, \"nnz\": %lld, \"lanczos\": [", static_cast<long long>(H.rows()),
                    H.blocks.size(), static_cast<long long>(rp.back()));
        for (size_t i = 0; i < ev.size(); ++i) std::printf("%s%.17g", i ? ", " : "", ev[i].real());
        // residual norms |H v - e v| of the returned eigenvectors, through the materialised CSR
        std::printf("], \"residual\": [");
        for (size_t q = 0; q < ev.size(); ++q) {
            double r2 = 0.0, v2 = 0.0;
            for (int64_t r = 0; r + 1 < static_cast<int64_t>(rp.size()); ++r) {
                double hv = 0.0;
                for (int64_t e = rp[r]; e < rp[r + 1]; ++e) hv += val[e] * vecs(col[e], q).real();
                const double d = hv - ev[q].real() * vecs(r, q).real();
                r2 += d * d;
                v2 += vecs(r, q).real() * vecs(r, q).real();
            }
            std::printf("%s%.6g", q ? ", " : "", std::sqrt(r2 / v2));
        }
        std::printf("]}\n");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}

static int run_hmf(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: bhx-host-example --hmf PSI P MU J Q\n");
        return 1;
    }
    const double psi = std::atof(argv[2]), mu = std::atof(argv[4]), J = std::atof(argv[5]);
    const int p = std::atoi(argv[3]), q = std::atoi(argv[6]);
    try {
        bhx::MatrixXd h;
        BH::h_MF(psi, p, mu, J, q, h);
        const int N = 2 * p + 1;
        std::vector<double> v(N);
        for (int i = 0; i < N; ++i) v[i] = 1.0 / (i + 1);
        std::printf("{\"h\": [");
        for (int i = 0; i < N; ++i)
            for (int c = 0; c < N; ++c) std::printf("%s%.17g", i || c ? ", " : "", h(i, c));
        std::printf("], \"sf_density\": %.17g}\n", Analysis::SF_density(v, p));
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}

static int run_scmf(int argc, char** argv) {
    if (argc < 7) {
        std::fprintf(stderr, "usage: bhx-host-example --scmf MU J Q PSI0 PRECISION\n");
        return 1;
    }
    try {
        const double psi = Analysis::SCMF(std::atof(argv[2]), std::atof(argv[3]), std::atoi(argv[4]), std::atof(argv[5]),
                                          std::atoi(argv[6]));
        std::printf("{\"psi\": %.17g}\n", psi);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}

static int run_folm(int argc, char** argv) {
    if (argc < 8) {
        std::fprintf(stderr, "usage: bhx-host-example --folm M N J U MU NITER\n");
        return 1;
    }
    const int m = std::atoi(argv[2]), n = std::atoi(argv[3]), nit = std::atoi(argv[7]);
    const double J = std::atof(argv[4]), U = std::atof(argv[5]), mu = std::atof(argv[6]);
    try {
        Neighbours nb(m);
        nb.chain_neighbours();
        const auto nei = nb.getNeighbours();
        auto [tags, basis] = BH::fixed_set_basis(m, n);
        const BH::SparseMatrix H = BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 1, 0, 0) * J +
                                   BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 0, 1, 0) * U +
                                   BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 0, 0, 1) * mu;
        std::vector<int64_t> rp;
        std::vector<int32_t> col;
        std::vector<double> val;
        H.to_csr(rp, col, val);
        bhx::MatrixXd vecs;
        const auto ev = Op::FOLM_eigen(H, nit, vecs);
        std::printf("{\"steps\": %zu, \"values\": [", ev.size());
        for (size_t i = 0; i < ev.size(); ++i) std::printf("%s%.17g", i ? ", " : "", ev[i]);
        std::printf("], \"residual\": [");
        for (size_t q = 0; q < ev.size(); ++q) {
            double r2 = 0.0;
            for (int64_t r = 0; r + 1 < static_cast<int64_t>(rp.size()); ++r) {
                double hv = 0.0;
                for (int64_t e = rp[r]; e < rp[r + 1]; ++e) hv += val[e] * vecs(col[e], q);
                const double d = hv - ev[q] * vecs(r, q);
                r2 += d * d;
            }
            std::printf("%s%.6g", q ? ", " : "", std::sqrt(r2));
        }
        std::printf("]}\n");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}

int main(int argc, char** argv) {
    if (argc > 1 && std::string(argv[1]) == "--max") return run_max(argc, argv);
    if (argc > 1 && std::string(argv[1]) == "--folm") return run_folm(argc, argv);
    if (argc > 1 && std::string(argv[1]) == "--hmf") return run_hmf(argc, argv);
    if (argc > 1 && std::string(argv[1]) == "--scmf") return run_scmf(argc, argv);
    if (argc < 6) {
        std::fprintf(stderr, "usage: bhx-host-example M N J U MU [K]\n");
        return 1;
    }
    const int m = std::atoi(argv[1]), n = std::atoi(argv[2]);
    const double J = std::atof(argv[3]), U = std::atof(argv[4]), mu = std::atof(argv[5]);
    const int k = argc > 6 ? std::atoi(argv[6]) : 4;
    try {
        Neighbours nb(m);
        nb.chain_neighbours();
        const auto nei = nb.getNeighbours();
        auto [tags, basis] = BH::fixed_set_basis(m, n);
        BH::SparseMatrix JH = BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 1, 0, 0);
        BH::SparseMatrix UH = BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 0, 1, 0);
        BH::SparseMatrix uH = BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 0, 0, 1);
        BH::SparseMatrix H = JH * J + UH * U + uH * mu;
        std::vector<int64_t> rp;
        std::vector<int32_t> col;
        std::vector<double> val;
        H.to_csr(rp, col, val);
        bhx::MatrixXcd vecs;
        const auto ev = Op::IRLM_eigen(H, k, vecs);
        std::printf("{\"D\": %lld, \"nnz\": %lld, \"tags\": [", static_cast<long long>(basis.cols()),
                    static_cast<long long>(rp.back()));
        for (size_t i = 0; i < tags.size() && i < 4; ++i) std::printf("%s%.17g", i ? ", " : "", tags[i]);
        std::printf("], \"state0\": [");
        for (int i = 0; i < m; ++i) std::printf("%s%g", i ? ", " : "", basis(i, 0));
        std::printf("], \"lanczos\": [");
        for (size_t i = 0; i < ev.size(); ++i) std::printf("%s%.17g", i ? ", " : "", ev[i].real());
        const bhx::MatrixXcd rho = Analysis::SPDM(basis, tags, vecs);
        std::printf("], \"condensate_fraction\": %.17g, \"coherence\": %.17g, \"spdm_trace\": %.17g",
                    Analysis::condensate_fraction(rho), Analysis::coherence(rho), [&] {
                        double t = 0.0;
                        for (int i = 0; i < m; ++i) t += rho(i, i).real();
                        return t;
                    }());
        std::printf(", \"vec_rows\": %lld, \"exact\": [", static_cast<long long>(vecs.rows()));
        double exres = 0.0;
        if (H.rows() <= 8000) {
            // the dense solve: its lowest values and the largest residual |H v - e v| over its first vectors
            bhx::MatrixXd ex;
            const auto w = Op::exact_eigen(H, ex);
            for (size_t i = 0; i < w.size() && i < 12; ++i) std::printf("%s%.17g", i ? ", " : "", w[i]);
            for (int64_t q = 0; q < ex.cols() && q < 12; ++q) {
                double r2 = 0.0;
                for (int64_t r = 0; r + 1 < static_cast<int64_t>(rp.size()); ++r) {
                    double hv = 0.0;
                    for (int64_t e = rp[r]; e < rp[r + 1]; ++e) hv += val[e] * ex(col[e], q);
                    const double d = hv - w[q] * ex(r, q);
                    r2 += d * d;
                }
                exres = std::max(exres, std::sqrt(r2));
            }
        }
        std::printf("], \"exact_residual\": %.6g}\n", exres);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}
