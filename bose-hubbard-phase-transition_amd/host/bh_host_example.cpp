// bh_host_example.cpp -- the reference's call sequence through the C++ host surface, as a caller writes it:
//   Neighbours -> BH::fixed_set_basis -> BH::fixed_bosons_hamiltonian (one term each) -> H = JH*J + UH*U
//   + uH*mu -> Op::IRLM_eigen (and Op::exact_eigen for small sectors) -> Analysis::SPDM, condensate fraction,
//   coherence.
// Prints one JSON object: dimension, the first tags, the CSR size of H, the Lanczos levels, the ground
// state's density-matrix observables and, when the sector is small, the lowest exact eigenvalues.
//   bhx-host-example M N J U MU [K]
#include <cstdio>
#include <cstdlib>
#include <string>

#include "bh_host.hpp"

int main(int argc, char** argv) {
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
        if (H.rows() <= 1500) {
            bhx::MatrixXd ex;
            const auto w = Op::exact_eigen(H, ex);
            for (size_t i = 0; i < w.size() && i < 8; ++i) std::printf("%s%.17g", i ? ", " : "", w[i]);
        }
        std::printf("]}\n");
    } catch (const std::exception& e) {
        std::fprintf(stderr, "error: %s\n", e.what());
        return 3;
    }
    return 0;
}
