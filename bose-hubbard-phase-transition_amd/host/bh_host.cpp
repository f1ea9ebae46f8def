// bh_host.cpp -- the reference's host classes over the C ABI (see bh_host.hpp).
#include "bh_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <iostream>
#include <limits>
#include <map>
#include <mutex>
#include <numeric>
#include <tuple>

#include "../csrc/bh_tridiag.hpp"

namespace bhx {

Settings& settings() {
    static Settings s = [] {
        Settings t;
        if (const char* e = std::getenv("BHX_DEVICE")) t.device = std::atoi(e);
        return t;
    }();
    return s;
}

namespace {
// The context is not destroyed by a static destructor: at process exit the HIP runtime may already be torn
// down (its own exit handlers are registered after this object is built), and the process's device memory is
// released with the process anyway.  bhx::shutdown() releases it explicitly.
struct CtxHolder {
    bh_ctx* ctx = nullptr;
};
CtxHolder& holder() {
    static CtxHolder h;
    return h;
}
std::mutex& mu() {
    static std::mutex m;
    return m;
}
}  // namespace

bh_ctx* context() {
    std::lock_guard<std::mutex> lk(mu());
    CtxHolder& h = holder();
    if (!h.ctx) {
        const int rc = bh_ctx_create(settings().device, 0, &h.ctx);
        if (rc != BH_OK) throw std::runtime_error("bhx: cannot create the device context (status " + std::to_string(rc) + ")");
    }
    return h.ctx;
}

void shutdown() {
    std::lock_guard<std::mutex> lk(mu());
    CtxHolder& h = holder();
    if (h.ctx) bh_ctx_destroy(h.ctx);
    h.ctx = nullptr;
}

void check(int rc, const char* what) {
    if (rc == BH_OK) return;
    const char* msg = holder().ctx ? bh_last_error(holder().ctx) : "";
    throw std::runtime_error(std::string(what) + ": " + (msg ? msg : "") + " (status " + std::to_string(rc) + ")");
}

Sector::~Sector() {
    if (ham) bh_ham_destroy(ham);
    if (basis) bh_basis_destroy(basis);
}

std::shared_ptr<Sector> sector(const std::vector<std::vector<int>>& neighbours, int m, int n) {
    using Key = std::tuple<std::vector<int>, std::vector<int>, int, int, unsigned>;
    static std::map<Key, std::weak_ptr<Sector>> cache;
    static std::mutex cm;
    if (static_cast<int>(neighbours.size()) != m) throw std::invalid_argument("neighbour lists must have one entry per site");
    std::vector<int> off(1, 0), sites;
    for (const auto& l : neighbours) {
        for (int s : l) sites.push_back(s);
        off.push_back(static_cast<int>(sites.size()));
    }
    const Key key{off, sites, m, n, settings().flags};
    std::lock_guard<std::mutex> lk(cm);
    auto it = cache.find(key);
    if (it != cache.end())
        if (auto sp = it->second.lock()) return sp;
    auto s = std::make_shared<Sector>();
    s->m = m;
    s->n = n;
    s->off = off;
    s->sites = sites;
    bh_ctx* ctx = context();
    check(bh_basis_build(ctx, m, n, BH_ORDER_REFERENCE, &s->basis), "basis");
    check(bh_hamiltonian_build(ctx, s->basis, s->off.data(), s->sites.empty() ? nullptr : s->sites.data(),
                               settings().flags, &s->ham),
          "hamiltonian");
    cache[key] = s;
    return s;
}

}  // namespace bhx

// ---- Neighbours --------------------------------------------------------------------------------------

Neighbours::Neighbours(int m_) : m(m_) {}
Neighbours::~Neighbours() = default;

static std::vector<std::vector<int>> lists_of(int kind, int Lx, int Ly, bool closed) {
    const int m = Lx * Ly;
    std::vector<int> off(static_cast<size_t>(m) + 1), sites(static_cast<size_t>(4 * m + 4));
    const int ne = bh_lattice_neighbours(kind, Lx, Ly, closed ? 1 : 0, off.data(), sites.data());
    if (ne < 0) throw std::invalid_argument("invalid lattice");
    std::vector<std::vector<int>> out(static_cast<size_t>(m));
    for (int i = 0; i < m; ++i) out[i].assign(sites.begin() + off[i], sites.begin() + off[i + 1]);
    return out;
}

void Neighbours::chain_neighbours(bool closed) {
    neighbours = m >= 1 ? lists_of(0, m, 1, closed) : std::vector<std::vector<int>>();
}

void Neighbours::square_neighbours(bool closed) {
    const int L = static_cast<int>(std::lround(std::sqrt(static_cast<double>(m))));
    if (L * L != m) throw std::invalid_argument("The number of sites must be a perfect square.");
    neighbours = lists_of(1, L, L, closed);
}

std::vector<std::vector<int>> Neighbours::getNeighbours() const { return neighbours; }

// ---- BH ------------------------------------------------------------------------------------------------

namespace BH {

int64_t SparseMatrix::rows() const {
    if (!blocks.empty()) {
        int64_t d = 0;
        for (const auto& b : blocks) d += b.rows();
        return d;
    }
    return sec ? bh_ham_dim(sec->ham) : 0;
}

int64_t SparseMatrix::nonZeros() const {
    if (!blocks.empty()) {
        int64_t z = 0;
        for (const auto& b : blocks) z += b.nonZeros();
        return z;
    }
    if (!sec) return 0;
    int64_t z = 0;
    if (J != 0.0) z += bh_ham_nnz(sec->ham);
    if (U != 0.0 || Mu != 0.0) z += rows();
    return z;
}

SparseMatrix SparseMatrix::operator*(double f) const {
    if (blocks.empty()) return SparseMatrix(sec, J * f, U * f, Mu * f);
    std::vector<SparseMatrix> b;
    for (const auto& x : blocks) b.push_back(x * f);
    return SparseMatrix(std::move(b));
}

SparseMatrix SparseMatrix::operator+(const SparseMatrix& o) const {
    if (!blocks.empty() || !o.blocks.empty()) {
        if (blocks.empty() && !sec) return o;
        if (o.blocks.empty() && !o.sec) return *this;
        if (blocks.size() != o.blocks.size())
            throw std::invalid_argument("block-diagonal matrices of different block structure cannot be added");
        std::vector<SparseMatrix> b;
        for (size_t i = 0; i < blocks.size(); ++i) b.push_back(blocks[i] + o.blocks[i]);
        return SparseMatrix(std::move(b));
    }
    if (!sec) return o;
    if (!o.sec) return *this;
    if (sec != o.sec) throw std::invalid_argument("sparse matrices of different sectors cannot be added");
    return SparseMatrix(sec, J + o.J, U + o.U, Mu + o.Mu);
}

void SparseMatrix::to_csr(std::vector<int64_t>& row_ptr, std::vector<int32_t>& col, std::vector<double>& val) const {
    if (!blocks.empty()) {
        row_ptr.assign(1, 0);
        col.clear();
        val.clear();
        int64_t off = 0;
        for (const auto& b : blocks) {
            std::vector<int64_t> rp;
            std::vector<int32_t> c;
            std::vector<double> v;
            b.to_csr(rp, c, v);
            const int64_t base = static_cast<int64_t>(col.size());
            for (size_t q = 0; q < c.size(); ++q) {
                col.push_back(static_cast<int32_t>(c[q] + off));
                val.push_back(v[q]);
            }
            for (size_t r = 1; r < rp.size(); ++r) row_ptr.push_back(rp[r] + base);
            off += b.rows();
        }
        return;
    }
    const int64_t D = rows();
    const int64_t nt = sec ? bh_ham_nnz(sec->ham) : 0;
    std::vector<int64_t> rp(static_cast<size_t>(D) + 1, 0);
    std::vector<int32_t> c(static_cast<size_t>(std::max<int64_t>(nt, 1)));
    std::vector<double> v(c.size()), d(static_cast<size_t>(D));
    if (sec) bhx::check(bh_ham_copy_csr(sec->ham, rp.data(), c.data(), v.data(), d.data()), "copy_csr");
    row_ptr.assign(static_cast<size_t>(D) + 1, 0);
    col.clear();
    val.clear();
    const double nshift = sec ? Mu * sec->n : 0.0;
    for (int64_t r = 0; r < D; ++r) {
        const double dv = U * d[r] + nshift;
        bool placed = U == 0.0 && Mu == 0.0;  // the diagonal is stored whenever a diagonal term is present
        for (int64_t q = (J != 0.0 ? rp[r] : rp[r + 1]); q < rp[r + 1]; ++q) {
            if (!placed && c[q] > r) {
                col.push_back(static_cast<int32_t>(r));
                val.push_back(dv);
                placed = true;
            }
            col.push_back(c[q]);
            val.push_back(J * v[q]);
        }
        if (!placed) {
            col.push_back(static_cast<int32_t>(r));
            val.push_back(dv);
        }
        row_ptr[r + 1] = static_cast<int64_t>(col.size());
    }
}

std::pair<std::vector<double>, bhx::MatrixXd> fixed_set_basis(int m, int n) {
    bh_basis* b = nullptr;
    bhx::check(bh_basis_build(bhx::context(), m, n, BH_ORDER_REFERENCE, &b), "fixed_set_basis");
    const int64_t D = bh_basis_dim(b);
    std::vector<uint8_t> st(static_cast<size_t>(D * m));
    std::vector<double> tags(static_cast<size_t>(D));
    int rc = bh_basis_copy_states(b, st.data());
    if (rc == BH_OK) rc = bh_basis_copy_tags(b, tags.data());
    bh_basis_destroy(b);
    bhx::check(rc, "fixed_set_basis");
    bhx::MatrixXd basis(m, D);
    for (int64_t k = 0; k < D; ++k)
        for (int i = 0; i < m; ++i) basis(i, k) = st[static_cast<size_t>(k * m + i)];
    return {tags, basis};
}

std::pair<std::vector<double>, bhx::MatrixXd> max_set_basis(int m, int n) {
    std::vector<double> tags;
    std::vector<std::vector<double>> cols;
    for (int N = 1; N <= n; ++N) {
        auto [t, b] = fixed_set_basis(m, N);
        for (int64_t k = 0; k < b.cols(); ++k) {
            tags.push_back(t[static_cast<size_t>(k)]);
            std::vector<double> c(m);
            for (int i = 0; i < m; ++i) c[i] = b(i, k);
            cols.push_back(std::move(c));
        }
    }
    // the reference's sort: argsort by tag, then its cycle-swap loop (which realises the inverse permutation)
    std::vector<int> idx(tags.size());
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&tags](int a, int b) { return tags[a] < tags[b]; });
    for (int i = 0; i < static_cast<int>(idx.size()); ++i) {
        while (idx[i] != i) {
            const int j = idx[i];
            std::swap(tags[i], tags[j]);
            std::swap(cols[i], cols[j]);
            std::swap(idx[i], idx[j]);
        }
    }
    bhx::MatrixXd basis(m, static_cast<int64_t>(cols.size()));
    for (size_t k = 0; k < cols.size(); ++k)
        for (int i = 0; i < m; ++i) basis(i, static_cast<int64_t>(k)) = cols[k][i];
    return {tags, basis};
}

SparseMatrix max_bosons_hamiltonian(const std::vector<std::vector<int>>& neighbours, int m, int n_min, int n_max, double J,
                                    double U, double mu) {
    if (n_min < 0) n_min = 0;
    if (n_max < n_min) n_max = n_min;
    std::vector<SparseMatrix> b;
    for (int N = n_min; N <= n_max; ++N) {
        auto [t, basis] = fixed_set_basis(m, N);
        b.push_back(fixed_bosons_hamiltonian(neighbours, basis, t, m, N, J, U, mu));
    }
    return SparseMatrix(std::move(b));
}

void h_MF(double psi, int p, double mu, double J, int q, bhx::MatrixXd& h) {
    if (p < 1) throw std::invalid_argument("h_MF: need p >= 1");
    const int64_t N = 2 * static_cast<int64_t>(p) + 1;
    if (h.rows() < N || h.cols() < N) h.resize(N, N);
    for (int64_t i = 0; i < N; ++i) {
        const double x = static_cast<double>(i);
        h(i, i) = -mu * x + 0.5 * x * static_cast<double>(i - 1) + q * J * psi * psi;
    }
    for (int64_t i = 0; i + 1 < N; ++i) {
        const double e = -q * J * psi * std::sqrt(static_cast<double>(i + 1));
        h(i + 1, i) = e;
        h(i, i + 1) = e;
    }
}

SparseMatrix fixed_bosons_hamiltonian(const std::vector<std::vector<int>>& neighbours, const bhx::MatrixXd& basis,
                                      const std::vector<double>& tags, int m, int n, double J, double U, double mu) {
    auto s = bhx::sector(neighbours, m, n);
    const int64_t D = bh_ham_dim(s->ham);
    if (basis.rows() != m || basis.cols() != D || static_cast<int64_t>(tags.size()) != D)
        throw std::invalid_argument("basis and tags do not belong to this (m, n) sector");
    const double eps = std::numeric_limits<double>::epsilon();
    if (std::abs(J) > eps) return SparseMatrix(s, J, 0.0, 0.0);
    if (std::abs(U) > eps) return SparseMatrix(s, 0.0, U, 0.0);
    if (std::abs(mu) > eps) return SparseMatrix(s, 0.0, 0.0, -mu);  // -mu N
    std::cerr << "Error: At least one of the parameters J, U, mu must be different from zero." << std::endl;
    return SparseMatrix(s, 0.0, 0.0, 0.0);
}

}  // namespace BH

// ---- Op --------------------------------------------------------------------------------------------------

namespace Op {

namespace {
constexpr int KMAX = 64;  // eigenvalues per Lanczos solve (bh_lanczos)
}  // namespace

std::vector<std::complex<double>> IRLM_eigen(const BH::SparseMatrix& O, int nb_eigen, bhx::MatrixXcd& eigenvectors) {
    if (!O.blocks.empty()) {
        // block by block: the lowest nb_eigen eigenvalues of each block (all of a small block), merged
        if (nb_eigen < 1) throw std::invalid_argument("nev must satisfy 1 <= nev <= n - 2, n is the size of matrix");
        struct Lv {
            double e;
            size_t blk;
            int64_t col;
        };
        std::vector<Lv> all;
        std::vector<bhx::MatrixXcd> bv(O.blocks.size());
        std::vector<int64_t> off(O.blocks.size() + 1, 0);
        for (size_t q = 0; q < O.blocks.size(); ++q) {
            const auto& blk = O.blocks[q];
            off[q + 1] = off[q] + blk.rows();
            const int kb = static_cast<int>(std::min<int64_t>(nb_eigen, blk.rows()));
            if (kb < 1) continue;
            std::vector<std::complex<double>> e;
            if (2 * static_cast<int64_t>(kb) + 1 > blk.rows() || kb > KMAX) {
                // a block the Lanczos solver cannot take (down to the one-state vacuum): dense solve
                bhx::MatrixXd V;
                const auto w = exact_eigen(blk, V);
                bv[q].resize(blk.rows(), kb);
                for (int i = 0; i < kb; ++i) {
                    e.emplace_back(w[i], 0.0);
                    for (int64_t r = 0; r < blk.rows(); ++r) bv[q](r, i) = V(r, i);
                }
            } else {
                e = IRLM_eigen(blk, kb, bv[q]);
            }
            for (size_t i = 0; i < e.size(); ++i) all.push_back({e[i].real(), q, static_cast<int64_t>(i)});
        }
        std::stable_sort(all.begin(), all.end(), [](const Lv& a, const Lv& b) { return a.e < b.e; });
        const int k = static_cast<int>(std::min<size_t>(static_cast<size_t>(nb_eigen), all.size()));
        if (k < 1) throw std::runtime_error("Eigenvalue computation failed.");
        eigenvectors.resize(O.rows(), k);
        std::vector<std::complex<double>> out;
        for (int i = 0; i < k; ++i) {
            out.emplace_back(all[i].e, 0.0);
            const auto& V = bv[all[i].blk];
            for (int64_t r = 0; r < V.rows(); ++r) eigenvectors(off[all[i].blk] + r, i) = V(r, all[i].col);
        }
        return out;
    }
    if (!O.sec) throw std::invalid_argument("empty operator");
    const int64_t D = O.rows();
    // Spectra's argument checks for nev = nb_eigen, ncv = 2 nb_eigen + 1 (GenEigsBase.h:335-339)
    if (nb_eigen < 1 || nb_eigen > D - 2)
        throw std::invalid_argument("nev must satisfy 1 <= nev <= n - 2, n is the size of matrix");
    if (2 * static_cast<int64_t>(nb_eigen) + 1 > D)
        throw std::invalid_argument("ncv must satisfy nev + 2 <= ncv <= n, n is the size of matrix");
    if (nb_eigen > KMAX) throw std::invalid_argument("IRLM_eigen: nb_eigen above 64 is not supported");
    const int k = nb_eigen;
    bh_lanczos_opts o;
    bh_lanczos_default_opts(&o);
    o.k = k;
    o.prec = bhx::settings().prec;
    o.want_vectors = 1;
    std::vector<double> ev(k), rs(k), ee(k);
    int it = 0, uf = 0, st = 0;
    bh_lanczos_result r{ev.data(), rs.data(), ee.data(), &it, &uf, &st};
    bh_ctx* ctx = bhx::context();
    const int rc = bh_lanczos(ctx, O.sec->ham, 1, &O.J, &O.U, &o, &r);
    if (rc != BH_OK || st != BH_OK) throw std::runtime_error("Eigenvalue computation failed.");
    std::vector<double> vec(static_cast<size_t>(k * D));
    bhx::check(bh_lanczos_copy_vectors(ctx, 1, k, vec.data()), "IRLM_eigen vectors");
    std::vector<std::complex<double>> out;
    eigenvectors.resize(D, k);
    for (int i = 0; i < k; ++i) {
        out.emplace_back(ev[i] + O.Mu * O.sec->n, 0.0);
        for (int64_t q = 0; q < D; ++q) eigenvectors(q, i) = vec[static_cast<size_t>(i * D + q)];
    }
    return out;
}

std::vector<double> FOLM_eigen(const BH::SparseMatrix& O, int nb_iter, bhx::MatrixXd& eigenvectors) {
    if (!O.blocks.empty() || !O.sec) throw std::invalid_argument("FOLM_eigen: one sector operator expected");
    if (nb_iter < 1) throw std::invalid_argument("FOLM_eigen: nb_iter must be positive");
    const int64_t D = O.rows();
    const int n = static_cast<int>(std::min<int64_t>(nb_iter, D));
    std::vector<double> a(n), b(n), basis(static_cast<size_t>(n) * D);
    int steps = 0;
    bhx::check(bh_lanczos_tridiag(bhx::context(), O.sec->ham, O.J, O.U, nb_iter, 0, a.data(), b.data(), &steps, basis.data()),
               "FOLM_eigen");
    // every eigenpair of T (steps x steps) by Jacobi rotations; T's vectors carried to H's space by V
    std::vector<double> T(static_cast<size_t>(steps) * steps, 0.0), w(steps), S(static_cast<size_t>(steps) * steps);
    for (int t = 0; t < steps; ++t) {
        T[static_cast<size_t>(t) * steps + t] = a[t];
        if (t > 0) T[static_cast<size_t>(t) * steps + t - 1] = T[static_cast<size_t>(t - 1) * steps + t] = b[t];
    }
    bhtri::jacobi_eig(steps, T.data(), w.data(), S.data());
    eigenvectors.resize(D, steps);
    for (int c = 0; c < steps; ++c)
        for (int t = 0; t < steps; ++t) {
            const double f = S[static_cast<size_t>(t) * steps + c];
            const double* v = basis.data() + static_cast<size_t>(t) * D;
            for (int64_t r = 0; r < D; ++r) eigenvectors(r, c) += f * v[r];
        }
    for (auto& x : w) x += O.Mu * O.sec->n;
    return w;
}

std::vector<double> exact_eigen(const BH::SparseMatrix& O, bhx::MatrixXd& eigenvectors) {
    const int64_t D = O.rows();
    std::vector<int64_t> rp;
    std::vector<int32_t> c;
    std::vector<double> v;
    O.to_csr(rp, c, v);
    std::vector<double> w(static_cast<size_t>(D));
    eigenvectors.resize(D, D);
    const int rc = bh_dense_eigen(bhx::context(), D, rp.data(), c.data(), v.data(), w.data(), eigenvectors.v.data());
    if (rc != BH_OK) throw std::runtime_error("Eigenvalue computation failed.");
    return w;
}

}  // namespace Op

// ---- Analysis ----------------------------------------------------------------------------------------------

namespace Analysis {

double SF_density(const std::vector<double>& phi0, int p) {
    if (static_cast<int64_t>(phi0.size()) < 2 * static_cast<int64_t>(p) + 1)
        throw std::invalid_argument("SF_density: phi0 shorter than 2p+1");
    double a = 0.0;
    for (int i = 1; i < 2 * p + 1; ++i) a += std::sqrt(static_cast<double>(i)) * phi0[i] * phi0[i - 1];
    return a;
}

namespace {
constexpr int MF_MAX_OUTER = 100000;
constexpr int MF_PMAX = 32;
}  // namespace

double SCMF(double mu, double J, int q, double psi0, int precision) {
    const double eps = std::pow(10.0, -precision);
    double psi = 0.0;
    int32_t st = 0;
    bhx::check(bh_mean_field(bhx::context(), 1, &mu, &J, &psi0, q, eps, MF_MAX_OUTER, MF_PMAX, &psi, &st, nullptr,
                             nullptr),
               "SCMF");
    if (st != 0) throw std::runtime_error("SCMF: no convergence within the iteration caps");
    return psi;
}

int64_t mean_field_parameters(int n, int precision, const std::string& out, uint64_t seed) {
    if (n < 1) throw std::invalid_argument("mean_field_parameters: need n >= 1");
    const double J_min = 0.0, J_max = 0.25, mu_min = 0.0, mu_max = 4.0;
    const double dJ = (J_max - J_min) / n, dmu = (mu_max - mu_min) / n;
    std::vector<double> mus, Js;
    for (double x = mu_min; x < mu_max; x += dmu) mus.push_back(x);
    for (double x = J_min; x < J_max; x += dJ) Js.push_back(x);
    const size_t npts = mus.size() * Js.size();
    std::vector<double> mu(npts), J(npts), psi0(npts), psi(npts);
    std::vector<int32_t> st(npts);
    for (size_t a = 0, k = 0; a < mus.size(); ++a)
        for (size_t b = 0; b < Js.size(); ++b, ++k) {
            mu[k] = mus[a];
            J[k] = Js[b];
            // SplitMix64: output k of the stream seeded with `seed`, top 53 bits
            uint64_t z = seed + (static_cast<uint64_t>(k) + 1) * 0x9E3779B97F4A7C15ull;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            z ^= z >> 31;
            psi0[k] = static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
        }
    const double eps = std::pow(10.0, -precision);
    bhx::check(bh_mean_field(bhx::context(), static_cast<int64_t>(npts), mu.data(), J.data(), psi0.data(), 1, eps,
                             MF_MAX_OUTER, MF_PMAX, psi.data(), st.data(), nullptr, nullptr),
               "mean_field_parameters");
    std::ofstream f(out);
    if (!f) throw std::runtime_error("cannot write " + out);
    int64_t bad = 0;
    for (size_t k = 0; k < npts; ++k) {
        f << mu[k] << " " << J[k] << " " << psi[k] << "\n";
        bad += st[k] != 0;
    }
    return bad;
}

std::vector<double> gap_ratios(const std::vector<std::complex<double>>& eigenvalues, int nb_eigen) {
    const int ne = std::min<int>(nb_eigen, static_cast<int>(eigenvalues.size()));
    std::vector<double> e(ne);
    for (int i = 0; i < ne; ++i) e[i] = eigenvalues[i].real();
    std::sort(e.begin(), e.end());
    std::vector<double> r;
    for (int i = 1; i + 1 < ne; ++i) {
        const double a = e[i] - e[i - 1], b = e[i + 1] - e[i];
        const double mn = std::min(a, b), mx = std::max(a, b);
        r.push_back(mx != 0 ? mn / mx : 0.0);
    }
    return r;
}

namespace {
// rho = <phi| b_i^+ b_j |phi> on the device (m x m, row-major)
std::vector<double> spdm_device(const bh_basis* b, int m, const double* phi) {
    std::vector<double> rho(static_cast<size_t>(m) * m);
    bhx::check(bh_spdm(bhx::context(), b, phi, rho.data()), "spdm");
    return rho;
}
double coherence_of(const std::vector<double>& rho, int m) {
    double all = 0.0, diag = 0.0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            const double v = rho[static_cast<size_t>(i) * m + j] * rho[static_cast<size_t>(j) * m + i];
            all += v;
            if (i == j) diag += v;
        }
    return (all - diag) / all;
}
double condensate_fraction_of(const std::vector<double>& rho, int m) {
    std::vector<double> A(static_cast<size_t>(m) * m), w(m), V(static_cast<size_t>(m) * m);
    double tr = 0.0;
    for (int i = 0; i < m; ++i) {
        tr += rho[static_cast<size_t>(i) * m + i];
        for (int j = 0; j < m; ++j)
            A[static_cast<size_t>(i) * m + j] = 0.5 * (rho[static_cast<size_t>(i) * m + j] + rho[static_cast<size_t>(j) * m + i]);
    }
    bhtri::jacobi_eig(m, A.data(), w.data(), V.data());
    double big = 0.0;
    for (double x : w)
        if (std::abs(x) > std::abs(big)) big = x;
    return std::abs(big / tr);
}
}  // namespace

bhx::MatrixXcd SPDM(const bhx::MatrixXd& basis, const std::vector<double>& tags, bhx::MatrixXcd& eigenvectors) {
    (void)tags;  // positions are exact here; the reference needs the tags for its search
    const int m = static_cast<int>(basis.rows());
    const int64_t D = basis.cols();
    if (m < 1 || D < 1 || eigenvectors.rows() != D || eigenvectors.cols() < 1)
        throw std::invalid_argument("SPDM: basis and eigenvectors do not match");
    int n = 0;
    for (int i = 0; i < m; ++i) n += static_cast<int>(std::lround(basis(i, 0)));
    bh_basis* b = nullptr;
    bhx::check(bh_basis_build(bhx::context(), m, n, BH_ORDER_REFERENCE, &b), "basis");
    std::vector<double> phi(static_cast<size_t>(D));
    for (int64_t q = 0; q < D; ++q) phi[static_cast<size_t>(q)] = eigenvectors(q, 0).real();
    std::vector<double> rho;
    try {
        rho = spdm_device(b, m, phi.data());
    } catch (...) {
        bh_basis_destroy(b);
        throw;
    }
    bh_basis_destroy(b);
    bhx::MatrixXcd out(m, m);
    const double scale = 1.0 / static_cast<double>(eigenvectors.cols());
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) out(i, j) = rho[static_cast<size_t>(i) * m + j] * scale;
    return out;
}

double coherence(const bhx::MatrixXcd& spdm) {
    const int m = static_cast<int>(spdm.rows());
    double all = 0.0, diag = 0.0;
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) {
            const double v = std::real(spdm(i, j) * spdm(j, i));
            all += v;
            if (i == j) diag += v;
        }
    return (all - diag) / all;
}

double condensate_fraction(const bhx::MatrixXcd& spdm) {
    const int m = static_cast<int>(spdm.rows());
    std::vector<double> rho(static_cast<size_t>(m) * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) rho[static_cast<size_t>(i) * m + j] = spdm(i, j).real();
    return condensate_fraction_of(rho, m);
}

void exact_parameters(int m, int n, double J, double U, double mu, double s, double r, std::string fixed_param,
                      const std::string& out) {
    constexpr int KMAX_SWEEP = 64;  // eigenvalues per Lanczos solve (bh_lanczos)
    const double eps = std::numeric_limits<double>::epsilon();
    if (std::abs(J) < eps && std::abs(U) < eps && std::abs(mu) < eps) {
        std::cerr << "Error: At least one of the parameters J, U, mu must be different from zero." << std::endl;
        return;
    }
    Neighbours nb(m);
    nb.chain_neighbours();
    const auto nei = nb.getNeighbours();
    auto [tags, basis] = BH::fixed_set_basis(m, n);
    BH::SparseMatrix JH = BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 1, 0, 0);
    BH::SparseMatrix UH = BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 0, 1, 0);
    BH::SparseMatrix uH = BH::fixed_bosons_hamiltonian(nei, basis, tags, m, n, 0, 0, 1);
    // parameter ranges and their wiring as src/analysis.cpp:241-256 has them: the first grid axis always
    // runs over [J, J + r]; the second over [mu, mu + r] (fixed J or fixed mu) or [U, U + r] (fixed U)
    const double J_min = J, J_max = J + r, mu_min = mu, mu_max = mu + r, U_min = U, U_max = U + r;
    BH::SparseMatrix Hf, H1, H2;
    double fixed_value, p1_min, p1_max, p2_min, p2_max;
    if (fixed_param == "J") {
        Hf = JH * J, H1 = UH, H2 = uH, fixed_value = J;
        p1_min = J_min, p1_max = J_max, p2_min = mu_min, p2_max = mu_max;
    } else if (fixed_param == "U") {
        Hf = UH * U, H1 = JH, H2 = uH, fixed_value = U;
        p1_min = J_min, p1_max = J_max, p2_min = U_min, p2_max = U_max;
    } else {
        Hf = uH * mu, H1 = JH, H2 = UH, fixed_value = mu;
        p1_min = J_min, p1_max = J_max, p2_min = mu_min, p2_max = mu_max;
    }
    const int n1 = static_cast<int>((p1_max - p1_min) / s) + 1, n2 = static_cast<int>((p2_max - p2_min) / s) + 1;
    const int npt = n1 * n2;
    std::vector<double> p1v(npt), p2v(npt), gr(npt);
    std::vector<BH::SparseMatrix> H(npt);
    // Point counts are truncated as the reference truncates them (src/analysis.cpp:281-282), so rounding can
    // give the two axes different counts.  Rows are i-major; the reference's index i * n1 + j
    // (src/analysis.cpp:349) leaves its arrays whenever n2 > n1, and is not reproduced.
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
            const int idx = i * n2 + j;
            p1v[idx] = p1_min + i * s;
            p2v[idx] = p2_min + j * s;
            H[idx] = Hf + H1 * p1v[idx] + H2 * p2v[idx];
        }
    // distinct (J, U) pairs are solved once, batched 64 per call; N = n I only shifts the levels
    std::map<std::pair<double, double>, std::vector<double>> lev;
    for (const auto& h : H) lev[{h.J, h.U}];
    int nb_eigen = 20;
    auto sec = JH.sec;
    const int64_t Dsec = sec->basis ? bh_basis_dim(sec->basis) : 0;
    for (;;) {
        const int k = nb_eigen;
        std::vector<std::pair<double, double>> keys;
        for (auto& kv : lev) keys.push_back(kv.first);
        for (size_t q0 = 0; q0 < keys.size(); q0 += 64) {
            const int B = static_cast<int>(std::min<size_t>(64, keys.size() - q0));
            std::vector<double> Jb(B), Ub(B), ev(static_cast<size_t>(B) * k);
            std::vector<int> st(B);
            for (int b = 0; b < B; ++b) {
                Jb[b] = keys[q0 + b].first;
                Ub[b] = keys[q0 + b].second;
            }
            bh_lanczos_opts o;
            bh_lanczos_default_opts(&o);
            o.k = k;
            o.prec = bhx::settings().prec;
            bh_lanczos_result res{ev.data(), nullptr, nullptr, nullptr, nullptr, st.data()};
            const int rc = bh_lanczos(bhx::context(), sec->ham, B, Jb.data(), Ub.data(), &o, &res);
            if (rc != BH_OK && rc != BH_ENOCONV) throw std::runtime_error("Eigenvalue computation failed.");
            for (int b = 0; b < B; ++b) lev[keys[q0 + b]].assign(ev.begin() + static_cast<int64_t>(b) * k, ev.begin() + static_cast<int64_t>(b + 1) * k);
        }
        for (int idx = 0; idx < npt; ++idx) {
            const auto& L = lev[{H[idx].J, H[idx].U}];
            std::vector<std::complex<double>> e;
            for (double x : L)
                if (std::isfinite(x)) e.emplace_back(x + H[idx].Mu * n, 0.0);
            const auto rr = gap_ratios(e, static_cast<int>(e.size()));
            gr[idx] = rr.empty() ? 0.0 : std::accumulate(rr.begin(), rr.end(), 0.0) / rr.size();
        }
        // the reference recomputes with 5 more eigenvalues while the gap ratios barely vary
        // (src/analysis.cpp:364-381); its solver throws once 2 nb_eigen + 1 exceeds D (GenEigsBase.h:335-339),
        // where this loop stops with the last grid it solved (also at the solver's 64 eigenvalues)
        const double mean = std::accumulate(gr.begin(), gr.end(), 0.0) / npt;
        double var = 0.0;
        for (double g : gr) var += (g - mean) * (g - mean);
        var /= npt;
        if (var > 1e-8 * mean) break;
        if (nb_eigen + 5 > KMAX_SWEEP || 2 * static_cast<int64_t>(nb_eigen + 5) + 1 > Dsec) break;
        nb_eigen += 5;
    }
    // condensate fraction and coherence from the ground state of each distinct (J, U) pair (-mu N leaves the
    // eigenvectors unchanged); ground states in batches of 64, one vector each
    std::map<std::pair<double, double>, std::pair<double, double>> obs;
    {
        std::vector<std::pair<double, double>> keys;
        for (auto& kv : lev) keys.push_back(kv.first);
        const int64_t D = sec->basis ? bh_basis_dim(sec->basis) : 0;
        for (size_t q0 = 0; q0 < keys.size(); q0 += 64) {
            const int B = static_cast<int>(std::min<size_t>(64, keys.size() - q0));
            std::vector<double> Jb(B), Ub(B), ev(B);
            std::vector<int> st(B);
            for (int b = 0; b < B; ++b) {
                Jb[b] = keys[q0 + b].first;
                Ub[b] = keys[q0 + b].second;
            }
            bh_lanczos_opts o;
            bh_lanczos_default_opts(&o);
            o.k = 1;
            o.prec = bhx::settings().prec;
            o.want_vectors = 1;
            o.distinct = 1;  // one ground state vector: the pass that finds E0 is enough
            bh_lanczos_result res{ev.data(), nullptr, nullptr, nullptr, nullptr, st.data()};
            const int rc = bh_lanczos(bhx::context(), sec->ham, B, Jb.data(), Ub.data(), &o, &res);
            if (rc != BH_OK && rc != BH_ENOCONV) throw std::runtime_error("Eigenvalue computation failed.");
            std::vector<double> vec(static_cast<size_t>(B) * D);
            bhx::check(bh_lanczos_copy_vectors(bhx::context(), B, 1, vec.data()), "ground states");
            for (int b = 0; b < B; ++b) {
                const auto rho = spdm_device(sec->basis, m, vec.data() + static_cast<size_t>(b) * D);
                obs[keys[q0 + b]] = {condensate_fraction_of(rho, m), coherence_of(rho, m)};
            }
        }
    }
    std::ofstream f(out);
    f << fixed_param << " " << fixed_value << std::endl;
    for (int idx = 0; idx < npt; ++idx) {
        const auto& ob = obs[{H[idx].J, H[idx].U}];
        f << p1v[idx] << " " << p2v[idx] << " " << gr[idx] << " " << ob.first << " " << ob.second << std::endl;
    }
}

}  // namespace Analysis
