// bh_host.hpp -- C++ host surface of the exact-diagonalisation path with the reference's names.
//
// Mirrors the classes and namespaces a caller of the reference uses on this path:
//   class Neighbours               include/neighbours.hpp:12-62   (src/neighbours.cpp:21-71, 126-128)
//   BH::fixed_set_basis            include/hamiltonian.hpp:64     (src/hamiltonian.cpp:143-149)
//   BH::fixed_bosons_hamiltonian   include/hamiltonian.hpp:87     (src/hamiltonian.cpp:238-256)
//   Op::IRLM_eigen, Op::FOLM_eigen, Op::exact_eigen  include/operator.hpp:34, 45, 55
//                                                     (src/operator.cpp:22-33, 39-87, 92-101)
//   Analysis::gap_ratios, Analysis::exact_parameters  include/analysis.hpp (src/analysis.cpp:207-387,433-454)
// over the C ABI of include/bh_capi.h.  Eigen types are replaced by the small dense containers below
// (column-major, as Eigen's defaults).  A sparse matrix is held as a linear combination
// cJ T + cU V + cMu N of one sector's device-resident terms (T hopping, V interaction, N = n I), so that
// the reference's composition H_fixed + H1 * p1 + H2 * p2 (src/analysis.cpp:311) keeps its meaning and
// every solve runs on the device.  Errors are thrown as the reference throws them: std::runtime_error
// from the eigensolvers (src/operator.cpp:28), std::invalid_argument from Neighbours
// (src/neighbours.cpp:43,80).
#pragma once

#include <complex>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/bh_capi.h"

namespace bhx {

// dense column-major matrix
template <typename T>
struct Matrix {
    int64_t nrows = 0, ncols = 0;
    std::vector<T> v;
    Matrix() = default;
    Matrix(int64_t r, int64_t c) : nrows(r), ncols(c), v(static_cast<size_t>(r * c), T()) {}
    int64_t rows() const { return nrows; }
    int64_t cols() const { return ncols; }
    T& operator()(int64_t i, int64_t j) { return v[static_cast<size_t>(i + j * nrows)]; }
    const T& operator()(int64_t i, int64_t j) const { return v[static_cast<size_t>(i + j * nrows)]; }
    T coeff(int64_t i, int64_t j) const { return (*this)(i, j); }
    void resize(int64_t r, int64_t c) {
        nrows = r;
        ncols = c;
        v.assign(static_cast<size_t>(r * c), T());
    }
};
using MatrixXd = Matrix<double>;
using MatrixXcd = Matrix<std::complex<double>>;

// Process-wide runtime of the shims: one device context (device BHX_DEVICE, default 0), the lattice and
// convention used for the sectors the shims build.
struct Settings {
    int device = 0;
    unsigned flags = 0;      // BH_CONV_* (0: the physical conventions)
    int prec = BH_PREC_MIXED;
};
Settings& settings();
bh_ctx* context();
void shutdown();  // releases the context (sectors still held by callers must be released first)
void check(int rc, const char* what);  // throws std::runtime_error with the library's message

// one sector: (neighbour lists, m, n) with its basis in the reference order and its Hamiltonian terms
struct Sector {
    int m = 0, n = 0;
    std::vector<int> off, sites;
    bh_basis* basis = nullptr;
    bh_ham* ham = nullptr;
    ~Sector();
};
std::shared_ptr<Sector> sector(const std::vector<std::vector<int>>& neighbours, int m, int n);

}  // namespace bhx

class Neighbours {
   public:
    explicit Neighbours(int m);
    ~Neighbours();
    // chain of m sites: i -> [i-1 if i > 0, i+1 if i < m-1], then the periodic pair (src/neighbours.cpp:21-34)
    void chain_neighbours(bool closed = true);
    // L x L square lattice (m = L^2), left/right/top/bottom with wrap when closed; the reference writes
    // these lists to a shadowing local (src/neighbours.cpp:45) and returns them empty; here they are filled
    void square_neighbours(bool closed = true);
    std::vector<std::vector<int>> getNeighbours() const;

   private:
    int m;
    std::vector<std::vector<int>> neighbours;
};

namespace BH {

// Sparse operator cJ T + cU V + cMu N on one sector (D x D, reference basis order), or a block-diagonal
// composition of such operators over consecutive boson numbers (max_bosons_hamiltonian: `blocks` set, rows of
// block b after those of blocks 0..b-1)
class SparseMatrix {
   public:
    SparseMatrix() = default;
    SparseMatrix(std::shared_ptr<bhx::Sector> s, double cJ, double cU, double cMu) : sec(std::move(s)), J(cJ), U(cU), Mu(cMu) {}
    explicit SparseMatrix(std::vector<SparseMatrix> b) : blocks(std::move(b)) {}
    int64_t rows() const;
    int64_t cols() const { return rows(); }
    int64_t nonZeros() const;  // stored entries of the materialised matrix (off-diagonal plus diagonal)
    SparseMatrix operator+(const SparseMatrix& o) const;
    SparseMatrix operator*(double f) const;
    // CSR materialisation on the host (row_ptr D+1, col, val), rows in the reference basis order
    void to_csr(std::vector<int64_t>& row_ptr, std::vector<int32_t>& col, std::vector<double>& val) const;
    std::shared_ptr<bhx::Sector> sec;
    double J = 0.0, U = 0.0, Mu = 0.0;
    std::vector<SparseMatrix> blocks;
};
inline SparseMatrix operator*(double f, const SparseMatrix& a) { return a * f; }

// (tags, basis): tags[k] = sum_i n_i ln p_i and column k of the m x D basis, in the reference order
std::pair<std::vector<double>, bhx::MatrixXd> fixed_set_basis(int m, int n);

// Tags and states of every boson number 1 .. n (src/hamiltonian.cpp:151-166): the fixed-N bases one after
// the other, then the reference's sort_basis (src/hamiltonian.cpp:108-123: argsort of the tags, applied by
// its cycle-swap loop) over the concatenation.
std::pair<std::vector<double>, bhx::MatrixXd> max_set_basis(int m, int n);

// Block-diagonal Hamiltonian over the boson numbers n_min .. n_max (src/hamiltonian.cpp:259-287): block N is
// fixed_bosons_hamiltonian(neighbours, ..., m, N, J, U, mu) with its one-term selection, in the fixed-N
// reference order; n_min < 0 is taken as 0 and n_max < n_min as n_min, as the reference does.
SparseMatrix max_bosons_hamiltonian(const std::vector<std::vector<int>>& neighbours, int m, int n_min, int n_max, double J,
                                    double U, double mu);

// The mean-field single-site Hamiltonian (src/hamiltonian.cpp:290-309), U = 1: fills the tridiagonal of the
// leading (2p+1) x (2p+1) block of h, d_i = -mu i + i(i-1)/2 + q J psi^2 and h(i, i+1) = h(i+1, i) =
// -q J psi sqrt(i+1); entries outside it are left as they are.  h is resized (zero-filled) when smaller.
void h_MF(double psi, int p, double mu, double J, int q, bhx::MatrixXd& h);

// The reference's term selection (src/hamiltonian.cpp:242-255): J != 0 gives the hopping term alone,
// else U != 0 the interaction term alone, else mu != 0 the chemical term alone, else a message on stderr
// and the zero matrix.  basis and tags must be the pair fixed_set_basis(m, n) returned.
SparseMatrix fixed_bosons_hamiltonian(const std::vector<std::vector<int>>& neighbours, const bhx::MatrixXd& basis,
                                      const std::vector<double>& tags, int m, int n, double J, double U, double mu);

}  // namespace BH

namespace Op {

// Lowest nb_eigen eigenvalues of O counted with their multiplicity, ascending, and eigenvectors
// (D x nb_eigen), by the device Lanczos solver (deflation passes, bh_lanczos).  As Spectra's GenEigsSolver
// with nev = nb_eigen and ncv = 2 nb_eigen + 1 (src/operator.cpp:24, GenEigsBase.h:335-339) it throws
// std::invalid_argument unless 1 <= nb_eigen <= D - 2 and 2 nb_eigen + 1 <= D; nb_eigen above 64 is refused
// the same way (the reference has no such limit); std::runtime_error("Eigenvalue computation failed.") when
// the solve does not converge (src/operator.cpp:26-29).  A block-diagonal O is solved block by block (a block
// too small for the Lanczos solver by the dense solve) and the lowest nb_eigen values over all blocks are
// returned.
std::vector<std::complex<double>> IRLM_eigen(const BH::SparseMatrix& O, int nb_eigen, bhx::MatrixXcd& eigenvectors);

// nb_iter Lanczos steps from a start vector of the Park-Miller sequence (bh_lanczos_tridiag), then every
// eigenpair of the tridiagonal T: its eigenvalues ascending, and eigenvectors V S (D x steps) for the Lanczos
// basis V and T's eigenvectors S (src/operator.cpp:74-87).  Fewer than nb_iter values when the recurrence
// stops at beta < 1e-6.
std::vector<double> FOLM_eigen(const BH::SparseMatrix& O, int nb_iter, bhx::MatrixXd& eigenvectors);

// All eigenvalues (ascending) and eigenvectors (columns) of O by a dense FP64 solve on the device
// (bh_dense_eigen); std::runtime_error when it fails (src/operator.cpp:92-101).
std::vector<double> exact_eigen(const BH::SparseMatrix& O, bhx::MatrixXd& eigenvectors);

}  // namespace Op

namespace Analysis {

// r_i = min(delta_i, delta_{i+1}) / max(delta_i, delta_{i+1}) over the sorted real parts, 0 when the
// larger gap is 0 (src/analysis.cpp:433-454)
std::vector<double> gap_ratios(const std::vector<std::complex<double>>& eigenvalues, int nb_eigen);

// Single-particle density matrix of the ground state (first column of `eigenvectors`, real part; basis and
// tags as BH::fixed_set_basis returns them), computed on the device: rho_ij = <phi0| b_i^+ b_j |phi0> with the
// hop amplitude sqrt((n_i+1) n_j), divided by the number of columns as the reference does
// (src/analysis.cpp:497-512; its braket, :562-594, takes the amplitude after the hop and finds the target with
// a tolerance search -- not reproduced).  m x m.
bhx::MatrixXcd SPDM(const bhx::MatrixXd& basis, const std::vector<double>& tags, bhx::MatrixXcd& eigenvectors);

// (sum_ij rho_ij rho_ji - sum_i rho_ii^2) / sum_ij rho_ij rho_ji (src/analysis.cpp:541-556)
double coherence(const bhx::MatrixXcd& spdm);

// |largest-magnitude eigenvalue of Re rho| / trace rho, as the reference computes it inline
// (src/analysis.cpp:330-333)
double condensate_fraction(const bhx::MatrixXcd& spdm);

// <phi0| a |phi0> = sum_{i=1}^{2p} sqrt(i) phi0_i phi0_{i-1} over the occupations 0..2p (src/analysis.cpp:95-105)
double SF_density(const std::vector<double>& phi0, int p);

// Self-consistent mean-field order parameter psi at (mu, J), U = 1, q neighbours, from the ansatz psi0, both
// stopping thresholds 10^-precision (src/analysis.cpp:111-178), on the device (bh_mean_field).  Throws
// std::runtime_error when the point does not converge within the caps (100000 self-consistency iterations,
// p <= 32); the reference loops without a cap.
double SCMF(double mu, double J, int q, double psi0, int precision);

// The mean-field phase diagram of the reference CLI (src/analysis.cpp:58-89): mu over [0, 4) in steps 4/n and
// J over [0, 0.25) in steps 0.25/n (accumulated in double as the reference's range loop does), q = 1, one
// batch of SCMF solves on the device, written to `out` (default mean_field.txt) as "mu J psi" per point, mu
// outer.  The reference draws each psi0 uniformly in [0, 1) from a nondeterministic seed; here psi0 of the
// k-th point (file order) is output k of the SplitMix64 stream seeded with `seed`.  Returns the number of
// points that did not converge within the caps (their last iterate is written).
int64_t mean_field_parameters(int n, int precision, const std::string& out = "mean_field.txt", uint64_t seed = 0);

// The exact sweep of the reference CLI (src/analysis.cpp:207-387): periodic chain of m sites, n bosons,
// a grid over two parameters with the third fixed, mean gap ratio per grid point, written to `out`
// (default phase.txt) as "fixed_param fixed_value" then "p1 p2 gap_ratio condensate_fraction coherence"
// per point (condensate fraction and coherence from the ground state's single-particle density matrix).
void exact_parameters(int m, int n, double J, double U, double mu, double s, double r, std::string fixed_param,
                      const std::string& out = "phase.txt");

}  // namespace Analysis
