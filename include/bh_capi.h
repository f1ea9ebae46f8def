/*
 * bh_capi.h -- C ABI of libbhx.so, the MI355X (gfx950) exact-diagonalisation engine for the
 * fixed-N Bose-Hubbard Hamiltonian  H = J T + U V  (T: hopping over a neighbour list, V: on-site
 * interaction; the chemical term -mu N is a constant in a fixed-N sector and is added by callers).
 *
 * The reference project has no FFI; its hot path sits behind C++ functions with Eigen types.
 * Each entry point below names the reference interface it replaces (paths relative to the
 * reference repository).  The C++ host layer (BH::, Op::, Analysis::, Neighbours, the CLI) is built
 * on these calls.
 *
 * Conventions of the ABI:
 *   - every function returns an int status (BH_OK == 0); no exception crosses the boundary;
 *     bh_last_error(ctx) gives a message for the last failure on that context;
 *   - handles are opaque and owned by the library; host arrays are owned by the caller;
 *     device memory is owned by the library and released by the matching _destroy;
 *   - a context serialises its calls with an internal mutex; use one context per host thread
 *     for concurrency;
 *   - the basis order is the reference's (see BH_ORDER_REFERENCE).
 */
#ifndef BH_CAPI_H
#define BH_CAPI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------------------------------- */
#define BH_OK 0
#define BH_EINVAL 1   /* invalid argument (sizes, neighbour list, options) */
#define BH_ENOMEM 2   /* device or host allocation failed */
#define BH_EBUDGET 3  /* request exceeds the context's HBM budget */
#define BH_ENOCONV 4  /* eigensolver did not converge within max_iter */
#define BH_EHIP 5     /* HIP runtime error */
#define BH_ERANGE 6   /* sector too large for an exact 128-bit ordering key or the 32-bit index formats */
#define BH_EAGAIN 7   /* serving: result not ready yet */

/* ---- Hamiltonian conventions (flags of bh_hamiltonian_build) ------------------------------------ */
#define BH_CONV_REF_HOP 1u         /* hop amplitude x2: each bond summed twice (src/hamiltonian.cpp:184-185,190) */
#define BH_CONV_REF_INTERACTION 2u /* V = sum n(n+1) (src/hamiltonian.cpp:206) instead of sum n(n-1)/2 */

/* ---- basis orders ------------------------------------------------------------------------------- */
/* BH_ORDER_REFERENCE: the order BH::fixed_set_basis actually produces.  The states are enumerated in
 * descending lexicographic order (src/hamiltonian.cpp:60-85), argsorted by tag = sum n_i ln p_i
 * (:91-114), and the argsort is applied by a cycle-swap loop (:115-122) that realises the inverse
 * permutation: the state of lexicographic rank p lands at position sortperm[p].
 * BH_ORDER_ASCENDING: ascending tag (= ascending prod p_i^n_i), the order the code intended. */
#define BH_ORDER_REFERENCE 0
#define BH_ORDER_ASCENDING 1
/* BH_ORDER_LEX: descending lexicographic enumeration order itself (the states before any sort); no
 * reference function returns it, it is offered as a solver layout whose hops stay closer together. */
#define BH_ORDER_LEX 2
/* BH_ORDER_CUSTOM: a permutation set by bh_basis_permute.  BH_ORDER_BAND: the reverse Cuthill-McKee order of a
 * lattice's hopping graph (bh_basis_reorder_band); no reference function returns either, they are solver
 * layouts like BH_ORDER_LEX.  bh_basis_build accepts neither. */
#define BH_ORDER_CUSTOM 3
#define BH_ORDER_BAND 4

/* ---- precisions --------------------------------------------------------------------------------- */
/* BH_PREC_FP64: FP64 vectors and arithmetic -- the plain comparator.  The SpMV entry points read an FP64
 *   CSR (double value + int32 column); the Lanczos engine reads the packed words below, whose FP64 table
 *   gives the same exact values in 4 bytes per entry.
 * BH_PREC_MIXED: one 32-bit word per entry (column << code_bits | code), value = -sqrt(code) read
 *   from an exact FP64 table, FP32-stored vectors, FP64 accumulation and reductions, FP64 finish. */
#define BH_PREC_FP64 0
#define BH_PREC_MIXED 1

typedef struct bh_ctx bh_ctx;
typedef struct bh_basis bh_basis;
typedef struct bh_ham bh_ham;
typedef struct bh_server bh_server;

/* ---- context -------------------------------------------------------------------------------------- */
/* hbm_budget: bytes this context may allocate on the device (0 = 90% of free device memory).
 * Replaces the memory-budgeted thread cap Resource::set_omp_threads (src/resource.cpp:77-82). */
int bh_ctx_create(int device, size_t hbm_budget, bh_ctx** out);
int bh_ctx_destroy(bh_ctx* ctx);
const char* bh_last_error(const bh_ctx* ctx);
int bh_ctx_memory(const bh_ctx* ctx, size_t* used, size_t* budget);
/* Kernel timing with HIP events on the library's stream (0 = off).  bh_ctx_kernel_stats returns
 * the launch count and total milliseconds recorded for a kernel name since the last reset. */
int bh_ctx_set_timing(bh_ctx* ctx, int on);
int bh_ctx_kernel_stats(bh_ctx* ctx, const char* name, int64_t* count, double* total_ms, int reset);
int bh_device_name(int device, char* buf, int len);

/* ---- lattice ------------------------------------------------------------------------------------ */
/* Neighbour lists as CSR: site i's neighbours are sites[off[i] .. off[i+1]-1].
 * kind 0: chain of Lx sites (Neighbours::chain_neighbours, src/neighbours.cpp:21-34);
 * kind 1: Lx x Ly square lattice, site = row*Lx + col, neighbours left/right/top/bottom
 *         (Neighbours::square_neighbours, src/neighbours.cpp:40-71, whose shadowing local at :45
 *         leaves the reference's lists empty; here they are filled and Lx may differ from Ly).
 * off needs m+1 entries, sites 4m+4.  Returns the number of entries written, or < 0 on error. */
int bh_lattice_neighbours(int kind, int Lx, int Ly, int periodic, int* off, int* sites);

/* ---- basis ------------------------------------------------------------------------------------ */
/* Replaces BH::fixed_set_basis(int m, int n) (include/hamiltonian.hpp:64, src/hamiltonian.cpp:143-149).
 * Built on the device: unranking of lexicographic ranks, exact integer keys prod p_i^n_i, LSD radix
 * sort, permutation and inverse permutation.  m <= 32, n <= 255, D = C(m+n-1, n) < 2^31. */
int bh_basis_build(bh_ctx* ctx, int m, int n, int order, bh_basis** out);
int bh_basis_destroy(bh_basis* b);
int64_t bh_basis_dim(const bh_basis* b);
/* states: D*m bytes (state k at k*m), the reference's column k of the basis matrix */
int bh_basis_copy_states(const bh_basis* b, uint8_t* states);
/* tags: D doubles, sum_i n_i * log(p_i) accumulated in site order (the reference's tags vector) */
int bh_basis_copy_tags(const bh_basis* b, double* tags);
/* lex_of_pos[k]: lexicographic rank of the state at position k; pos_of_lex: the inverse */
int bh_basis_copy_perm(const bh_basis* b, int32_t* lex_of_pos, int32_t* pos_of_lex);
/* Re-orders a basis: the state at position k becomes the state of lexicographic rank lex_of_pos[k] (D entries, a
 * permutation of 0 .. D-1, else BH_EINVAL); the order becomes BH_ORDER_CUSTOM.  Hamiltonians built on the basis
 * before the call keep the former order: build them after it.  No reference counterpart: the reference has one
 * order (BH::fixed_set_basis, src/hamiltonian.cpp:60-122) and the spectra do not depend on it. */
int bh_basis_permute(bh_ctx* ctx, bh_basis* b, const int32_t* lex_of_pos);
/* Re-orders a basis into BH_ORDER_BAND for the lattice (nbr_off, nbr_sites as for bh_hamiltonian_build): the
 * reverse Cuthill-McKee order of the hopping graph, computed on the host from a Hamiltonian built in the current
 * order, so that every hop lies close to the diagonal and the gathers of the operator pass stay in L2-sized
 * windows.  Deterministic; same conditions as bh_basis_permute. */
int bh_basis_reorder_band(bh_ctx* ctx, bh_basis* b, const int* nbr_off, const int* nbr_sites);

/* Replaces Analysis::SPDM and Analysis::braket (include/analysis.hpp:56-57, src/analysis.cpp:497-512,
 * 562-594): the single-particle density matrix rho[i*m + j] = <phi| b_i^+ b_j |phi> of a real state phi
 * (D doubles in the basis's position order, e.g. a Ritz vector of bh_lanczos_copy_vectors for a Hamiltonian
 * built on this basis).  Amplitude sqrt((n_i + 1) n_j) of the hop and exact target positions; the reference's
 * amplitude taken after the hop (src/analysis.cpp:586-588), its tolerance tag search and its 1/nb_eigen
 * scale (:511) are not reproduced.  The trace is n <phi|phi>. */
int bh_spdm(bh_ctx* ctx, const bh_basis* b, const double* phi, double* rho);

/* Replaces Analysis::SCMF with Analysis::SF_density and BH::h_MF (include/analysis.hpp:13-16,
 * src/analysis.cpp:95-178, include/hamiltonian.hpp:103-110, src/hamiltonian.cpp:290-309): the self-consistent
 * mean-field order parameter psi = <a> at npts independent points (mu[k], J[k]), U = 1, q neighbours per site,
 * starting from psi0[k].  Each point repeats: ground state of h_MF on 2p+1 occupations for p = 1, 2, ... until
 * the ground energy changes by at most eps; psi <- <a> of that state; until psi changes by at most eps (the
 * reference's eps = 10^-precision).  Arrays are host memory of npts entries.  Outputs: psi (required);
 * status (0 converged, 1 max_outer self-consistency iterations reached without convergence, psi then being
 * the last iterate, 2 p would exceed pmax), outer iterations run and the last p (each may be NULL).  The
 * reference has no iteration cap and a 1000 x 1000 buffer (p <= 499); here 2 <= pmax <= 32.  A ground state
 * that is degenerate (psi = 0 or J = 0 with tied diagonal entries) is taken as the first lowest occupation
 * state; the reference's Spectra solve returns an unspecified vector of the eigenspace there. */
int bh_mean_field(bh_ctx* ctx, int64_t npts, const double* mu, const double* J, const double* psi0, int q,
                  double eps, int max_outer, int pmax, double* psi, int32_t* status, int32_t* iters,
                  int32_t* p_last);

/* ---- Hamiltonian ------------------------------------------------------------------------------- */
/* Replaces BH::fixed_bosons_hamiltonian(nbr, basis, tags, m, n, J, U, mu) (include/hamiltonian.hpp:87,
 * src/hamiltonian.cpp:170-256) for the J and U terms: two-pass count / exclusive scan / fill CSR
 * build of T (off-diagonal, columns ascending per row, value -sqrt((n_i+1) n_j) per neighbour entry,
 * repeated entries summed) and the diagonal V.  The neighbour list must be symmetric as a multiset
 * and free of self-loops.  flags: BH_CONV_*. */
int bh_hamiltonian_build(bh_ctx* ctx, const bh_basis* b, const int* nbr_off, const int* nbr_sites, unsigned flags,
                         bh_ham** out);
int bh_ham_destroy(bh_ham* h);
int64_t bh_ham_dim(const bh_ham* h);
int64_t bh_ham_nnz(const bh_ham* h);
/* bytes the Hamiltonian occupies on the device */
int64_t bh_ham_device_bytes(const bh_ham* h);
/* host copy: row_ptr (D+1), col (nnz), val (nnz, T at J=1), diag (D, V at U=1); any may be NULL */
int bh_ham_copy_csr(const bh_ham* h, int64_t* row_ptr, int32_t* col, double* val, double* diag);
/* the packed words (column << code_bits | code) and code_bits of the mixed format */
int bh_ham_copy_packed(const bh_ham* h, uint32_t* words, int* code_bits);
/* Materialise the FP64 CSR of the FP64 SpMV entry points (needed once before their BH_PREC_FP64 use, and before
 * bh_spmv_noalloc in FP64; done lazily by bh_spmv otherwise; the Lanczos engines do not use it). */
int bh_ham_prepare(bh_ham* h, int prec);
/* No counterpart in the reference (its product has one fixed kernel): the operator-kernel variant the
 * autotuner chose for this sector at (kind, prec, B), once the sector has been used that way (kind 0: the
 * Lanczos operator pass at engine width B, kind 1: bh_spmv).  *variant: 0 = LDS-staged row blocks;
 * V = 1, 2, 4 (and 8, kind 0 only) = one row per B / V threads holding V vectors each (64 V / B rows per
 * wavefront), four entries of a row in flight per thread; V + 16 and V + 32 (kind 0, V <= 2, B <= 16), and
 * 20 (kind 0, V = 4, FP32, 4 <= B <= 32): the same with eight or sixteen entries in flight;
 * 100 + G = G threads per (row, vector).  cand_var / cand_us (capacity cap, may be NULL) receive each
 * timed candidate and its microseconds per launch, *ncand their number (0 when the default variant was
 * kept untimed: BHX_TUNE=0, BHX_ENGINE set, or a single candidate).  BH_EINVAL when not used yet. */
int bh_ham_tuned(const bh_ham* h, int kind, int prec, int B, int* variant, int* ncand, int* cand_var, double* cand_us,
                 int cap);
/* No counterpart in the reference: pins the variant used for (kind, prec, B) on this sector instead of the
 * autotuner's choice (engines already running keep theirs).  BH_EINVAL when the variant does not fit:
 * 0 needs the longest row times B within the 2048-product stage; V in {1, 2, 4, 8} needs V <= B, and
 * V = 4 and V = 8 FP32 vectors (V = 8 kind 0 only); V + 16, V + 32 need V <= 2, V <= B <= 16 (kind 0);
 * 20 needs FP32, 4 <= B <= 32 (kind 0);
 * 100 + G (kind 1 only) needs G B <= 64. */
int bh_ham_set_tuned(bh_ham* h, int kind, int prec, int B, int variant);

/* ---- operator application ---------------------------------------------------------------------- */
/* Replaces the sparse product of SparseGenMatProd::perform_op
 * (external/spectra/include/Spectra/MatOp/SparseGenMatProd.h:81-86): y_b = J_b T x_b + U_b V x_b for
 * B vectors stored row-interleaved ([D][B]).  x, y are device pointers: double for BH_PREC_FP64,
 * float for BH_PREC_MIXED (FP64 accumulation).  stream: a hipStream_t or NULL for the library's. */
int bh_spmv(bh_ctx* ctx, bh_ham* h, int prec, int B, const double* J, const double* U, const void* x, void* y,
            void* stream);
/* bh_spmv that allocates nothing.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops that forbid hipMalloc).  J and U
 * are staged in a library-owned device array of 2 x 64 doubles (one copy per GPU, held until the process
 * ends) instead of per-call buffers, so B <= 64, and calls that overlap on one GPU race on it.  The
 * handle must already hold what the precision needs (FP64: call bh_ham_prepare first; BH_EINVAL
 * otherwise); the variant is the autotuner's choice for (h, prec, B) if one was made, else the default,
 * and nothing is tuned.  Same results as bh_spmv with the same variant.  Outside stream capture the call
 * returns after its work on `stream` is done. */
int bh_spmv_noalloc(bh_ctx* ctx, bh_ham* h, int prec, int B, const double* J, const double* U, const void* x,
                    void* y, void* stream);
/* Host-array convenience: x, y are host doubles [D][B]; FP64 accumulation in both precisions. */
int bh_spmv_host(bh_ctx* ctx, bh_ham* h, int prec, int B, const double* J, const double* U, const double* x,
                 double* y);
/* Times `reps` back-to-back launches of the SpMV kernel used by bh_spmv (device-resident data,
 * HIP events on the library stream).  avg_ms: average per launch; bytes: algorithmic bytes per launch. */
int bh_spmv_bench(bh_ctx* ctx, bh_ham* h, int prec, int B, int reps, double* avg_ms, double* bytes);

/* ---- Lanczos eigensolver ------------------------------------------------------------------------ */
/* Replaces Op::IRLM_eigen(SparseMatrix O, int nb_eigen, MatrixXcd& evecs) (include/operator.hpp:34,
 * src/operator.cpp:22-33, Spectra GenEigs nev/ncv=2nev+1/SmallestReal/tol 1e-10).
 * Lanczos with selective orthogonalisation against converged Ritz vectors, FP64 alpha/beta
 * reductions, FP64 tridiagonal analysis; B independent solves (J_b, U_b) share one CSR pass.
 * Results are the lowest k eigenvalues of H_b counted with their multiplicity, ascending, as the
 * reference's Spectra solve returns them (the two members of each +-momentum pair of a periodic chain
 * appear twice).  A single start vector meets each degenerate eigenspace once, so a lane whose pass has
 * converged freezes the vectors it found and starts a further pass from a fresh vector orthogonal to
 * them; the solve ends with a pass that finds nothing new among the lowest k.  With distinct = 1 the
 * result is instead the lowest k distinct levels (one pass; values within dedup_rel*||T|| count as one). */
typedef struct bh_lanczos_opts {
    int k;             /* wanted eigenvalues, 1 <= k <= 64 (distinct levels: k <= 16) */
    int prec;          /* BH_PREC_FP64 or BH_PREC_MIXED */
    int max_iter;      /* Lanczos steps (operator applications) per solve, over all restarts (0: 10000) */
    int chunk;         /* steps between convergence checks (0: 16) */
    double tol;        /* Spectra rule: |beta s_last| < tol * max(eps^(2/3), |theta|) (0: 1e-10) */
    double dedup_rel;  /* level merge tolerance relative to ||T|| (0: 1e-9) */
    double accept_rel; /* mixed: FP64-finish acceptance of E0, ||r||^2/gap <= accept_rel*|E0| (0: 1e-9) */
    uint64_t seed;     /* Park-Miller start vector seed (0 -> 1, Spectra's convention) */
    int fallback;      /* mixed: rerun failing solves in FP64 (1, default; an FP64 rerun that is not certified
                          either is solved again with full reorthogonalisation) or report them (-1) */
    int want_vectors;  /* 1: keep the k Ritz vectors (bh_lanczos_copy_vectors) */
    int ncv;           /* Krylov vectors kept per solve (Spectra's ncv); a solve that fills them is thick-
                          restarted with its lowest Ritz vectors (0: min(D, 512); at least 2k + 8) */
    int fullre_d;      /* sectors with D <= fullre_d run FP64 Lanczos with full reorthogonalisation whatever
                          prec asks (0: 16384; -1: never); used_fp64 reports it.  With multiplicities, k > 6
                          also runs that way */
    int distinct;      /* 0 (default): eigenvalues with multiplicity; 1: distinct levels */
} bh_lanczos_opts;

typedef struct bh_lanczos_result {
    double* evals;   /* [B][k] eigenvalues (FP64-finished Rayleigh quotients), NaN if missing */
    double* resid;   /* [B][k] FP64 residual norms ||H y - theta y|| (may be NULL) */
    double* errest;  /* [B][k] a posteriori eigenvalue error bound ||r||^2 / gap (may be NULL) */
    int* iters;      /* [B] Lanczos steps taken (may be NULL) */
    int* used_fp64;  /* [B] 1 when an FP64 engine produced the result: asked for, a small sector
                        (fullre_d), or the fallback of a mixed solve (may be NULL) */
    int* status;     /* [B] per-solve BH_OK / BH_ENOCONV (may be NULL) */
} bh_lanczos_result;

void bh_lanczos_default_opts(bh_lanczos_opts* o);
int bh_lanczos(bh_ctx* ctx, bh_ham* h, int B, const double* J, const double* U, const bh_lanczos_opts* opts,
               bh_lanczos_result* res);
/* Ritz vectors of the last bh_lanczos call with want_vectors (host [B][k][D] doubles). */
int bh_lanczos_copy_vectors(bh_ctx* ctx, int B, int k, double* vecs);

/* Replaces Op::FOLM_diag (include/operator.hpp:45 with FOLM_eigen, src/operator.cpp:39-71): nb_iter steps of
 * Lanczos on J T + U V from the normalised Park-Miller start vector of `seed` (the reference draws
 * Eigen's VectorXd::Random, :77), FP64 with full reorthogonalisation and no convergence test.  Outputs the
 * tridiagonal T: alpha[t] (t < *steps) on the diagonal, beta[t] coupling t-1 and t (beta[0] = 0), and, when
 * basis is not NULL, the normalised Lanczos vectors as basis[t*D + r].  *steps is nb_iter (at most D) or,
 * when some beta_t < 1e-6 (the reference's stopping test, :56), the first such t.  The reference's defects
 * (alpha v_i never subtracted, beta stored one row early, T left uninitialised after a stop: :47-67) are
 * not reproduced.  alpha, beta need nb_iter entries, basis nb_iter * D. */
int bh_lanczos_tridiag(bh_ctx* ctx, bh_ham* h, double J, double U, int nb_iter, uint64_t seed, double* alpha, double* beta,
                       int* steps, double* basis);

/* Replaces Op::exact_eigen (include/operator.hpp:55, src/operator.cpp:92-101: a dense copy of the sparse
 * matrix and Eigen's SelfAdjointEigenSolver): all n eigenvalues of the symmetric n x n matrix given as a host
 * CSR (row_ptr[n+1], col, val; entries summed where repeated), ascending, and, when evecs is not NULL, the
 * orthonormal eigenvectors as its columns (column-major: evecs[c*n + r] is component r of eigenvector c).
 * The matrix is densified in HBM and solved by rocSOLVER's divide-and-conquer dsyevd (n*n doubles on the
 * device). */
int bh_dense_eigen(bh_ctx* ctx, int64_t n, const int64_t* row_ptr, const int32_t* col, const double* val, double* evals,
                   double* evecs);

/* ---- batched phase-diagram sweep ----------------------------------------------------------------- */
/* Replaces the grid loop of Analysis::exact_parameters / calculate_and_save (src/analysis.cpp:207-387),
 * which assembles H = H_fixed + H1 p1 + H2 p2 per grid point and solves each point on its own.
 * Here the sectors n-1, n, n+1 of one lattice are built once each (one CSR per sector) and all J
 * values of a sector are solved together, one CSR pass per Lanczos step for the whole batch.  The
 * chemical potential enters afterwards as the N-dependent shift -mu N.  Outputs per sector q
 * (0: n-1, 1: n, 2: n+1) and J point i:
 *   E[(q*nJ + i)*k + l]   lowest k eigenvalues of J T + U V (no mu term), ascending, with multiplicity
 *                         (distinct levels when opts->distinct)
 *   gap_ratio[q*nJ + i]   mean of min/max adjacent gaps over those levels (src/analysis.cpp:433-454)
 * and per J point i and mu point c:
 *   density[i*nmu + c]          N* / m with N* = argmin_{N in n-1..n+1} E0(N) - mu N (ties: smaller N)
 *   compressibility[i*nmu + c]  1 / (m (mu_plus - mu_minus)) when N* = n, NaN otherwise
 *   mu_plus[i] = E0(n+1) - E0(n), mu_minus[i] = E0(n) - E0(n-1)  (charge-gap edges of sector n)
 * Sector n-1 is skipped (NaN) when n = 0. */
typedef struct bh_sweep_spec {
    int kind;            /* lattice kind of bh_lattice_neighbours: 0 chain, 1 square */
    int Lx, Ly;          /* chain: Lx = m */
    int periodic;
    int n;               /* central boson number */
    int nJ;              /* J points (<= 64), solved in batches of `width` per sector */
    const double* J;
    double U;
    int nmu;
    const double* mu;
    unsigned flags;      /* BH_CONV_* */
    int order;           /* basis order used for the solves (BH_ORDER_LEX is the fastest) */
    int width;           /* J points per batched solve: 0 = chosen per sector by timing the operator at
                          * widths 8 .. 64 (the narrowest within 10% of the best cost per vector), else a
                          * power of two <= 64 */
} bh_sweep_spec;

typedef struct bh_sweep_out {
    double* E;               /* [3][nJ][k] */
    double* gap_ratio;       /* [3][nJ] */
    double* density;         /* [nJ][nmu] */
    double* compressibility; /* [nJ][nmu] */
    double* mu_plus;         /* [nJ] */
    double* mu_minus;        /* [nJ] */
    int* status;             /* [3][nJ] */
    int* used_fp64;          /* [3][nJ] (may be NULL) */
    int* width;              /* [3] batch width used per sector (may be NULL; 0 for an empty sector) */
} bh_sweep_out;

int bh_sweep(bh_ctx* ctx, const bh_sweep_spec* spec, const bh_lanczos_opts* opts, bh_sweep_out* out);

/* ---- serving --------------------------------------------------------------------------------------- */
/* No counterpart in the reference (its nearest analogues are the dynamic OpenMP schedule of independent
 * solves, src/analysis.cpp:302, and the memory-budgeted thread cap, src/resource.cpp:77-82).
 * A server owns a worker thread and its own device context.  Requests of uneven (lattice, m, n) are
 * grouped per sector; each sector in flight has a Lanczos engine of L lanes on its own HIP stream and a
 * cached CSR.  Lanes are refilled as soon as their solve finishes (continuous batching), sector
 * engines run concurrently on the device, and new sectors are admitted while the HBM budget (CSR plus
 * Krylov bytes) allows; idle sectors are evicted least-recently-used first.  A lane's arithmetic and
 * its analysis schedule do not depend on the other lanes, so a request returns the same bits alone or
 * in a mix served with the same lane count.  Mixed-precision results that fail the FP64 acceptance
 * test of E0 are rerun in FP64. */
typedef struct bh_request {
    int kind, Lx, Ly, periodic; /* lattice as in bh_lattice_neighbours (chain: Lx = m) */
    int n;                      /* bosons */
    double J, U, mu;            /* the levels returned include the shift -mu n */
} bh_request;

typedef struct bh_server_opts {
    int lanes;              /* lanes per sector engine, upper bound (power of two <= 64; 0: 16) */
    int prec;               /* BH_PREC_MIXED or BH_PREC_FP64 */
    size_t hbm_budget;      /* device bytes for CSR cache + engines (0: 80% of free memory) */
    size_t engine_bytes;    /* cap per engine; lanes shrink for large sectors (0: 16 GiB) */
    unsigned flags;         /* BH_CONV_* */
    bh_lanczos_opts lanczos;
} bh_server_opts;

typedef struct bh_result {
    uint64_t id;
    int status;             /* BH_OK, BH_ENOCONV, BH_EBUDGET (sector can never fit), BH_EINVAL */
    int k;
    double evals[16];       /* lowest k levels of J T + U V - mu n */
    double resid[16];
    int iters;
    int used_fp64;
    double t_submit, t_start, t_done; /* seconds on the server's clock */
} bh_result;

typedef struct bh_server_stats {
    int64_t submitted, completed, queued, in_flight;
    int engines;            /* sector engines alive */
    size_t hbm_used;        /* bytes held by the server's context */
    double p50_latency, p99_latency, mean_latency; /* seconds, completed requests */
    double busy_seconds;    /* first submit to last completion */
} bh_server_stats;

int bh_server_create(bh_ctx* ctx, const bh_server_opts* opts, bh_server** out);
int bh_server_submit(bh_server* srv, const bh_request* req, uint64_t* id);
/* BH_EAGAIN while the request is pending; the result is kept until taken once */
int bh_server_poll(bh_server* srv, uint64_t id, bh_result* out);
int bh_server_wait(bh_server* srv, uint64_t id, double timeout_s, bh_result* out);
/* waits until every submitted request has completed (BH_EAGAIN on timeout) */
int bh_server_drain(bh_server* srv, double timeout_s);
int bh_server_stats_get(bh_server* srv, bh_server_stats* st);
int bh_server_destroy(bh_server* srv);

/* Dense FP64 eigenvalues of a symmetric tridiagonal matrix (QL), exposed for tests of the host
 * analysis: d[n] diagonal, e[n-1] off-diagonal; evals ascending. */
int bh_tridiag_eigvals(int n, const double* d, const double* e, double* evals);

#ifdef __cplusplus
}
#endif
#endif /* BH_CAPI_H */
