/*
 * bh_oracle.c -- CPU statement of the reference's exact-diagonalisation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke check of
 * __graft_entry__.py and the cpu_baseline leg of bench.py load it (as oracle/liboracle.so)
 * to verify the HIP library and to time a CPU baseline.  The product library never
 * links, loads or calls it, and nothing here runs on a GPU.
 *
 * What it restates (reference file:line, paths relative to the reference repository):
 *   orc_dimension            src/hamiltonian.cpp:39-54     C(m+n-1, n)
 *   orc_chain_neighbours     src/neighbours.cpp:21-34      chain, open/periodic
 *   orc_square_neighbours    src/neighbours.cpp:40-71      square lattice; the reference fills a
 *                                                          shadowing local (line 45) and returns empty
 *                                                          lists -- here the lists are filled, and the
 *                                                          side is generalised to Lx x Ly
 *   orc_basis                src/hamiltonian.cpp:60-85     descending lexicographic enumeration
 *                            src/hamiltonian.cpp:91-123    tag = sum n_i ln p_i and its argsort; the
 *                                                          tag order equals the order of the exact
 *                                                          integer prod p_i^n_i, which is what is
 *                                                          sorted here.  The reference then applies the
 *                                                          argsort with a cycle-swap loop (:115-122)
 *                                                          that realises the INVERSE permutation: the
 *                                                          state of lexicographic rank p lands at
 *                                                          position sortperm[p] (the lexicographic rank
 *                                                          of the p-th smallest key).  That is the
 *                                                          reference order (order 0); order 1 is the
 *                                                          ascending-key order the code intended.
 *   orc_ham_count/_fill      src/hamiltonian.cpp:170-211   hopping T and interaction V.  The reference
 *                                                          loop has three defects (neighbour position
 *                                                          used as site, out-of-range tag read, 1e-3
 *                                                          tolerance search); this follows the intended
 *                                                          operator: for every site i and every entry j
 *                                                          of i's neighbour list with n_j >= 1, the
 *                                                          state a_i^+ a_j |k> gets -sqrt((n_i+1) n_j),
 *                                                          found by an exact key search.
 *   orc_lanczos              src/operator.cpp:22-33 with the Spectra start-vector convention
 *                            (Util/SimpleRandom.h:30-96, LinAlg/Arnoldi.h:147-154) and convergence
 *                            rule (GenEigsBase.h:141-158); plain Lanczos with full
 *                            reorthogonalisation instead of the implicitly restarted Arnoldi
 *   orc_lanczos_mult         src/operator.cpp:22-33: the lowest k eigenvalues with their multiplicity, as the
 *                            reference's Spectra solve returns them (deflation passes from fresh start
 *                            vectors, full reorthogonalisation, a final Rayleigh-Ritz step)
 *   orc_gap_ratios           src/analysis.cpp:433-454
 *   orc_mf_ground            src/hamiltonian.cpp:290-309 (BH::h_MF) and the ground state the reference asks
 *                            of Spectra's SymEigsSolver (src/analysis.cpp:128-133), with SF_density
 *                            (:95-105) of it: lowest eigenvalue by a bracketed Newton iteration on the
 *                            LDL^T pivots, <a> by Hellmann-Feynman from the same pivots
 *   orc_scmf                 src/analysis.cpp:95-178 (Analysis::SCMF, SF_density): the self-consistent
 *                            loop with the reference's two stopping rules; a cap on the outer iterations
 *                            and on p (the reference has none; status 1 / 2 when reached)
 *   orc_spdm                 src/analysis.cpp:497-512, 562-594 (Analysis::SPDM / braket): <phi| b_i^+ b_j
 *                            |phi> with the intended amplitude sqrt((n_i+1) n_j) and exact key search (the
 *                            reference multiplies by the occupations after the hop, :586-588, finds the
 *                            target with its tolerance search and divides by nb_eigen, :511)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

#define ORC_OK 0
#define ORC_EINVAL 1
#define ORC_ENOMEM 2
#define ORC_ERANGE 6
#define ORC_ENOCONV 4

/* conventions (same bit values as the library's BH_CONV_* flags) */
#define ORC_CONV_REF_HOP 1u         /* each bond counted twice, as the reference's triplet fill does */
#define ORC_CONV_REF_INTERACTION 2u /* diagonal sum n(n+1) (reference) instead of n(n-1)/2 */

static const int orc_primes[32] = {2,  3,  5,  7,  11, 13, 17, 19, 23,  29,  31,  37,  41,  43,  47,  53,
                                   59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131};

int orc_max_sites(void) { return 32; }

/* ---- combinatorics ------------------------------------------------------------------------ */

int64_t orc_binom(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    u128 r = 1;
    for (int i = 1; i <= k; ++i) {
        r = r * (u128)(n - k + i) / (u128)i;
        if (r > (u128)INT64_MAX) return -1;
    }
    return (int64_t)r;
}

int64_t orc_dimension(int m, int n) { return orc_binom(m + n - 1, n); }

/* ---- neighbours ---------------------------------------------------------------------------- */

/* off has m+1 entries; sites must hold at least 2m+2.  Returns the number of entries. */
int orc_chain_neighbours(int m, int periodic, int *off, int *sites) {
    if (m < 1) return -1;
    int cnt[64][4];
    int len[64];
    if (m > 64) return -1;
    for (int i = 0; i < m; ++i) {
        len[i] = 0;
        if (i > 0) cnt[i][len[i]++] = i - 1;
        if (i < m - 1) cnt[i][len[i]++] = i + 1;
    }
    if (periodic) {
        cnt[0][len[0]++] = m - 1;
        cnt[m - 1][len[m - 1]++] = 0;
    }
    int t = 0;
    for (int i = 0; i < m; ++i) {
        off[i] = t;
        for (int a = 0; a < len[i]; ++a) sites[t++] = cnt[i][a];
    }
    off[m] = t;
    return t;
}

/* Lx columns, Ly rows, site = row*Lx + col.  Order per site: left, right, top, bottom. */
int orc_square_neighbours(int Lx, int Ly, int periodic, int *off, int *sites) {
    if (Lx < 1 || Ly < 1 || Lx * Ly > 64) return -1;
    int t = 0;
    for (int i = 0; i < Ly; ++i) {
        for (int j = 0; j < Lx; ++j) {
            int idx = i * Lx + j;
            off[idx] = t;
            if (j > 0) sites[t++] = idx - 1;
            else if (periodic) sites[t++] = idx + Lx - 1;
            if (j < Lx - 1) sites[t++] = idx + 1;
            else if (periodic) sites[t++] = idx - Lx + 1;
            if (i > 0) sites[t++] = idx - Lx;
            else if (periodic) sites[t++] = idx + (Ly - 1) * Lx;
            if (i < Ly - 1) sites[t++] = idx + Lx;
            else if (periodic) sites[t++] = idx - (Ly - 1) * Lx;
        }
    }
    off[Lx * Ly] = t;
    return t;
}

/* ---- basis --------------------------------------------------------------------------------- */

static int key_fits(int m, int n) {
    /* the largest key is p_{m-1}^n (all bosons on the last site); it must stay below 2^128 */
    double bits = (double)n * log2((double)orc_primes[m - 1]);
    return bits < 127.5;
}

int orc_key_fits(int m, int n) { return (m >= 1 && m <= 32 && n >= 0) ? key_fits(m, n) : 0; }

static u128 state_key(const uint8_t *s, int m) {
    u128 k = 1;
    for (int i = 0; i < m; ++i)
        for (int a = 0; a < s[i]; ++a) k *= (u128)orc_primes[i];
    return k;
}

typedef struct {
    u128 key;
    int64_t lex;
} keyed_t;

static int cmp_keyed(const void *a, const void *b) {
    const keyed_t *x = (const keyed_t *)a, *y = (const keyed_t *)b;
    if (x->key < y->key) return -1;
    if (x->key > y->key) return 1;
    return (x->lex < y->lex) ? -1 : (x->lex > y->lex); /* keys are distinct; kept stable anyway */
}

/*
 * Fixed-N basis.
 *   order  : 0 = the reference's order (see the header), 1 = ascending prod p_i^n_i
 *   states : D*m bytes, state k at states[k*m ...]
 *   tags   : D doubles, sum n_i * log(p_i) accumulated in site order (may be NULL)
 *   lex    : D int64, lexicographic rank of the k-th state (may be NULL)
 */
int orc_basis(int m, int n, int order, uint8_t *states, double *tags, int64_t *lex) {
    if (m < 1 || m > 32 || n < 0 || n > 255) return ORC_EINVAL;
    if (!key_fits(m, n)) return ORC_ERANGE;
    int64_t D = orc_dimension(m, n);
    if (D <= 0) return ORC_EINVAL;
    uint8_t *lexs = (uint8_t *)malloc((size_t)D * m);
    keyed_t *kv = (keyed_t *)malloc(sizeof(keyed_t) * (size_t)D);
    if (!lexs || !kv) {
        free(lexs);
        free(kv);
        return ORC_ENOMEM;
    }
    /* descending lexicographic enumeration starting at |n,0,...,0> */
    uint8_t s[64];
    memset(s, 0, sizeof s);
    s[0] = (uint8_t)n;
    int64_t c = 0;
    for (;;) {
        memcpy(lexs + (size_t)c * m, s, m);
        ++c;
        int k = m - 2, moved = 0;
        for (; k >= 0; --k) {
            if (s[k] != 0) {
                int pre = 0;
                s[k] -= 1;
                for (int a = 0; a <= k; ++a) pre += s[a];
                s[k + 1] = (uint8_t)(n - pre);
                for (int a = k + 2; a < m; ++a) s[a] = 0;
                moved = 1;
                break;
            }
        }
        if (!moved) break;
    }
    if (c != D) {
        free(lexs);
        free(kv);
        return ORC_EINVAL;
    }
    for (int64_t r = 0; r < D; ++r) {
        kv[r].key = state_key(lexs + (size_t)r * m, m);
        kv[r].lex = r;
    }
    qsort(kv, (size_t)D, sizeof(keyed_t), cmp_keyed);
    /* position k holds: order 1 -> the k-th smallest key; order 0 -> the state of lex rank p
       with sortperm[p] == k */
    int64_t *src_lex = (int64_t *)malloc(sizeof(int64_t) * (size_t)D);
    if (!src_lex) {
        free(lexs);
        free(kv);
        return ORC_ENOMEM;
    }
    for (int64_t k = 0; k < D; ++k) {
        if (order == 1) src_lex[k] = kv[k].lex;
        else src_lex[kv[k].lex] = k; /* ref[sortperm[p]] = L(p) */
    }
    for (int64_t k = 0; k < D; ++k) {
        const uint8_t *src = lexs + (size_t)src_lex[k] * m;
        if (states) memcpy(states + (size_t)k * m, src, m);
        if (lex) lex[k] = src_lex[k];
        if (tags) {
            double t = 0.0;
            for (int i = 0; i < m; ++i) t += (double)src[i] * log((double)orc_primes[i]);
            tags[k] = t;
        }
    }
    free(src_lex);
    free(lexs);
    free(kv);
    return ORC_OK;
}

/* ---- Hamiltonian terms --------------------------------------------------------------------- */

/* (key, position) pairs of the states array sorted by key: an exact lookup for any basis order */
static int64_t find_key(const keyed_t *keys, int64_t D, u128 k) {
    int64_t lo = 0, hi = D - 1;
    while (lo <= hi) {
        int64_t mid = lo + (hi - lo) / 2;
        if (keys[mid].key == k) return keys[mid].lex;
        if (keys[mid].key < k) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

static int check_nbr(int m, const int *off, const int *sites) {
    if (off[0] != 0) return 0;
    for (int i = 0; i < m; ++i) {
        if (off[i + 1] < off[i]) return 0;
        for (int a = off[i]; a < off[i + 1]; ++a)
            if (sites[a] < 0 || sites[a] >= m || sites[a] == i) return 0;
    }
    return 1;
}

typedef struct {
    int64_t col;
    double val;
} entry_t;

static int cmp_entry(const void *a, const void *b) {
    const entry_t *x = (const entry_t *)a, *y = (const entry_t *)b;
    return (x->col < y->col) ? -1 : (x->col > y->col);
}

/*
 * Builds row k of T (tag order) into ent[]; duplicate targets (repeated neighbour entries)
 * are summed, as the reference's setFromTriplets sums repeated triplets.  Returns the count.
 */
static int build_row(int m, const uint8_t *s, const int *off, const int *sites, const keyed_t *keys, int64_t D,
                     double hop, entry_t *ent) {
    int cnt = 0;
    uint8_t t[64];
    for (int i = 0; i < m; ++i) {
        for (int a = off[i]; a < off[i + 1]; ++a) {
            int j = sites[a];
            if (s[j] < 1) continue;
            memcpy(t, s, m);
            t[i] += 1;
            t[j] -= 1;
            int64_t c = find_key(keys, D, state_key(t, m));
            if (c < 0) return -1;
            double v = -hop * sqrt((double)(s[i] + 1) * (double)s[j]);
            ent[cnt].col = c;
            ent[cnt].val = v;
            ++cnt;
        }
    }
    qsort(ent, (size_t)cnt, sizeof(entry_t), cmp_entry);
    int w = 0;
    for (int a = 0; a < cnt; ++a) {
        if (w > 0 && ent[w - 1].col == ent[a].col) ent[w - 1].val += ent[a].val;
        else ent[w++] = ent[a];
    }
    return w;
}

static keyed_t *make_keys(const uint8_t *states, int m, int64_t D) {
    keyed_t *keys = (keyed_t *)malloc(sizeof(keyed_t) * (size_t)D);
    if (!keys) return NULL;
    for (int64_t k = 0; k < D; ++k) {
        keys[k].key = state_key(states + (size_t)k * m, m);
        keys[k].lex = k;
    }
    qsort(keys, (size_t)D, sizeof(keyed_t), cmp_keyed);
    return keys;
}

/* row_ptr has D+1 entries.  Returns nnz (>= 0) or a negative error. */
int64_t orc_ham_count(int m, int n, const uint8_t *states, const int *off, const int *sites, int64_t *row_ptr) {
    int64_t D = orc_dimension(m, n);
    if (!check_nbr(m, off, sites)) return -ORC_EINVAL;
    keyed_t *keys = make_keys(states, m, D);
    if (!keys) return -ORC_ENOMEM;
    int nmax = off[m];
    entry_t *ent = (entry_t *)malloc(sizeof(entry_t) * (size_t)(nmax + 1));
    int64_t tot = 0;
    row_ptr[0] = 0;
    for (int64_t k = 0; k < D; ++k) {
        int c = build_row(m, states + (size_t)k * m, off, sites, keys, D, 1.0, ent);
        if (c < 0) {
            free(ent);
            free(keys);
            return -ORC_EINVAL;
        }
        tot += c;
        row_ptr[k + 1] = tot;
    }
    free(ent);
    free(keys);
    return tot;
}

/*
 * Fills the CSR of T (the hopping term at J=1, off-diagonal, columns ascending) and the
 * diagonal V (U=1): V = sum n(n-1)/2, or sum n(n+1) under ORC_CONV_REF_INTERACTION.
 */
int orc_ham_fill(int m, int n, const uint8_t *states, const int *off, const int *sites, unsigned flags,
                 const int64_t *row_ptr, int32_t *col, double *val, double *diag) {
    int64_t D = orc_dimension(m, n);
    if (!check_nbr(m, off, sites)) return ORC_EINVAL;
    keyed_t *keys = make_keys(states, m, D);
    if (!keys) return ORC_ENOMEM;
    double hop = (flags & ORC_CONV_REF_HOP) ? 2.0 : 1.0;
    int nmax = off[m];
    entry_t *ent = (entry_t *)malloc(sizeof(entry_t) * (size_t)(nmax + 1));
    for (int64_t k = 0; k < D; ++k) {
        const uint8_t *s = states + (size_t)k * m;
        int c = build_row(m, s, off, sites, keys, D, hop, ent);
        if (c < 0 || c != row_ptr[k + 1] - row_ptr[k]) {
            free(ent);
            free(keys);
            return ORC_EINVAL;
        }
        for (int a = 0; a < c; ++a) {
            col[row_ptr[k] + a] = (int32_t)ent[a].col;
            val[row_ptr[k] + a] = ent[a].val;
        }
        double v = 0.0;
        for (int i = 0; i < m; ++i) {
            double ni = (double)s[i];
            v += (flags & ORC_CONV_REF_INTERACTION) ? ni * (ni + 1.0) : 0.5 * ni * (ni - 1.0);
        }
        diag[k] = v;
    }
    free(ent);
    free(keys);
    return ORC_OK;
}

/* ---- operator application ------------------------------------------------------------------ */

/* y = J * T x + U * (V o x), rows in order, entries of a row in column order */
void orc_spmv(int64_t D, const int64_t *row_ptr, const int32_t *col, const double *val, const double *diag,
              double J, double U, const double *x, double *y) {
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < D; ++r) {
        double acc = 0.0;
        for (int64_t a = row_ptr[r]; a < row_ptr[r + 1]; ++a) acc += val[a] * x[col[a]];
        y[r] = J * acc + U * diag[r] * x[r];
    }
}

/* ---- Spectra's start vector ---------------------------------------------------------------- */

/* Park-Miller minimal standard generator, x <- 16807 x mod (2^31-1), seed 0 mapped to 1;
 * values x/(2^31-1) - 0.5 (Util/SimpleRandom.h:30-96). */
void orc_start_vector(int64_t D, uint64_t seed, double *r) {
    const uint64_t p = 2147483647ull;
    uint64_t x = seed ? (seed & p) : 1;
    for (int64_t i = 0; i < D; ++i) {
        x = (16807ull * x) % p;
        r[i] = (double)x / (double)p - 0.5;
    }
}

/* ---- symmetric tridiagonal eigensolver (implicit QL with Wilkinson shifts) ------------------- */

/*
 * d[0..n-1] diagonal, e[0..n-2] off-diagonal.  On return d holds the eigenvalues (ascending) and,
 * when z != NULL, z (n*n, column-major, z[i + n*c] = component i of vector c) the eigenvectors.
 * The implicit QL algorithm with Wilkinson shifts of Bowdler, Martin, Reinsch and Wilkinson (Numer. Math.
 * 11, 293-306, 1968): the procedures tql1 (values) and tql2 (vectors) of the Handbook for Automatic
 * Computation vol. II and of EISPACK, whose step sequence the loop below follows.
 */
int orc_tridiag_eig(int n, double *d, const double *e_in, double *z) {
    double *e = (double *)calloc((size_t)n + 1, sizeof(double));
    if (!e) return ORC_ENOMEM;
    for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
    if (z) {
        memset(z, 0, sizeof(double) * (size_t)n * n);
        for (int i = 0; i < n; ++i) z[i + (size_t)n * i] = 1.0;
    }
    for (int l = 0; l < n; ++l) {
        int iter = 0;
        for (;;) {
            int mm;
            for (mm = l; mm < n - 1; ++mm) {
                double dd = fabs(d[mm]) + fabs(d[mm + 1]);
                if (fabs(e[mm]) <= 1e-300 + 2.2e-16 * dd) break;
            }
            if (mm == l) break;
            if (++iter > 200) {
                free(e);
                return ORC_ENOCONV;
            }
            /* Wilkinson shift from the leading 2x2 block */
            double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(g, 1.0);
            g = d[mm] - d[l] + e[l] / (g + (g >= 0 ? r : -r));
            double s = 1.0, c = 1.0, p = 0.0;
            int i;
            int deflated = 0;
            for (i = mm - 1; i >= l; --i) {
                double f = s * e[i], b = c * e[i];
                r = hypot(f, g);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[mm] = 0.0;
                    deflated = 1;
                    break;
                }
                s = f / r;
                c = g / r;
                g = d[i + 1] - p;
                r = (d[i] - g) * s + 2.0 * c * b;
                p = s * r;
                d[i + 1] = g + p;
                g = c * r - b;
                if (z) {
                    for (int k = 0; k < n; ++k) {
                        double zi1 = z[k + (size_t)n * (i + 1)], zi = z[k + (size_t)n * i];
                        z[k + (size_t)n * (i + 1)] = s * zi + c * zi1;
                        z[k + (size_t)n * i] = c * zi - s * zi1;
                    }
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = g;
            e[mm] = 0.0;
        }
    }
    free(e);
    /* selection sort ascending, carrying vectors */
    for (int i = 0; i < n - 1; ++i) {
        int k = i;
        for (int j = i + 1; j < n; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            double t = d[i];
            d[i] = d[k];
            d[k] = t;
            if (z)
                for (int a = 0; a < n; ++a) {
                    double u = z[a + (size_t)n * i];
                    z[a + (size_t)n * i] = z[a + (size_t)n * k];
                    z[a + (size_t)n * k] = u;
                }
        }
    }
    return ORC_OK;
}

/* ---- Lanczos ------------------------------------------------------------------------------- */

static double dot(int64_t D, const double *a, const double *b) {
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int64_t i = 0; i < D; ++i) s += a[i] * b[i];
    return s;
}

/*
 * Lowest k distinct eigenvalues of H = J*T + U*V (distinct: Ritz values closer than
 * dedup_rel * ||T_j|| count once; a single start vector sees each degenerate level once).
 * FP64 throughout, full reorthogonalisation (two classical Gram-Schmidt passes per step).
 * Start: r from orc_start_vector(seed), v0 = H r / ||H r|| (Spectra's Arnoldi init).
 * Converged when |beta_j s_{j,i}| < tol * max(eps^(2/3), |theta_i|) for the k wanted levels.
 * Returns ORC_OK, or ORC_ENOCONV after max_iter steps.
 */
int orc_lanczos(int64_t D, const int64_t *row_ptr, const int32_t *col, const double *val, const double *diag,
                double J, double U, int k, double tol, int max_iter, uint64_t seed, double dedup_rel, double *evals,
                double *bounds, int *iters) {
    if (k < 1 || max_iter < 1 || D < 1) return ORC_EINVAL;
    if (max_iter > D) max_iter = (int)D;
    double *V = (double *)malloc(sizeof(double) * (size_t)D * (size_t)(max_iter + 1));
    double *w = (double *)malloc(sizeof(double) * (size_t)D);
    double *alpha = (double *)malloc(sizeof(double) * (size_t)(max_iter + 1));
    double *beta = (double *)malloc(sizeof(double) * (size_t)(max_iter + 2));
    double *h = (double *)malloc(sizeof(double) * (size_t)(max_iter + 1));
    double *td = (double *)malloc(sizeof(double) * (size_t)(max_iter + 1));
    double *z = NULL;
    int rc = ORC_ENOCONV;
    if (!V || !w || !alpha || !beta || !h || !td) {
        rc = ORC_ENOMEM;
        goto out;
    }
    orc_start_vector(D, seed, w);
    orc_spmv(D, row_ptr, col, val, diag, J, U, w, V);
    {
        double nrm = sqrt(dot(D, V, V));
        if (nrm == 0.0) {
            /* H r = 0: fall back to r itself */
            orc_start_vector(D, seed, V);
            nrm = sqrt(dot(D, V, V));
        }
        for (int64_t i = 0; i < D; ++i) V[i] /= nrm;
    }
    const double eps = 2.220446049250313e-16;
    const double eps23 = pow(eps, 2.0 / 3.0);
    beta[0] = 0.0;
    int j;
    for (j = 0; j < max_iter; ++j) {
        double *vj = V + (size_t)D * j;
        orc_spmv(D, row_ptr, col, val, diag, J, U, vj, w);
        /* two passes of classical Gram-Schmidt against v_0..v_j */
        for (int pass = 0; pass < 2; ++pass) {
            for (int t = 0; t <= j; ++t) h[t] = dot(D, V + (size_t)D * t, w);
            if (pass == 0) alpha[j] = h[j];
            else alpha[j] += h[j];
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < D; ++i) {
                double s = 0.0;
                for (int t = 0; t <= j; ++t) s += V[(size_t)D * t + i] * h[t];
                w[i] -= s;
            }
        }
        double b = sqrt(dot(D, w, w));
        beta[j + 1] = b;
        int nstep = j + 1;
        /* an invariant subspace (tiny beta) or the whole space ends the iteration */
        double anorm = fabs(alpha[0]);
        for (int t = 0; t < nstep; ++t) anorm = fmax(anorm, fabs(alpha[t]) + beta[t + 1]);
        int exhausted = (b <= 1e-12 * anorm) || nstep == D;
        int check = exhausted || nstep == max_iter || (nstep >= k + 2 && nstep % 5 == 0);
        if (check) {
            free(z);
            z = (double *)malloc(sizeof(double) * (size_t)nstep * nstep);
            for (int t = 0; t < nstep; ++t) td[t] = alpha[t];
            if (orc_tridiag_eig(nstep, td, beta + 1, z) != ORC_OK) {
                rc = ORC_ENOCONV;
                goto out;
            }
            double tnorm = fmax(fabs(td[0]), fabs(td[nstep - 1]));
            /* walk distinct levels from the bottom */
            int got = 0, ok = 1;
            for (int t = 0; t < nstep && got < k; ++t) {
                if (got > 0 && fabs(td[t] - evals[got - 1]) <= dedup_rel * tnorm) continue;
                double bnd = fabs(b * z[(nstep - 1) + (size_t)nstep * t]);
                double thr = tol * fmax(eps23, fabs(td[t]));
                evals[got] = td[t];
                if (bounds) bounds[got] = bnd;
                if (bnd >= thr) ok = 0;
                ++got;
            }
            if ((got == k && ok) || exhausted || nstep == max_iter) {
                for (int t = got; t < k; ++t) {
                    evals[t] = NAN;
                    if (bounds) bounds[t] = NAN;
                }
                *iters = nstep;
                rc = (got == k && (ok || exhausted)) ? ORC_OK : ORC_ENOCONV;
                goto out;
            }
        }
        if (j + 1 < max_iter) {
            double *vn = V + (size_t)D * (j + 1);
            for (int64_t i = 0; i < D; ++i) vn[i] = w[i] / b;
        }
    }
    *iters = max_iter;
out:
    free(V);
    free(w);
    free(alpha);
    free(beta);
    free(h);
    free(td);
    free(z);
    return rc;
}

/* ---- eigenvalues with multiplicity ---------------------------------------------------------- */

/* x_{off+1} .. x_{off+D} of the Park-Miller sequence of orc_start_vector (x_i = 16807^i x_0 mod 2^31-1) as
 * values x / (2^31 - 1) - 0.5: the sequence continued past its first off values */
void orc_start_vector_at(int64_t D, uint64_t seed, uint64_t off, double *r) {
    const uint64_t p = 2147483647ull;
    uint64_t e = off, base = 16807ull, pw = 1;
    while (e) {
        if (e & 1) pw = (pw * base) % p;
        base = (base * base) % p;
        e >>= 1;
    }
    uint64_t x = (pw * (seed ? (seed & p) : 1)) % p;
    for (int64_t i = 0; i < D; ++i) {
        x = (16807ull * x) % p;
        r[i] = (double)x / (double)p - 0.5;
    }
}

/* z -= sum_c <Q_c, z> Q_c over the nq columns of Q and the nr columns of R (D x n, column-major), with both
 * sets of coefficients taken from the same z (two passes); returns ||z||.  Projecting out one set after the
 * other would let the second subtraction put back components along the first. */
static double orc_cgs2(int64_t D, const double *Q, int nq, const double *R, int nr, double *z, double *h) {
    for (int pass = 0; pass < 2; ++pass) {
        for (int c = 0; c < nq; ++c) h[c] = dot(D, Q + (size_t)D * c, z);
        for (int c = 0; c < nr; ++c) h[nq + c] = dot(D, R + (size_t)D * c, z);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < D; ++i) {
            double s = 0.0;
            for (int c = 0; c < nq; ++c) s += Q[(size_t)D * c + i] * h[c];
            for (int c = 0; c < nr; ++c) s += R[(size_t)D * c + i] * h[nq + c];
            z[i] -= s;
        }
    }
    return sqrt(dot(D, z, z));
}

/* eigenpairs of a small symmetric matrix A (n x n, row-major, destroyed) by cyclic Jacobi rotations:
 * w ascending, V[i*n + c] = component i of eigenvector c */
static void orc_jacobi(int n, double *A, double *w, double *V) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, tot = 0.0;
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                tot += A[i * n + j] * A[i * n + j];
                if (i != j) off += A[i * n + j] * A[i * n + j];
            }
        if (off <= 1e-32 * tot || off == 0.0) break;
        for (int p = 0; p < n; ++p)
            for (int q = p + 1; q < n; ++q) {
                const double apq = A[p * n + q];
                if (apq == 0.0) continue;
                const double th = (A[q * n + q] - A[p * n + p]) / (2.0 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
                for (int r = 0; r < n; ++r) {
                    const double arp = A[r * n + p], arq = A[r * n + q];
                    A[r * n + p] = c * arp - sn * arq;
                    A[r * n + q] = sn * arp + c * arq;
                }
                for (int r = 0; r < n; ++r) {
                    const double apr = A[p * n + r], aqr = A[q * n + r];
                    A[p * n + r] = c * apr - sn * aqr;
                    A[q * n + r] = sn * apr + c * aqr;
                }
                for (int r = 0; r < n; ++r) {
                    const double vrp = V[r * n + p], vrq = V[r * n + q];
                    V[r * n + p] = c * vrp - sn * vrq;
                    V[r * n + q] = sn * vrp + c * vrq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
    /* ascending order (selection sort of values and vector columns) */
    for (int i = 0; i < n; ++i) {
        int m = i;
        for (int j = i + 1; j < n; ++j)
            if (w[j] < w[m]) m = j;
        if (m == i) continue;
        double t = w[i];
        w[i] = w[m];
        w[m] = t;
        for (int r = 0; r < n; ++r) {
            t = V[r * n + i];
            V[r * n + i] = V[r * n + m];
            V[r * n + m] = t;
        }
    }
}

static int cmp_double(const void *a, const void *b) {
    const double x = *(const double *)a, y = *(const double *)b;
    return (x > y) - (x < y);
}

/*
 * Lowest k eigenvalues of H = J*T + U*V counted with their multiplicity, as Op::IRLM_eigen returns them
 * (src/operator.cpp:22-33: Spectra GenEigs, nev = k, SmallestReal, tol).  FP64, full reorthogonalisation.
 * A single start vector meets each eigenspace once, so the method runs passes: pass 0 starts from
 * H r / ||H r|| (Spectra's Arnoldi init, r = orc_start_vector(seed)), pass p > 0 from the Park-Miller sequence
 * continued past p*D values; every vector of a pass is orthogonalised (two classical Gram-Schmidt passes)
 * against the vectors frozen by the earlier passes and against the pass's own basis.  A pass ends when its
 * distinct levels (Ritz values closer than 1e-10 ||T|| count once) at or below kappa -- the k-th lowest of the
 * frozen values and the pass's levels -- and its lowest level have converged (Spectra's rule
 * |beta s_last| < tol max(eps^(2/3), |theta|), GenEigsBase.h:141-158).  Its levels at or below kappa are then
 * frozen (Ritz vectors, orthonormalised against the frozen set; frozen values now above kappa are dropped)
 * and a new pass starts; a pass that adds nothing at or below kappa ends the solve.  A last Rayleigh-Ritz step over the frozen vectors and the last
 * pass's levels gives evals (ascending) and resid (||H y - theta y||).  max_iter bounds the steps of a pass.
 */
int orc_lanczos_mult(int64_t D, const int64_t *row_ptr, const int32_t *col, const double *val, const double *diag,
                     double J, double U, int k, double tol, int max_iter, uint64_t seed, double *evals, double *resid,
                     int *iters) {
    if (k < 1 || max_iter < 1 || D < 1) return ORC_EINVAL;
    if (max_iter > D) max_iter = (int)D;
    const int fcap = 2 * k + 8;
    const double eps = 2.220446049250313e-16, eps23 = pow(eps, 2.0 / 3.0);
    double *F = (double *)malloc(sizeof(double) * (size_t)D * (size_t)fcap);   /* frozen vectors */
    double *V = (double *)malloc(sizeof(double) * (size_t)D * (size_t)(max_iter + 1));
    double *w = (double *)malloc(sizeof(double) * (size_t)D);
    double *alpha = (double *)malloc(sizeof(double) * (size_t)(max_iter + 1));
    double *beta = (double *)malloc(sizeof(double) * (size_t)(max_iter + 2));
    double *h = (double *)malloc(sizeof(double) * (size_t)(max_iter + fcap + 1));
    double *td = (double *)malloc(sizeof(double) * (size_t)(max_iter + 1));
    double *fth = (double *)malloc(sizeof(double) * (size_t)fcap);
    double *z = NULL, *mv = NULL;
    int *lev = (int *)malloc(sizeof(int) * (size_t)(max_iter + 1));
    int rc = ORC_ENOCONV, nf = 0, total = 0, last_n = 0, nlast = 0;
    int *last = (int *)malloc(sizeof(int) * (size_t)(k + 1));
    if (!F || !V || !w || !alpha || !beta || !h || !td || !fth || !lev || !last) {
        rc = ORC_ENOMEM;
        goto out;
    }
    for (int pass = 0;; ++pass) {
        if (pass == 0) {
            orc_start_vector(D, seed, w);
            orc_spmv(D, row_ptr, col, val, diag, J, U, w, V);
            if (sqrt(dot(D, V, V)) == 0.0) orc_start_vector(D, seed, V);
        } else {
            orc_start_vector_at(D, seed, (uint64_t)pass * (uint64_t)D, V);
        }
        double nrm = orc_cgs2(D, F, nf, NULL, 0, V, h);
        if (nrm <= 1e-10) { /* the frozen vectors span the whole space */
            rc = ORC_OK;
            last_n = 0;
            break;
        }
        for (int64_t i = 0; i < D; ++i) V[i] /= nrm;
        beta[0] = 0.0;
        int done = 0, nstep = 0;
        for (int j = 0; j < max_iter && !done; ++j) {
            double *vj = V + (size_t)D * j;
            orc_spmv(D, row_ptr, col, val, diag, J, U, vj, w);
            alpha[j] = dot(D, vj, w);
            /* against the frozen vectors and the pass's basis together */
            const double b = orc_cgs2(D, F, nf, V, j + 1, w, h);
            beta[j + 1] = b;
            nstep = j + 1;
            ++total;
            double anorm = fabs(alpha[0]);
            for (int t = 0; t < nstep; ++t) anorm = fmax(anorm, fabs(alpha[t]) + beta[t + 1]);
            const int exhausted = (b <= 1e-12 * anorm) || nstep + nf >= D;
            if (exhausted || nstep == max_iter || (nstep >= 3 && nstep % 5 == 0)) {
                free(z);
                z = (double *)malloc(sizeof(double) * (size_t)nstep * nstep);
                for (int t = 0; t < nstep; ++t) td[t] = alpha[t];
                if (!z || orc_tridiag_eig(nstep, td, beta + 1, z) != ORC_OK) {
                    rc = ORC_ENOCONV;
                    goto out;
                }
                const double tnorm = fmax(fabs(td[0]), fabs(td[nstep - 1]));
                int nl = 0;
                for (int t = 0; t < nstep && nl < k; ++t) {
                    if (nl > 0 && fabs(td[t] - td[lev[nl - 1]]) <= 1e-10 * tnorm) continue;
                    lev[nl++] = t;
                }
                free(mv);
                mv = (double *)malloc(sizeof(double) * (size_t)(nf + nl + 1));
                for (int q = 0; q < nf; ++q) mv[q] = fth[q];
                for (int q = 0; q < nl; ++q) mv[nf + q] = td[lev[q]];
                qsort(mv, (size_t)(nf + nl), sizeof(double), cmp_double);
                const double kappa = nf + nl >= k ? mv[k - 1] : INFINITY;
                const double marg = 1e-10 * tnorm;
                int conv = nf + nl >= k || exhausted; /* every Ritz value of T is analysed */
                for (int q = 0; q < nl; ++q) {
                    const double bnd = fabs(b * z[(nstep - 1) + (size_t)nstep * lev[q]]);
                    const int below = td[lev[q]] <= kappa + marg;
                    if ((below || q == 0) && bnd >= tol * fmax(eps23, fabs(td[lev[q]])) && !exhausted) conv = 0;
                }
                if (!conv && nstep < max_iter) goto next_step;
                /* the pass's levels at or below kappa */
                int nnew = 0;
                for (int q = 0; q < nl; ++q)
                    if (td[lev[q]] <= kappa + marg) last[nnew < k ? nnew : k] = lev[q], ++nnew;
                if (nnew > k) nnew = k;
                int nkeep = 0;
                for (int q = 0; q < nf; ++q) nkeep += fth[q] <= kappa + marg;
                if (nnew == 0 || nkeep + nnew > fcap || nkeep + nnew >= D || nstep == max_iter) {
                    /* the solve ends: the last pass's levels at or below kappa (and its lowest) join the RR */
                    last_n = nstep;
                    nlast = 0;
                    for (int q = 0; q < nl && nlast < k; ++q)
                        if (q == 0 || td[lev[q]] <= kappa + marg) last[nlast++] = lev[q];
                    rc = (nstep == max_iter && !conv) || (nnew > 0 && nkeep + nnew > fcap) ? ORC_ENOCONV : ORC_OK;
                    done = 2;
                    break;
                }
                /* frozen values above kappa are no longer among the lowest k: their vectors are dropped (a later
                 * pass may meet them again, above kappa) */
                {
                    int keep = 0;
                    for (int q = 0; q < nf; ++q) {
                        if (fth[q] > kappa + marg) continue;
                        if (keep != q) {
                            memcpy(F + (size_t)D * keep, F + (size_t)D * q, sizeof(double) * (size_t)D);
                            fth[keep] = fth[q];
                        }
                        ++keep;
                    }
                    nf = keep;
                }
                /* freeze: Ritz vectors V z, orthonormalised against the frozen set */
                for (int q = 0; q < nnew; ++q) {
                    double *y = F + (size_t)D * nf;
                    const double *zc = z + (size_t)nstep * last[q];
#pragma omp parallel for schedule(static)
                    for (int64_t i = 0; i < D; ++i) {
                        double sacc = 0.0;
                        for (int t = 0; t < nstep; ++t) sacc += V[(size_t)D * t + i] * zc[t];
                        y[i] = sacc;
                    }
                    const double yn = orc_cgs2(D, F, nf, NULL, 0, y, h);
                    for (int64_t i = 0; i < D; ++i) y[i] /= yn;
                    fth[nf++] = td[last[q]];
                }
                done = 1;
                break;
            }
        next_step:
            if (j + 1 < max_iter) {
                double *vn = V + (size_t)D * (j + 1);
                for (int64_t i = 0; i < D; ++i) vn[i] = w[i] / b;
            }
        }
        if (done == 2) break;
        if (!done) {
            rc = ORC_ENOCONV;
            last_n = 0;
            break;
        }
    }
    {
        /* Rayleigh-Ritz over the frozen vectors and the last pass's Ritz vectors */
        const int m = nf + (last_n > 0 ? nlast : 0);
        double *Y = (double *)malloc(sizeof(double) * (size_t)D * (size_t)(m > 0 ? m : 1));
        double *HY = (double *)malloc(sizeof(double) * (size_t)D * (size_t)(m > 0 ? m : 1));
        double *A = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m * m : 1));
        double *S = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m * m : 1));
        double *th = (double *)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
        if (!Y || !HY || !A || !S || !th) {
            free(Y), free(HY), free(A), free(S), free(th);
            rc = ORC_ENOMEM;
            goto out;
        }
        memcpy(Y, F, sizeof(double) * (size_t)D * (size_t)nf);
        int mm = nf;
        for (int q = 0; last_n > 0 && q < nlast; ++q) {
            double *y = Y + (size_t)D * mm;
            const double *zc = z + (size_t)last_n * last[q];
#pragma omp parallel for schedule(static)
            for (int64_t i = 0; i < D; ++i) {
                double sacc = 0.0;
                for (int t = 0; t < last_n; ++t) sacc += V[(size_t)D * t + i] * zc[t];
                y[i] = sacc;
            }
            const double yn = orc_cgs2(D, Y, mm, NULL, 0, y, h);
            if (yn <= 1e-8) continue;
            for (int64_t i = 0; i < D; ++i) y[i] /= yn;
            ++mm;
        }
        for (int c = 0; c < mm; ++c) orc_spmv(D, row_ptr, col, val, diag, J, U, Y + (size_t)D * c, HY + (size_t)D * c);
        for (int a = 0; a < mm; ++a)
            for (int c = a; c < mm; ++c) A[a * mm + c] = A[c * mm + a] = dot(D, Y + (size_t)D * a, HY + (size_t)D * c);
        orc_jacobi(mm, A, th, S);
        for (int q = 0; q < k; ++q) {
            evals[q] = NAN;
            if (resid) resid[q] = NAN;
        }
        for (int q = 0; q < k && q < mm; ++q) {
            evals[q] = th[q];
            if (resid) {
#pragma omp parallel for schedule(static)
                for (int64_t i = 0; i < D; ++i) {
                    double sacc = 0.0;
                    for (int c = 0; c < mm; ++c) sacc += (HY[(size_t)D * c + i] - th[q] * Y[(size_t)D * c + i]) * S[c * mm + q];
                    w[i] = sacc;
                }
                resid[q] = sqrt(dot(D, w, w));
            }
        }
        if (mm < k && rc == ORC_OK && nf + last_n < D) rc = ORC_ENOCONV;
        free(Y), free(HY), free(A), free(S), free(th);
    }
    *iters = total;
out:
    free(F);
    free(V);
    free(w);
    free(alpha);
    free(beta);
    free(h);
    free(td);
    free(fth);
    free(z);
    free(mv);
    free(lev);
    free(last);
    return rc;
}

/* ---- spectral statistic --------------------------------------------------------------------- */

/* r_i = min(d_i, d_{i+1}) / max(d_i, d_{i+1}) over ascending-sorted ev, 0 when max == 0;
 * returns the mean over the nb-2 ratios (src/analysis.cpp:433-454, mean at :317-318). */
double orc_gap_ratio(int nb, const double *ev_in, double *ratios) {
    if (nb < 3) return 0.0;
    double *ev = (double *)malloc(sizeof(double) * (size_t)nb);
    memcpy(ev, ev_in, sizeof(double) * (size_t)nb);
    for (int i = 1; i < nb; ++i) {
        double x = ev[i];
        int j = i - 1;
        while (j >= 0 && ev[j] > x) {
            ev[j + 1] = ev[j];
            --j;
        }
        ev[j + 1] = x;
    }
    double sum = 0.0;
    for (int i = 1; i < nb - 1; ++i) {
        double a = ev[i] - ev[i - 1], b = ev[i + 1] - ev[i];
        double mn = a < b ? a : b, mx = a < b ? b : a;
        double r = (mx != 0.0) ? mn / mx : 0.0;
        if (ratios) ratios[i - 1] = r;
        sum += r;
    }
    free(ev);
    return sum / (double)(nb - 2);
}

/*
 * Single-particle density matrix rho[i*m + j] = <phi| b_i^+ b_j |phi> of a real state phi given in the order
 * of `states` (D states of m sites).  Returns ORC_OK, or ORC_ENOMEM / ORC_ERANGE (a target not in the basis).
 */
int orc_spdm(int m, int64_t D, const uint8_t *states, const double *phi, double *rho) {
    if (m < 1 || m > 32 || D < 1) return ORC_EINVAL;
    keyed_t *keys = make_keys(states, m, D);
    if (!keys) return ORC_ENOMEM;
    for (int a = 0; a < m * m; ++a) rho[a] = 0.0;
    uint8_t t[64];
    for (int64_t k = 0; k < D; ++k) {
        const uint8_t *s = states + (size_t)k * m;
        const double pk = phi[k];
        for (int i = 0; i < m; ++i) {
            rho[i * m + i] += pk * pk * (double)s[i];
            for (int j = 0; j < m; ++j) {
                if (j == i || s[j] < 1) continue;
                memcpy(t, s, (size_t)m);
                t[i] += 1;
                t[j] -= 1;
                const int64_t c = find_key(keys, D, state_key(t, m));
                if (c < 0) {
                    free(keys);
                    return ORC_ERANGE;
                }
                rho[i * m + j] += pk * phi[c] * sqrt((double)(s[i] + 1) * (double)s[j]);
            }
        }
    }
    free(keys);
    return ORC_OK;
}

/* ---- mean-field single-site problem ---------------------------------------------------------- */

/*
 * h_MF on the 2p+1 occupations 0..2p (src/hamiltonian.cpp:290-309), U = 1:
 *   diagonal  d_i = -mu i + i(i-1)/2 + q J psi^2
 *   coupling  e_i = h(i, i+1) = -q J psi sqrt(i+1),  i = 0 .. 2p-1
 * sq[i] = sqrt(i) (correctly rounded); every expression keeps one operand order, and the file is built
 * without floating-point contraction, so each value is one fixed sequence of IEEE operations.
 */
static double mf_diag(int i, double mu, double w) { return -mu * (double)i + 0.5 * (double)i * (double)(i - 1) + w; }

/*
 * LDL^T pivots u_i of T - x I and their x-derivatives; returns 1 when every pivot is positive (x below the
 * lowest eigenvalue) with *S = sum_i u_i'/u_i = d/dx ln det(T - x I) = -sum_k 1/(lambda_k - x), else 0.
 * The squared couplings are taken as e_{i-1}^2 = c^2 i (c2 = c^2).
 *   u_0 = d_0 - x, u_i = (d_i - x) - e_{i-1}^2 / u_{i-1};  u_0' = -1, u_i' = -1 + e_{i-1}^2 u_{i-1}' / u_{i-1}^2
 */
static int mf_pivots(int N, double mu, double w, double c2, double x, double *S) {
    double u = mf_diag(0, mu, w) - x, du = -1.0;
    if (!(u > 0.0)) return 0;
    double r = 1.0 / u;
    double s = du * r;
    for (int i = 1; i < N; ++i) {
        const double ee = c2 * (double)i;
        du = -1.0 + ee * du * r * r;
        u = (mf_diag(i, mu, w) - x) - ee * r;
        if (!(u > 0.0)) return 0;
        r = 1.0 / u;
        s = s + du * r;
    }
    *S = s;
    return 1;
}

/*
 * Logarithmic derivatives of det(T - x I) in x and in the coupling scale c (diagonal held fixed), from the
 * pivots and their derivatives; 1 when every pivot is positive, else 0.  d(e_{i-1}^2)/dc = 2 c i (tc = 2c).
 *   u_i'(c) = -(2 c i) / u_{i-1} + e_{i-1}^2 u_{i-1}'(c) / u_{i-1}^2,  u_0'(c) = 0
 */
static int mf_logdet_deriv(int N, double mu, double w, double c2, double tc, double x, double *Sx, double *Sc) {
    double u = mf_diag(0, mu, w) - x, dux = -1.0, duc = 0.0;
    if (!(u > 0.0)) return 0;
    double r = 1.0 / u;
    double sx = dux * r, sc = 0.0;
    for (int i = 1; i < N; ++i) {
        const double ee = c2 * (double)i;
        const double dee = tc * (double)i;
        dux = -1.0 + ee * dux * r * r;
        duc = -(dee * r) + ee * duc * r * r;
        u = (mf_diag(i, mu, w) - x) - ee * r;
        if (!(u > 0.0)) return 0;
        r = 1.0 / u;
        sx = sx + dux * r;
        sc = sc + duc * r;
    }
    *Sx = sx;
    *Sc = sc;
    return 1;
}

/*
 * Ground state of h_MF(psi, p): returns its energy and *sf = <phi|a|phi> (what SF_density computes from the
 * eigenvector, src/analysis.cpp:95-105).
 *   energy: Newton's iteration on ln det(T - x I) from below (x <- x - 1/S).  Every root is real, so from
 *           below the lowest one each exact step stays below it and the iterates rise monotonically to it
 *           (quadratically once close).  The start is `hint` when its pivots are all positive (the caller
 *           passes the previous p's eigenvalue less 1e-3 (|lambda| + 1): more occupations only lower it), else
 *           the Gershgorin lower bound; hi = min_i d_i, or the rejected hint, lies at or above the root.  An
 *           iterate whose pivots are not all positive, or a step reaching hi, can only come from rounding at
 *           the root: the iterate (resp. hi) is returned; so is the last iterate when a step no longer moves x.
 *   <a>   : Hellmann-Feynman.  The couplings are c sqrt(i+1), so d lambda/dc = 2 sum_i sqrt(i+1) phi_i phi_{i+1}
 *           = 2 <a>, and d lambda/dc = -(d ln det/dc) / (d ln det/dx) as x -> lambda.  The ratio f(x) is
 *           evaluated at lambda - delta and lambda - 2 delta, delta = 1e-9 (|lambda| + 1), and extrapolated
 *           linearly to x = lambda (2 f(lambda - delta) - f(lambda - 2 delta)), which leaves an O(delta^2)
 *           error; no eigenvector is formed.
 *   psi = 0 or J = 0 (diagonal h): the first smallest diagonal entry and <a> = 0.
 */
double orc_mf_ground(double psi, int p, double mu, double J, int q, const double *sq, double hint, double *sf) {
    const int N = 2 * p + 1;
    const double c = (double)(-q) * J * psi;
    const double w = (double)q * J * psi * psi;
    double lo = 0.0, hi = 0.0;
    for (int i = 0; i < N; ++i) {
        const double d = mf_diag(i, mu, w);
        const double a0 = i > 0 ? fabs(c * sq[i]) : 0.0;
        const double a1 = i < N - 1 ? fabs(c * sq[i + 1]) : 0.0;
        const double g = d - a0 - a1;
        if (i == 0 || g < lo) lo = g;
        if (i == 0 || d < hi) hi = d;
    }
    *sf = 0.0;
    if (c == 0.0) return hi;
    const double c2 = c * c, tc = 2.0 * c;
    double S = 0.0;
    /* the start: the hint when its pivots are all positive, else the Gershgorin bound (moved down if rounding
       put it at the eigenvalue) */
    double xk = lo;
    int have_start = 0;
    if (hint > lo && hint < hi) {
        if (mf_pivots(N, mu, w, c2, hint, &S)) {
            xk = hint;
            have_start = 1;
        } else {
            hi = hint;
        }
    }
    if (!have_start) {
        for (int t = 0; t < 64 && !mf_pivots(N, mu, w, c2, lo, &S); ++t) lo = lo - (fabs(lo) + 1.0);
        xk = lo;
    }
    double lam = xk;
    for (int it = 0; it < 200; ++it) {
        if (!mf_pivots(N, mu, w, c2, xk, &S)) break; /* a Newton iterate at the root within rounding */
        const double xn = xk - 1.0 / S;
        if (!(xn < hi)) { /* the root lies at hi within rounding */
            lam = hi;
            break;
        }
        if (!(xn > xk)) break; /* no progress */
        xk = xn;
        lam = xk;
    }
    double delta = 1e-9 * (fabs(lam) + 1.0);
    for (int t = 0; t < 8; ++t) {
        double sx1, sc1, sx2, sc2;
        if (mf_logdet_deriv(N, mu, w, c2, tc, lam - delta, &sx1, &sc1) &&
            mf_logdet_deriv(N, mu, w, c2, tc, lam - 2.0 * delta, &sx2, &sc2)) {
            const double f1 = -0.5 * sc1 / sx1;
            const double f2 = -0.5 * sc2 / sx2;
            *sf = 2.0 * f1 - f2;
            break;
        }
        delta = 2.0 * delta;
    }
    return lam;
}

/* the start hint for p + 1 after the eigenvalue e of p */
static double mf_hint(double e) { return e - 1e-3 * (fabs(e) + 1.0); }

/*
 * Analysis::SCMF (src/analysis.cpp:111-178) at one (mu, J) point, U = 1:
 *   repeat: p = 1, E0 = ground(psi, 1); repeat p += 1, E = ground(psi, p) until |E0 - E| <= eps (E0 <- E);
 *           psi' = <a> of the last ground state; stop when |psi' - psi| <= eps (psi <- psi')
 * Outputs psi, the outer iterations run and the last p.  Returns 0 (converged), 1 (max_outer outer iterations
 * without convergence: psi is the last iterate), 2 (p would exceed pmax), -ORC_EINVAL or -ORC_ENOMEM.
 */
int orc_scmf(double mu, double J, int q, double psi0, double eps, int max_outer, int pmax, double *psi_out,
             int *iters_out, int *p_out) {
    if (pmax < 2 || pmax > 4096 || max_outer < 1) return -ORC_EINVAL;
    const int NMAX = 2 * pmax + 1;
    double *sq = (double *)malloc(sizeof(double) * (size_t)(NMAX + 1));
    if (!sq) return -ORC_ENOMEM;
    for (int i = 0; i <= NMAX; ++i) sq[i] = sqrt((double)i);
    double psi = psi0, sf = 0.0;
    int status = 1, outer = 0, p = 1;
    while (outer < max_outer) {
        ++outer;
        p = 1;
        double e0 = orc_mf_ground(psi, p, mu, J, q, sq, -INFINITY, &sf);
        double tmp;
        do {
            if (p + 1 > pmax) {
                status = 2;
                goto done;
            }
            ++p;
            const double e = orc_mf_ground(psi, p, mu, J, q, sq, mf_hint(e0), &sf);
            tmp = fabs(e0 - e);
            e0 = e;
        } while (tmp > eps);
        tmp = fabs(sf - psi);
        psi = sf;
        if (tmp <= eps) {
            status = 0;
            break;
        }
    }
done:
    *psi_out = psi;
    if (iters_out) *iters_out = outer;
    if (p_out) *p_out = p;
    free(sq);
    return status;
}
