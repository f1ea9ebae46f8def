// bh_lanczos.hip -- Lanczos eigensolver for B independent H_b = J_b T + U_b V sharing one CSR.
//
// Replaces Op::IRLM_eigen (src/operator.cpp:22-33), which runs Spectra's implicitly restarted
// Arnoldi (GenEigsBase.h:386-441; nev, ncv = 2 nev + 1, SmallestReal, tol 1e-10) with a serial
// sparse product per step.  Here:
//   step j (device, no host round trip):
//     K1  w = H v_j - beta_j v_{j-1}  (CSR pass fused with the diagonal and the recurrence),
//         alpha_j = <v_j, w>           (FP64 partials, deterministic two-level grid reduction)
//     KS  c_l = <y_l, w - alpha_j v_j> for the locked Ritz vectors y_l (only when some are locked)
//     K2  w'' = w - alpha_j v_j - sum_l c_l y_l, stored as v_{j+1} (unnormalised, storage precision),
//         beta_{j+1} = ||w''|| (FP64), scale_{j+1} = 1 / beta_{j+1}
//   every `chunk` steps (host): FP64 analysis of T_j (Sturm bisection + inverse iteration), the
//     Spectra convergence rule |beta_{j} s_last| < tol max(eps^(2/3), |theta|), and selective
//     orthogonalisation: a Ritz pair whose bound drops below sqrt(eps_store) ||T|| is locked, its Ritz
//     vector computed on the device, and later Lanczos vectors are kept orthogonal to it
//     (Parlett & Scott's selective orthogonalisation).
//   finish: Ritz vectors in FP64, Rayleigh quotients and residuals with the exact FP64 values of T.
// Start vector: Spectra's convention, Park-Miller LCG r (SimpleRandom.h:30-96), v0 = H r / ||H r||
// (Arnoldi.h:147-154).
// Mixed precision: Krylov vectors and locked vectors stored FP32, all reductions and the W buffer
// in FP64; the matrix values come exactly from the FP64 table of the packed format.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

#include "bh_kernels.hpp"
#include "bh_spmv.hpp"
#include "bh_tridiag.hpp"

constexpr int LZ_WG = 512;      // threads per workgroup of the Lanczos kernels (8 waves)
constexpr int LZ_GRID_MAX = 1024;
constexpr int LZ_GROUP = 32;    // workgroups per first-level reduction group
constexpr int LZ_LMAX = 16;     // locked Ritz vectors per solve
constexpr int LZ_NVMAX = 16;    // Ritz vectors formed per pass

// ---- deterministic grid reduction -------------------------------------------------------------------
// Each workgroup has written part[wg * nv + v] (v < nv).  The last workgroup of each group of
// LZ_GROUP sums its group into gpart[group * nv + v]; the last group sums the groups into out[v].
// Returns true in the single workgroup that wrote out[].  Summation order is fixed.
__device__ bool lz_grid_reduce(const double* part, double* gpart, int nv, unsigned* counters, double* out) {
    const int ngroups = (gridDim.x + LZ_GROUP - 1) / LZ_GROUP;
    const int grp = blockIdx.x / LZ_GROUP;
    unsigned* gc = counters + grp;
    unsigned* top = counters + LZ_GRID_MAX / LZ_GROUP;
    __shared__ int flag;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const int members = min(LZ_GROUP, static_cast<int>(gridDim.x) - grp * LZ_GROUP);
        const unsigned prev = __hip_atomic_fetch_add(gc, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        flag = (prev == static_cast<unsigned>(members - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (!flag) return false;
    __threadfence();
    const int w0 = grp * LZ_GROUP;
    const int w1 = min(w0 + LZ_GROUP, static_cast<int>(gridDim.x));
    for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        double s = 0.0;
        for (int w = w0; w < w1; ++w) s += bh_coherent_load(part + static_cast<size_t>(w) * nv + v);
        gpart[static_cast<size_t>(grp) * nv + v] = s;
    }
    if (threadIdx.x == 0) bh_reset_counter(gc);
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = __hip_atomic_fetch_add(top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        flag = (prev == static_cast<unsigned>(ngroups - 1)) ? 1 : 0;
    }
    __syncthreads();
    if (!flag) return false;
    __threadfence();
    for (int v = threadIdx.x; v < nv; v += blockDim.x) {
        double s = 0.0;
        for (int g = 0; g < ngroups; ++g) s += bh_coherent_load(gpart + static_cast<size_t>(g) * nv + v);
        out[v] = s;
    }
    if (threadIdx.x == 0) bh_reset_counter(top);
    __syncthreads();
    return true;
}

// per-workgroup sum of one value per vector b (lane b = lane % B), written to part[wg * nv + q*B + b]
template <int B>
__device__ __forceinline__ void lz_block_put(double v, int q, double* lds, double* part, int nv) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int o = B; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
    if (lane < B) lds[wid * B + lane] = v;
    __syncthreads();
    if (threadIdx.x < B) {
        double s = 0.0;
        for (int w = 0; w < LZ_WG / 64; ++w) s += lds[w * B + threadIdx.x];
        part[static_cast<size_t>(blockIdx.x) * nv + q * B + threadIdx.x] = s;
    }
    __syncthreads();
}

struct LzState {
    double* alpha;   // [maxit+1][B]
    double* beta;    // [maxit+2][B]
    double* scale;   // [maxit+2][B]  1 / ||stored vector||
    double* J;       // [B]
    double* U;       // [B]
    double* cl;      // [LZ_LMAX][B]
    double* red;     // [(1 + LZ_LMAX) * B] reduction results
    int* active;     // [B]
    int* nlock;      // [B]
    int* step;       // [1]
    unsigned* cnt;   // [LZ_GRID_MAX / LZ_GROUP + 1]
    double* part;    // [LZ_GRID_MAX][(1 + LZ_LMAX) * B]
    double* gpart;   // [LZ_GRID_MAX / LZ_GROUP][(1 + LZ_LMAX) * B]
};

template <typename XT>
struct LzArgs {
    int64_t D;
    const int32_t* row_ptr;
    const uint32_t* words;
    const int32_t* col64;
    const double* val64;
    const float* diag32;
    const double* diag64;
    const double* table;
    int cb;
    XT* V;         // [(maxit+1)][D][B]
    int64_t vs;    // D * B
    double* W;     // [D][B]
    XT* Y;         // [LZ_LMAX][D][B] locked Ritz vectors (unit norm)
    int64_t ntile;
    int64_t chunk; // tiles per workgroup
    int maxit;
    LzState st;
};

// K1: w = H v_j - beta_j v_{j-1}; alpha partials.  init: w = H r for the start vector r (double [D]).
template <typename XT, int B, int G, int FMT, bool INIT>
__global__ void __launch_bounds__(LZ_WG) k_lz_a(LzArgs<XT> a, const double* __restrict__ r0) {
    constexpr int GB = G * B;
    constexpr int SLOTS = LZ_WG / GB;
    extern __shared__ double tab[];
    __shared__ double lds[LZ_WG / 64 * B];
    if constexpr (FMT == BH_FMT_PACKED) {
        for (int i = threadIdx.x; i < (1 << a.cb); i += LZ_WG) tab[i] = a.table[i];
        __syncthreads();
    }
    const int lane = threadIdx.x;
    const int slot = lane / GB;
    const int g = (lane % GB) / B;
    const int b = lane % B;
    const int j = INIT ? 0 : *a.st.step;
    const XT* vj = a.V + static_cast<int64_t>(j) * a.vs;
    const XT* vjm = a.V + static_cast<int64_t>(j > 0 ? j - 1 : 0) * a.vs;
    const double sj = INIT ? 1.0 : a.st.scale[j * B + b];
    const double bj = (!INIT && j > 0) ? a.st.beta[j * B + b] * a.st.scale[(j - 1) * B + b] : 0.0;
    const bool act = a.st.active[b] != 0;
    const double Jb = a.st.J[b], Ub = a.st.U[b];
    const int64_t L = bh_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t0 = L * a.chunk, t1 = min(t0 + a.chunk, a.ntile);
    double pa = 0.0;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t r = t * SLOTS + slot;
        double acc = 0.0;
        if (r < a.D && act) {
            if constexpr (INIT) {
                // start vector shared by all b: x[c] = r0[c]
                if constexpr (FMT == BH_FMT_PACKED) {
                    const uint32_t mask = (1u << a.cb) - 1u;
                    for (int32_t k = a.row_ptr[r] + g; k < a.row_ptr[r + 1]; k += G) {
                        const uint32_t w = a.words[k];
                        acc += tab[w & mask] * r0[w >> a.cb];
                    }
                } else {
                    for (int32_t k = a.row_ptr[r] + g; k < a.row_ptr[r + 1]; k += G) acc += a.val64[k] * r0[a.col64[k]];
                }
            } else {
                acc = bh_row_partial<XT, B, G, FMT>(a.row_ptr[r], a.row_ptr[r + 1], g, b, a.words, a.col64, a.val64, tab,
                                                     a.cb, vj);
            }
        }
        acc = bh_group_sum<B, G>(acc);
        if (g == 0 && r < a.D && act) {
            const double d = a.diag32 ? static_cast<double>(a.diag32[r]) : a.diag64[r];
            const double hx = (FMT == BH_FMT_PACKED) ? -acc : acc;
            if constexpr (INIT) {
                a.W[r * B + b] = Jb * hx + Ub * d * r0[r];
            } else {
                const double xr = to_d(vj[r * B + b]) * sj;
                double w = Jb * hx * sj + Ub * d * xr;
                if (j > 0) w -= bj * to_d(vjm[r * B + b]);
                a.W[r * B + b] = w;
                pa += xr * w;
            }
        }
    }
    if constexpr (!INIT) {
        const int nv = B;
        lz_block_put<B>(pa, 0, lds, a.st.part, nv);
        if (lz_grid_reduce(a.st.part, a.st.gpart, nv, a.st.cnt, a.st.red)) {
            if (threadIdx.x < B && a.st.active[threadIdx.x]) a.st.alpha[j * B + threadIdx.x] = a.st.red[threadIdx.x];
        }
    }
}

// KS: c_l = <y_l, w - alpha_j v_j> for l < nlock[b]
template <typename XT, int B>
__global__ void __launch_bounds__(LZ_WG) k_lz_so(LzArgs<XT> a, int lmax_used) {
    __shared__ double lds[LZ_WG / 64 * B];
    const int b = threadIdx.x % B;
    const int j = *a.st.step;
    const XT* vj = a.V + static_cast<int64_t>(j) * a.vs;
    const double sj = a.st.scale[j * B + b];
    const double al = a.st.alpha[j * B + b];
    const int nl = a.st.active[b] ? a.st.nlock[b] : 0;
    double acc[LZ_LMAX];
#pragma unroll
    for (int l = 0; l < LZ_LMAX; ++l) acc[l] = 0.0;
    constexpr int ROWS = LZ_WG / B;
    const int64_t L = bh_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t0 = L * a.chunk, t1 = min(t0 + a.chunk, a.ntile * (LZ_WG / B) / ROWS);
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t r = t * ROWS + threadIdx.x / B;
        if (r < a.D && nl > 0) {
            const int64_t idx = r * B + b;
            const double w = a.W[idx] - al * to_d(vj[idx]) * sj;
#pragma unroll
            for (int l = 0; l < LZ_LMAX; ++l)
                if (l < nl) acc[l] += to_d(a.Y[static_cast<int64_t>(l) * a.vs + idx]) * w;
        }
    }
    const int nv = lmax_used * B;
    for (int l = 0; l < lmax_used; ++l) lz_block_put<B>(acc[l], l, lds, a.st.part, nv);
    if (lz_grid_reduce(a.st.part, a.st.gpart, nv, a.st.cnt, a.st.red)) {
        for (int i = threadIdx.x; i < nv; i += LZ_WG) a.st.cl[i] = a.st.red[i];
    }
}

// K2: v_{j+1} = w - alpha v_j - sum_l c_l y_l (stored), beta_{j+1} = ||v_{j+1}||.  INIT stores V[0].
template <typename XT, int B, bool INIT>
__global__ void __launch_bounds__(LZ_WG) k_lz_b(LzArgs<XT> a, int use_so) {
    __shared__ double lds[LZ_WG / 64 * B];
    const int b = threadIdx.x % B;
    const int j = INIT ? -1 : *a.st.step;
    const XT* vj = a.V + static_cast<int64_t>(j < 0 ? 0 : j) * a.vs;
    XT* vn = a.V + static_cast<int64_t>(j + 1) * a.vs;
    const bool act = a.st.active[b] != 0;
    const double sj = INIT ? 0.0 : a.st.scale[j * B + b];
    const double al = INIT ? 0.0 : a.st.alpha[j * B + b];
    const int nl = (use_so && act) ? a.st.nlock[b] : 0;
    double c[LZ_LMAX];
#pragma unroll
    for (int l = 0; l < LZ_LMAX; ++l) c[l] = (l < nl) ? a.st.cl[l * B + b] : 0.0;
    constexpr int ROWS = LZ_WG / B;
    const int64_t ntile_b = (a.D + ROWS - 1) / ROWS;
    const int64_t per = (ntile_b + gridDim.x - 1) / gridDim.x;
    const int64_t L = bh_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t0 = L * per, t1 = min(t0 + per, ntile_b);
    double pn = 0.0;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t r = t * ROWS + threadIdx.x / B;
        if (r < a.D && act) {
            const int64_t idx = r * B + b;
            double w = a.W[idx];
            if (!INIT) {
                w -= al * to_d(vj[idx]) * sj;
#pragma unroll
                for (int l = 0; l < LZ_LMAX; ++l)
                    if (l < nl) w -= c[l] * to_d(a.Y[static_cast<int64_t>(l) * a.vs + idx]);
            }
            const XT wr = static_cast<XT>(w);
            vn[idx] = wr;
            const double wd = to_d(wr);
            pn += wd * wd;
        }
    }
    lz_block_put<B>(pn, 0, lds, a.st.part, B);
    if (lz_grid_reduce(a.st.part, a.st.gpart, B, a.st.cnt, a.st.red)) {
        if (threadIdx.x < B) {
            const int bb = threadIdx.x;
            if (a.st.active[bb]) {
                const double nrm = std::sqrt(a.st.red[bb]);
                if (INIT) {
                    a.st.beta[bb] = 0.0;
                    a.st.scale[bb] = nrm > 0.0 ? 1.0 / nrm : 0.0;
                } else {
                    a.st.beta[(j + 1) * B + bb] = nrm;
                    a.st.scale[(j + 1) * B + bb] = nrm > 0.0 ? 1.0 / nrm : 0.0;
                }
            }
        }
        if (threadIdx.x == 0) *a.st.step = j + 1;
    }
}

// Ritz vectors: out[i][r][b] = sum_{t < nj[b]} V[t][r][b] * C[(t * nv + i) * B + b]   (C includes scale_t).
// Each run stops at its own step count: Krylov slots past it were never written for that run.
template <typename XT, typename OT, int B>
__global__ void __launch_bounds__(256) k_lz_ritz(const XT* __restrict__ V, int64_t vs, const int* __restrict__ njb,
                                                 const double* __restrict__ C, int nv, OT* __restrict__ out, int64_t os) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    const int nj = njb[b];
    double acc[LZ_NVMAX];
#pragma unroll
    for (int i = 0; i < LZ_NVMAX; ++i) acc[i] = 0.0;
    for (int t = 0; t < nj; ++t) {
        const double v = to_d(V[static_cast<int64_t>(t) * vs + idx]);
        const double* ct = C + (static_cast<int64_t>(t) * nv) * B + b;
#pragma unroll
        for (int i = 0; i < LZ_NVMAX; ++i)
            if (i < nv) acc[i] += v * ct[i * B];
    }
#pragma unroll
    for (int i = 0; i < LZ_NVMAX; ++i)
        if (i < nv) out[static_cast<int64_t>(i) * os + idx] = static_cast<OT>(acc[i]);
}

// per-(i, b) reductions over rows of [nv][D][B] arrays:
//   mode 0: out[(2i)B+b] = <y,y>, out[(2i+1)B+b] = <y,Hy>;  mode 1: out[iB+b] = ||Hy - theta y||^2;
//   mode 2: out[iB+b] = <y,y> (y in storage precision)
template <typename YT, int B>
__global__ void __launch_bounds__(LZ_WG) k_lz_dots(const YT* __restrict__ Y, const double* __restrict__ HY, int64_t D, int nv,
                                                   int mode, const double* __restrict__ theta, double* part, double* gpart,
                                                   unsigned* cnt, double* out) {
    __shared__ double lds[LZ_WG / 64 * B];
    const int b = threadIdx.x % B;
    constexpr int ROWS = LZ_WG / B;
    const int64_t vs = D * B;
    const int64_t ntile = (D + ROWS - 1) / ROWS;
    const int64_t per = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t t0 = blockIdx.x * per, t1 = min(t0 + per, ntile);
    const int nq = mode == 0 ? 2 * nv : nv;
    for (int i = 0; i < nv; ++i) {
        double s0 = 0.0, s1 = 0.0;
        const double th = mode == 1 ? theta[i * B + b] : 0.0;
        for (int64_t t = t0; t < t1; ++t) {
            const int64_t r = t * ROWS + threadIdx.x / B;
            if (r < D) {
                const int64_t idx = static_cast<int64_t>(i) * vs + r * B + b;
                const double y = to_d(Y[idx]);
                if (mode == 0) {
                    s0 += y * y;
                    s1 += y * HY[idx];
                } else if (mode == 1) {
                    const double rr = HY[idx] - th * y;
                    s0 += rr * rr;
                } else {
                    s0 += y * y;
                }
            }
        }
        if (mode == 0) {
            lz_block_put<B>(s0, 2 * i, lds, part, nq * B);
            lz_block_put<B>(s1, 2 * i + 1, lds, part, nq * B);
        } else {
            lz_block_put<B>(s0, i, lds, part, nq * B);
        }
    }
    (void)lz_grid_reduce(part, gpart, nq * B, cnt, out);
}

// y *= f[i*B + b] over [nv][D][B]
template <typename YT, int B>
__global__ void k_lz_scale(YT* __restrict__ Y, int64_t vs, int nv, const double* __restrict__ f) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs * nv) return;
    const int64_t i = idx / vs;
    const int b = static_cast<int>((idx % vs) % B);
    Y[idx] = static_cast<YT>(to_d(Y[idx]) * f[i * B + b]);
}

// ---- host driver --------------------------------------------------------------------------------------

namespace {

struct RunHost {
    std::vector<double> alpha, beta;     // alpha[t], beta[t] (beta[0] = 0)
    std::vector<double> lock_theta;      // locked Ritz values
    bool done = false;
    int iters = 0;
    int status = BH_ENOCONV;
    std::vector<double> theta;           // wanted levels (Ritz values of T)
    std::vector<std::vector<double>> s;  // their eigenvectors of T
    double tnorm = 0.0;
    std::vector<double> all_theta;       // lowest Ritz values of the last analysis (for gaps)
};

struct Analysis {
    std::vector<double> theta, bound;
    std::vector<std::vector<double>> s;
    double tnorm = 0.0;
};

// lowest p Ritz pairs of T_j with bounds |beta_j s_last|
void analyze(const RunHost& rh, int j, double beta_next, int p, Analysis& an) {
    const double* a = rh.alpha.data();
    const double* b = rh.beta.data();  // b[t] couples t-1, t; b[0] unused
    double lo, hi;
    bhtri::gershgorin(j, a, b, &lo, &hi);
    an.tnorm = std::fmax(std::fabs(lo), std::fabs(hi));
    p = std::min(p, j);
    an.theta.assign(p, 0.0);
    an.bound.assign(p, 0.0);
    an.s.assign(p, std::vector<double>(j));
    for (int i = 0; i < p; ++i) {
        an.theta[i] = bhtri::bisect(j, a, b, i, lo, hi);
        bhtri::inverse_iteration(j, a, b, an.theta[i], an.tnorm, an.s[i].data());
        an.bound[i] = std::fabs(beta_next * an.s[i][j - 1]);
    }
}

template <typename T>
int dalloc(bh_ctx* ctx, DevArray<T>& a, size_t n) {
    return a.alloc(ctx, n);
}

}  // namespace

template <typename XT, int B>
static int lanczos_impl(bh_ctx* ctx, bh_ham* h, const double* Jh, const double* Uh, int nreal,
                        const bh_lanczos_opts& o, bh_lanczos_result* res, std::vector<double>* vecs_out) {
    constexpr bool MIXED = std::is_same<XT, float>::value;
    constexpr int FMT = MIXED ? BH_FMT_PACKED : BH_FMT_CSR64;
    hipStream_t s = ctx->stream;
    const int64_t D = h->D;
    const int k = o.k;
    int maxit = o.max_iter > 0 ? o.max_iter : 1000;
    if (maxit > D) maxit = static_cast<int>(D);
    const int chunk = o.chunk > 0 ? o.chunk : 16;
    const double tol = o.tol > 0 ? o.tol : 1e-10;
    const double dedup = o.dedup_rel > 0 ? o.dedup_rel : 1e-9;
    const double eps_store = MIXED ? 5.96e-8 : 1.11e-16;
    const double eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
    const int64_t vs = D * B;

    // ---- device buffers ----
    DevArray<XT> V, Y;
    DevArray<double> W, st_d, part, gpart, r0;
    DevArray<int> st_i;
    DevArray<unsigned> cnt;
    BH_TRY(dalloc(ctx, V, static_cast<size_t>(maxit + 1) * vs));
    BH_TRY(dalloc(ctx, W, static_cast<size_t>(vs)));
    BH_TRY(dalloc(ctx, Y, static_cast<size_t>(LZ_LMAX) * vs));
    const size_t n_alpha = static_cast<size_t>(maxit + 1) * B, n_beta = static_cast<size_t>(maxit + 2) * B;
    const size_t nd = n_alpha + 2 * n_beta + 2 * B + LZ_LMAX * B + (1 + LZ_LMAX) * B * 2;
    BH_TRY(dalloc(ctx, st_d, nd));
    BH_TRY(dalloc(ctx, st_i, static_cast<size_t>(2 * B + 1)));
    const int nvmax = std::max((1 + LZ_LMAX) * B, 2 * LZ_NVMAX * B);
    BH_TRY(dalloc(ctx, part, static_cast<size_t>(LZ_GRID_MAX) * nvmax));
    BH_TRY(dalloc(ctx, gpart, static_cast<size_t>(LZ_GRID_MAX / LZ_GROUP) * nvmax));
    BH_TRY(dalloc(ctx, cnt, LZ_GRID_MAX / LZ_GROUP + 1));
    BH_TRY(dalloc(ctx, r0, static_cast<size_t>(D)));

    LzState st;
    st.alpha = st_d.p;
    st.beta = st.alpha + n_alpha;
    st.scale = st.beta + n_beta;
    st.J = st.scale + n_beta;
    st.U = st.J + B;
    st.cl = st.U + B;
    st.red = st.cl + LZ_LMAX * B;
    st.active = st_i.p;
    st.nlock = st_i.p + B;
    st.step = st_i.p + 2 * B;
    st.cnt = cnt.p;
    st.part = part.p;
    st.gpart = gpart.p;

    std::vector<double> Jb(B, 0.0), Ub(B, 0.0);
    std::vector<int> act(B, 0), nlock(B, 0);
    for (int b = 0; b < nreal; ++b) {
        Jb[b] = Jh[b];
        Ub[b] = Uh[b];
        act[b] = 1;
    }
    BH_HIP(ctx, hipMemsetAsync(st_d.p, 0, nd * sizeof(double), s));
    BH_HIP(ctx, hipMemsetAsync(cnt.p, 0, cnt.bytes(), s));
    BH_HIP(ctx, hipMemcpyAsync(st.J, Jb.data(), B * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(st.U, Ub.data(), B * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(st.active, act.data(), B * 4, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(st.nlock, nlock.data(), B * 4, hipMemcpyHostToDevice, s));
    {
        std::vector<double> r(static_cast<size_t>(D));
        const uint64_t p = 2147483647ull;
        uint64_t x = o.seed ? (o.seed & p) : 1;
        for (int64_t i = 0; i < D; ++i) {
            x = (16807ull * x) % p;
            r[i] = static_cast<double>(x) / static_cast<double>(p) - 0.5;
        }
        BH_HIP(ctx, hipMemcpyAsync(r0.p, r.data(), D * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipStreamSynchronize(s));  // r is a local
    }

    const int G = std::max(1, std::min(bh_pick_lanes(h, B), 64 / B));
    const int64_t slots = LZ_WG / (G * B);
    const int64_t ntile = (D + slots - 1) / slots;
    const int grid = static_cast<int>(std::min<int64_t>(ntile, LZ_GRID_MAX));
    const int64_t tchunk = (ntile + grid - 1) / grid;
    LzArgs<XT> a{D,
                 h->row_ptr,
                 h->words,
                 h->col64,
                 h->val64,
                 MIXED ? h->diag : nullptr,
                 MIXED ? nullptr : h->diag64,
                 h->table,
                 h->code_bits,
                 V.p,
                 vs,
                 W.p,
                 Y.p,
                 ntile,
                 tchunk,
                 maxit,
                 st};
    // the SO kernel walks rows in tiles of LZ_WG / B
    LzArgs<XT> aso = a;
    {
        const int64_t rows = LZ_WG / B;
        const int64_t nt = (D + rows - 1) / rows;
        aso.ntile = nt;
        aso.chunk = (nt + grid - 1) / grid;
    }
    const size_t shm = FMT == BH_FMT_PACKED ? (static_cast<size_t>(1) << h->code_bits) * sizeof(double) : 0;

    auto launch_a = [&](bool init) {
        KTimer kt(ctx, init ? "lanczos_init" : "lanczos_spmv", s);
#define LA(GG)                                                                                                     \
    if (G == GG) {                                                                                                 \
        if (init) hipLaunchKernelGGL((k_lz_a<XT, B, GG, FMT, true>), dim3(grid), dim3(LZ_WG), shm, s, a, r0.p);   \
        else hipLaunchKernelGGL((k_lz_a<XT, B, GG, FMT, false>), dim3(grid), dim3(LZ_WG), shm, s, a, r0.p);       \
        return;                                                                                                    \
    }
        if constexpr (B == 1) { LA(1) LA(2) LA(4) LA(8) LA(16) LA(32) LA(64) }
        else if constexpr (B == 2) { LA(1) LA(2) LA(4) LA(8) LA(16) LA(32) }
        else if constexpr (B == 4) { LA(1) LA(2) LA(4) LA(8) LA(16) }
        else if constexpr (B == 8) { LA(1) LA(2) LA(4) LA(8) }
        else if constexpr (B == 16) { LA(1) LA(2) LA(4) }
        else if constexpr (B == 32) { LA(1) LA(2) }
        else { LA(1) }
#undef LA
    };
    auto launch_b = [&](bool init, int use_so) {
        KTimer kt(ctx, "lanczos_update", s);
        if (init) hipLaunchKernelGGL((k_lz_b<XT, B, true>), dim3(grid), dim3(LZ_WG), 0, s, a, 0);
        else hipLaunchKernelGGL((k_lz_b<XT, B, false>), dim3(grid), dim3(LZ_WG), 0, s, a, use_so);
    };
    auto launch_so = [&](int lmax_used) {
        KTimer kt(ctx, "lanczos_so", s);
        hipLaunchKernelGGL((k_lz_so<XT, B>), dim3(grid), dim3(LZ_WG), 0, s, aso, lmax_used);
    };

    // start: W = H r, V[0] = W, scale_0 = 1/||W||
    launch_a(true);
    launch_b(true, 0);
    BH_HIP(ctx, hipGetLastError());

    std::vector<RunHost> rh(B);
    std::vector<double> ha(n_alpha), hb(n_beta);
    int j = 0;          // steps done
    int lmax_used = 0;  // max locks over runs
    int next_check = std::max(2 * k + 8, 20);
    int nleft = nreal;
    std::vector<double> Cbuf;
    DevArray<double> dC, dnorm;
    BH_TRY(dalloc(ctx, dnorm, static_cast<size_t>(2 * LZ_NVMAX * B)));

    DevArray<int> dnj;
    BH_TRY(dalloc(ctx, dnj, B));
    auto form_vectors = [&](const std::vector<int>& njb, int nv, const std::vector<double>& C, auto* out,
                            int64_t os) -> int {
        // C: [max nj][nv][B]; njb[b]: Krylov vectors used for run b (0 for runs not formed)
        BH_TRY(dC.alloc(ctx, C.size()));
        BH_HIP(ctx, hipMemcpyAsync(dC.p, C.data(), C.size() * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dnj.p, njb.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
        using OT = std::remove_pointer_t<decltype(out)>;
        KTimer kt(ctx, "lanczos_ritz", s);
        hipLaunchKernelGGL((k_lz_ritz<XT, OT, B>), dim3(static_cast<unsigned>((vs + 255) / 256)), dim3(256), 0, s, V.p, vs,
                           dnj.p, dC.p, nv, out, os);
        BH_HIP(ctx, hipGetLastError());
        BH_HIP(ctx, hipStreamSynchronize(s));
        return BH_OK;
    };

    while (j < maxit && nleft > 0) {
        const int steps = std::min(std::min(next_check, maxit) - j, maxit - j);
        for (int t = 0; t < steps; ++t) {
            launch_a(false);
            if (lmax_used > 0) launch_so(lmax_used);
            launch_b(false, lmax_used > 0 ? 1 : 0);
        }
        BH_HIP(ctx, hipGetLastError());
        j += steps;
        BH_HIP(ctx, hipMemcpyAsync(ha.data(), st.alpha, static_cast<size_t>(j) * B * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hb.data(), st.beta, static_cast<size_t>(j + 1) * B * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        next_check = j + chunk;

        // analysis per active run
        std::vector<std::vector<int>> new_locks(B);
        std::vector<Analysis> ans(B);
        bool any_lock = false, any_done = false;
        for (int b = 0; b < nreal; ++b) {
            RunHost& R = rh[b];
            if (R.done) continue;
            R.alpha.assign(j, 0.0);
            R.beta.assign(j + 1, 0.0);
            for (int t = 0; t < j; ++t) R.alpha[t] = ha[static_cast<size_t>(t) * B + b];
            for (int t = 0; t <= j; ++t) R.beta[t] = hb[static_cast<size_t>(t) * B + b];
            const double bnext = R.beta[j];
            Analysis& an = ans[b];
            const int p = std::min(j, 2 * k + 4);
            analyze(R, j, bnext, p, an);
            const bool exhausted = bnext <= 1e-12 * std::fmax(an.tnorm, 1e-300) || j >= D;
            // FP32-stored vectors: Ritz residuals cannot fall below the storage rounding, ~eps32 ||T||
            const double floor = MIXED ? 4.0 * eps_store * an.tnorm : 0.0;
            // walk distinct levels from the bottom
            std::vector<int> lev;
            for (int i = 0; i < static_cast<int>(an.theta.size()) && static_cast<int>(lev.size()) < k; ++i) {
                if (!lev.empty() && std::fabs(an.theta[i] - an.theta[lev.back()]) <= dedup * an.tnorm) continue;
                lev.push_back(i);
            }
            bool conv = static_cast<int>(lev.size()) == k || exhausted;
            for (int i : lev) {
                const double thr = std::fmax(tol * std::fmax(eps23, std::fabs(an.theta[i])), floor);
                if (an.bound[i] >= thr && !exhausted) conv = false;
            }
            if (conv || j >= maxit) {
                R.done = true;
                R.iters = j;
                R.status = conv && static_cast<int>(lev.size()) == k ? BH_OK : BH_ENOCONV;
                R.theta.clear();
                R.s.clear();
                for (int i : lev) {
                    R.theta.push_back(an.theta[i]);
                    R.s.push_back(an.s[i]);
                }
                R.tnorm = an.tnorm;
                R.all_theta = an.theta;
                act[b] = 0;
                any_done = true;
                --nleft;
                continue;
            }
            // selective orthogonalisation: lock converged wanted-region Ritz vectors
            const double lock_thr = std::sqrt(eps_store) * an.tnorm;
            for (int i : lev) {
                if (an.bound[i] > lock_thr) continue;
                if (nlock[b] + static_cast<int>(new_locks[b].size()) >= LZ_LMAX) break;
                bool have = false;
                for (double th : R.lock_theta)
                    if (std::fabs(th - an.theta[i]) <= dedup * an.tnorm) have = true;
                if (have) continue;
                new_locks[b].push_back(i);
                any_lock = true;
            }
        }
        if (any_lock) {
            int nv = 0;
            for (int b = 0; b < nreal; ++b) nv = std::max(nv, static_cast<int>(new_locks[b].size()));
            nv = std::min(nv, LZ_NVMAX);
            // C[t][i][b] = s_i[t] * scale_t[b]
            std::vector<double> hsc(static_cast<size_t>(j) * B);
            BH_HIP(ctx, hipMemcpyAsync(hsc.data(), st.scale, hsc.size() * 8, hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
            Cbuf.assign(static_cast<size_t>(j) * nv * B, 0.0);
            int base = 0;  // all runs write new vectors at slot nlock[b] + i; form into a staging area
            (void)base;
            // stage new Ritz vectors at Y slots [L0, L0+nv) where L0 = max nlock: runs with fewer locks are
            // compacted afterwards by writing at their own nlock offset (per-run offsets differ), so form
            // one vector index at a time per slot offset.
            for (int i = 0; i < nv; ++i) {
                for (int b = 0; b < nreal; ++b) {
                    if (i >= static_cast<int>(new_locks[b].size())) continue;
                    const std::vector<double>& sv = ans[b].s[new_locks[b][i]];
                    for (int t = 0; t < j; ++t)
                        Cbuf[(static_cast<size_t>(t) * nv + i) * B + b] = sv[t] * hsc[static_cast<size_t>(t) * B + b];
                }
            }
            // form into W-sized scratch slots at the top of Y, then move per run: Y slots are per (l, b);
            // writing vector i of run b to slot nlock[b] + i requires per-b offsets, so form into a
            // temporary [nv][D][B] and copy the columns on the device.
            DevArray<XT> tmp;
            BH_TRY(tmp.alloc(ctx, static_cast<size_t>(nv) * vs));
            std::vector<int> njb(B, 0);
            for (int b = 0; b < nreal; ++b)
                if (!new_locks[b].empty()) njb[b] = j;
            BH_TRY(form_vectors(njb, nv, Cbuf, tmp.p, vs));
            // normalise: ||y||^2 per (i, b)
            hipLaunchKernelGGL((k_lz_dots<XT, B>), dim3(grid), dim3(LZ_WG), 0, s, tmp.p, nullptr, D, nv, 2, nullptr, part.p,
                               gpart.p, cnt.p, dnorm.p);
            std::vector<double> nrm(static_cast<size_t>(nv) * B);
            BH_HIP(ctx, hipMemcpyAsync(nrm.data(), dnorm.p, nrm.size() * 8, hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
            for (auto& v : nrm) v = v > 0 ? 1.0 / std::sqrt(v) : 0.0;
            BH_HIP(ctx, hipMemcpyAsync(dnorm.p, nrm.data(), nrm.size() * 8, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((k_lz_scale<XT, B>), dim3(static_cast<unsigned>((vs * nv + 255) / 256)), dim3(256), 0, s,
                               tmp.p, vs, nv, dnorm.p);
            // scatter vector i of run b into Y slot nlock[b] + i (strided copy of one interleaved column)
            for (int b = 0; b < nreal; ++b) {
                for (int i = 0; i < static_cast<int>(new_locks[b].size()) && i < nv; ++i) {
                    const int slot = nlock[b] + i;
                    BH_HIP(ctx, hipMemcpy2DAsync(Y.p + static_cast<int64_t>(slot) * vs + b, B * sizeof(XT),
                                                 tmp.p + static_cast<int64_t>(i) * vs + b, B * sizeof(XT), sizeof(XT),
                                                 static_cast<size_t>(D), hipMemcpyDeviceToDevice, s));
                    rh[b].lock_theta.push_back(ans[b].theta[new_locks[b][i]]);
                }
                nlock[b] += std::min(static_cast<int>(new_locks[b].size()), nv);
                lmax_used = std::max(lmax_used, nlock[b]);
            }
            BH_HIP(ctx, hipMemcpyAsync(st.nlock, nlock.data(), B * 4, hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
        }
        if (any_done) {
            BH_HIP(ctx, hipMemcpyAsync(st.active, act.data(), B * 4, hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
        }
    }
    for (int b = 0; b < nreal; ++b) {
        if (!rh[b].done) {  // ran out of steps
            RunHost& R = rh[b];
            R.done = true;
            R.iters = j;
            R.status = BH_ENOCONV;
        }
    }

    // ---- FP64 finish: Ritz vectors, Rayleigh quotients and residuals with the exact values ----
    std::vector<double> hsc(static_cast<size_t>(j + 1) * B);
    BH_HIP(ctx, hipMemcpyAsync(hsc.data(), st.scale, hsc.size() * 8, hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipStreamSynchronize(s));
    int nvf = 0;
    for (int b = 0; b < nreal; ++b) nvf = std::max(nvf, static_cast<int>(rh[b].theta.size()));
    nvf = std::min(nvf, LZ_NVMAX);
    std::vector<double> theta_f(static_cast<size_t>(nvf) * B, 0.0), resid_f(static_cast<size_t>(nvf) * B, 0.0);
    if (nvf > 0) {
        int njmax = 0;
        for (int b = 0; b < nreal; ++b)
            if (!rh[b].s.empty()) njmax = std::max(njmax, static_cast<int>(rh[b].s[0].size()));
        Cbuf.assign(static_cast<size_t>(njmax) * nvf * B, 0.0);
        for (int b = 0; b < nreal; ++b)
            for (int i = 0; i < static_cast<int>(rh[b].s.size()) && i < nvf; ++i)
                for (int t = 0; t < static_cast<int>(rh[b].s[i].size()); ++t)
                    Cbuf[(static_cast<size_t>(t) * nvf + i) * B + b] = rh[b].s[i][t] * hsc[static_cast<size_t>(t) * B + b];
        DevArray<double> Yf, HY;
        BH_TRY(Yf.alloc(ctx, static_cast<size_t>(nvf) * vs));
        BH_TRY(HY.alloc(ctx, static_cast<size_t>(nvf) * vs));
        std::vector<int> njb(B, 0);
        for (int b = 0; b < nreal; ++b)
            if (!rh[b].s.empty()) njb[b] = static_cast<int>(rh[b].s[0].size());
        BH_TRY(form_vectors(njb, nvf, Cbuf, Yf.p, vs));
        // exact FP64 operator on each Ritz vector (packed words, FP64 table, FP64 vectors)
        DevArray<double> dJ, dU;
        BH_TRY(dJ.alloc(ctx, B));
        BH_TRY(dU.alloc(ctx, B));
        BH_HIP(ctx, hipMemcpyAsync(dJ.p, Jb.data(), B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dU.p, Ub.data(), B * 8, hipMemcpyHostToDevice, s));
        const int Gf = bh_pick_lanes(h, B);
        for (int i = 0; i < nvf; ++i) {
            KTimer kt(ctx, "lanczos_finish_spmv", s);
            SpmvArgs<double> sa{D, h->row_ptr, h->words, nullptr, nullptr, h->diag, nullptr, h->table, h->code_bits,
                                dJ.p, dU.p, Yf.p + static_cast<int64_t>(i) * vs, HY.p + static_cast<int64_t>(i) * vs, 1};
            bh_launch_spmv(sa, B, Gf, BH_FMT_PACKED, s);
        }
        BH_HIP(ctx, hipGetLastError());
        std::vector<double> dots(static_cast<size_t>(2 * nvf) * B);
        hipLaunchKernelGGL((k_lz_dots<double, B>), dim3(grid), dim3(LZ_WG), 0, s, Yf.p, HY.p, D, nvf, 0, nullptr, part.p,
                           gpart.p, cnt.p, dnorm.p);
        BH_HIP(ctx, hipMemcpyAsync(dots.data(), dnorm.p, dots.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (int i = 0; i < nvf; ++i)
            for (int b = 0; b < B; ++b) {
                const double yy = dots[(2 * i) * B + b], yhy = dots[(2 * i + 1) * B + b];
                theta_f[i * B + b] = yy > 0 ? yhy / yy : 0.0;
            }
        DevArray<double> dth;
        BH_TRY(dth.alloc(ctx, theta_f.size()));
        BH_HIP(ctx, hipMemcpyAsync(dth.p, theta_f.data(), theta_f.size() * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_lz_dots<double, B>), dim3(grid), dim3(LZ_WG), 0, s, Yf.p, HY.p, D, nvf, 1, dth.p, part.p,
                           gpart.p, cnt.p, dnorm.p);
        std::vector<double> rr(static_cast<size_t>(nvf) * B);
        BH_HIP(ctx, hipMemcpyAsync(rr.data(), dnorm.p, rr.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (int i = 0; i < nvf; ++i)
            for (int b = 0; b < B; ++b) {
                const double yy = dots[(2 * i) * B + b];
                resid_f[i * B + b] = yy > 0 ? std::sqrt(rr[i * B + b] / yy) : 0.0;
            }
        if (vecs_out) {
            // host copy [B][k][D] of the normalised Ritz vectors
            std::vector<double> hy(static_cast<size_t>(nvf) * vs);
            BH_HIP(ctx, hipMemcpyAsync(hy.data(), Yf.p, hy.size() * 8, hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
            vecs_out->assign(static_cast<size_t>(nreal) * k * D, 0.0);
            for (int b = 0; b < nreal; ++b)
                for (int i = 0; i < std::min(k, nvf); ++i) {
                    const double yy = dots[(2 * i) * B + b];
                    const double f = yy > 0 ? 1.0 / std::sqrt(yy) : 0.0;
                    for (int64_t r = 0; r < D; ++r)
                        (*vecs_out)[(static_cast<size_t>(b) * k + i) * D + r] = hy[static_cast<size_t>(i) * vs + r * B + b] * f;
                }
        }
    }

    // ---- results ----
    for (int b = 0; b < nreal; ++b) {
        RunHost& R = rh[b];
        const int nl = std::min(static_cast<int>(R.theta.size()), nvf);
        // levels after the FP64 finish, ascending (Ritz vectors of one degenerate level collapse)
        std::vector<std::pair<double, double>> lv;
        for (int i = 0; i < nl; ++i) lv.push_back({theta_f[i * B + b], resid_f[i * B + b]});
        std::sort(lv.begin(), lv.end());
        for (int i = 0; i < k; ++i) {
            const double ev = i < static_cast<int>(lv.size()) ? lv[i].first : NAN;
            const double rs = i < static_cast<int>(lv.size()) ? lv[i].second : NAN;
            // gap to the nearest other Ritz value of T (all computed low Ritz values)
            double gap = INFINITY;
            for (double th : R.all_theta)
                if (std::fabs(th - ev) > dedup * std::fmax(R.tnorm, 1e-300)) gap = std::fmin(gap, std::fabs(th - ev));
            const double err = std::isfinite(gap) ? rs * rs / gap : rs;
            if (res->evals) res->evals[b * k + i] = ev;
            if (res->resid) res->resid[b * k + i] = rs;
            if (res->errest) res->errest[b * k + i] = err;
        }
        if (res->iters) res->iters[b] = R.iters;
        if (res->status) res->status[b] = R.status;
        if (res->used_fp64) res->used_fp64[b] = MIXED ? 0 : 1;
    }
    return BH_OK;
}

template <typename XT>
static int lanczos_dispatch(bh_ctx* ctx, bh_ham* h, int Bp, const double* J, const double* U, int nreal,
                            const bh_lanczos_opts& o, bh_lanczos_result* r, std::vector<double>* vecs) {
    switch (Bp) {
        case 1: return lanczos_impl<XT, 1>(ctx, h, J, U, nreal, o, r, vecs);
        case 2: return lanczos_impl<XT, 2>(ctx, h, J, U, nreal, o, r, vecs);
        case 4: return lanczos_impl<XT, 4>(ctx, h, J, U, nreal, o, r, vecs);
        case 8: return lanczos_impl<XT, 8>(ctx, h, J, U, nreal, o, r, vecs);
        case 16: return lanczos_impl<XT, 16>(ctx, h, J, U, nreal, o, r, vecs);
        case 32: return lanczos_impl<XT, 32>(ctx, h, J, U, nreal, o, r, vecs);
        default: return lanczos_impl<XT, 64>(ctx, h, J, U, nreal, o, r, vecs);
    }
}

extern "C" {

void bh_lanczos_default_opts(bh_lanczos_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->k = 4;
    o->prec = BH_PREC_MIXED;
    o->max_iter = 1000;
    o->chunk = 16;
    o->tol = 1e-10;
    o->dedup_rel = 1e-9;
    o->accept_rel = 1e-9;
    o->seed = 0;
    o->fallback = 1;
    o->want_vectors = 0;
}

int bh_lanczos(bh_ctx* ctx, bh_ham* h, int B, const double* J, const double* U, const bh_lanczos_opts* opts,
               bh_lanczos_result* res) {
    if (!ctx || !h || !J || !U || !opts || !res) return BH_EINVAL;
    if (B < 1 || B > 64) return bh_set_error(ctx, BH_EINVAL, "lanczos: 1 <= B <= 64");
    if (opts->k < 1 || opts->k > LZ_NVMAX) return bh_set_error(ctx, BH_EINVAL, "lanczos: 1 <= k <= 16");
    if (opts->prec != BH_PREC_FP64 && opts->prec != BH_PREC_MIXED) return bh_set_error(ctx, BH_EINVAL, "lanczos: prec");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    int Bp = 1;
    while (Bp < B) Bp <<= 1;
    std::vector<double> vecs;
    std::vector<double>* vp = opts->want_vectors ? &vecs : nullptr;
    int rc;
    if (opts->prec == BH_PREC_FP64) {
        BH_TRY(bh_ham_prepare_locked(h, BH_PREC_FP64));
        rc = lanczos_dispatch<double>(ctx, h, Bp, J, U, B, *opts, res, vp);
    } else {
        rc = lanczos_dispatch<float>(ctx, h, Bp, J, U, B, *opts, res, vp);
    }
    if (rc != BH_OK) return rc;
    // mixed: acceptance of E0 by the FP64 finish; failing solves are rerun in FP64
    if (opts->prec == BH_PREC_MIXED && opts->fallback >= 0) {
        const double acc = opts->accept_rel > 0 ? opts->accept_rel : 1e-9;
        std::vector<int> bad;
        for (int b = 0; b < B; ++b) {
            const double e0 = res->evals ? res->evals[b * opts->k] : 0.0;
            const double er = res->errest ? res->errest[b * opts->k] : 0.0;
            const int stt = res->status ? res->status[b] : BH_OK;
            if (stt != BH_OK || !std::isfinite(e0) || !(er <= acc * std::fmax(std::fabs(e0), 1e-300))) bad.push_back(b);
        }
        if (!bad.empty()) {
            const int nb = static_cast<int>(bad.size());
            const int k = opts->k;
            std::vector<double> Jb(nb), Ub(nb), ev(nb * k), rs(nb * k), ee(nb * k);
            std::vector<int> it(nb), uf(nb), stt(nb);
            for (int i = 0; i < nb; ++i) {
                Jb[i] = J[bad[i]];
                Ub[i] = U[bad[i]];
            }
            bh_lanczos_opts o2 = *opts;
            o2.prec = BH_PREC_FP64;
            bh_lanczos_result r2{ev.data(), rs.data(), ee.data(), it.data(), uf.data(), stt.data()};
            BH_TRY(bh_ham_prepare_locked(h, BH_PREC_FP64));
            int Bp2 = 1;
            while (Bp2 < nb) Bp2 <<= 1;
            std::vector<double> v2;
            BH_TRY(lanczos_dispatch<double>(ctx, h, Bp2, Jb.data(), Ub.data(), nb, o2, &r2, vp ? &v2 : nullptr));
            for (int i = 0; i < nb; ++i) {
                const int b = bad[i];
                for (int q = 0; q < k; ++q) {
                    if (res->evals) res->evals[b * k + q] = ev[i * k + q];
                    if (res->resid) res->resid[b * k + q] = rs[i * k + q];
                    if (res->errest) res->errest[b * k + q] = ee[i * k + q];
                }
                if (res->iters) res->iters[b] += it[i];
                if (res->used_fp64) res->used_fp64[b] = 1;
                if (res->status) res->status[b] = stt[i];
                if (vp)
                    std::copy(v2.begin() + static_cast<int64_t>(i) * k * h->D, v2.begin() + static_cast<int64_t>(i + 1) * k * h->D,
                              vecs.begin() + static_cast<int64_t>(b) * k * h->D);
            }
        }
    }
    if (vp) {
        ctx->last_vecs.swap(vecs);
        ctx->last_vecs_B = B;
        ctx->last_vecs_k = opts->k;
        ctx->last_vecs_D = h->D;
    }
    int worst = BH_OK;
    if (res->status)
        for (int b = 0; b < B; ++b)
            if (res->status[b] != BH_OK) worst = BH_ENOCONV;
    if (worst != BH_OK) bh_set_error(ctx, worst, "lanczos: some solves did not converge within max_iter");
    return worst;
}

int bh_lanczos_copy_vectors(bh_ctx* ctx, int B, int k, double* vecs) {
    if (!ctx || !vecs) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (B != ctx->last_vecs_B || k != ctx->last_vecs_k || ctx->last_vecs.empty())
        return bh_set_error(ctx, BH_EINVAL, "lanczos_copy_vectors: no matching vectors kept (want_vectors)");
    std::memcpy(vecs, ctx->last_vecs.data(), ctx->last_vecs.size() * sizeof(double));
    return BH_OK;
}

int bh_tridiag_eigvals(int n, const double* d, const double* e, double* evals) {
    if (n < 1 || !d || !evals || (n > 1 && !e)) return BH_EINVAL;
    return bhtri::ql_eigvals(n, d, e, evals) == 0 ? BH_OK : BH_ENOCONV;
}

}  // extern "C"
