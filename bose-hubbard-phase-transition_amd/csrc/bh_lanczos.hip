// bh_lanczos.hip -- Lanczos eigensolver for B independent H_b = J_b T + U_b V sharing one CSR.
//
// Replaces Op::IRLM_eigen (src/operator.cpp:22-33), which runs Spectra's implicitly restarted
// Arnoldi (GenEigsBase.h:386-441; nev, ncv = 2 nev + 1, SmallestReal, tol 1e-10) with a serial
// sparse product per step.  Here, per Lanczos step j, four launches on one stream and no host
// round trip:
//   K1  w = H v_j - beta_j v_{j-1} (one CSR pass for all B vectors, fused with the diagonal and the
//       recurrence), written to W shifted by alpha_{j-1} v_j in storage precision (lz_sigma); per-workgroup
//       partials of alpha_j = <v_j, w> (FP64, unshifted) and, for the
//       locked Ritz vectors y_l, of <y_l, w> and <y_l, v_j>
//   R1  the partials summed in a fixed order (one wave per value)
//   K2  w'' = w - alpha_j v_j - sum_l c_l y_l with c_l = <y_l, w> - alpha_j <y_l, v_j>, stored as
//       v_{j+1} (unnormalised, storage precision); partials of ||w''||^2
//   R2  beta_{j+1} = ||w''||, scale_{j+1} = 1 / beta_{j+1}, step counter advanced
// Cross-workgroup sums go through these separate launches instead of a last-workgroup reduction:
// on a multi-XCD device each workgroup's device-scope release would write back its XCD's L2.
// Every `chunk` steps the host analyses T_j in FP64 (Sturm bisection + inverse iteration), applies
// Spectra's convergence rule |beta_j s_last| < tol max(eps^(2/3), |theta|), and performs selective
// orthogonalisation (Parlett & Scott): a Ritz pair of a wanted level whose bound is below
// sqrt(eps_store) ||T|| is locked, its Ritz vector formed on the device, and later Lanczos vectors
// are kept orthogonal to it.  The finish forms the Ritz vectors in FP64 and evaluates Rayleigh
// quotients and residuals with the exact FP64 values of T.
// Start vector: Spectra's Park-Miller LCG r (SimpleRandom.h:30-96) normalised, v0 = r / ||r||.
// Spectra projects it first (v0 = A r / ||A r||, Arnoldi.h:147-154); that projection removes the null
// space of H, and at J = 0 the unit-filling Mott state has E = 0 exactly, so it is not applied.
// Mixed precision: Krylov and locked vectors and the shifted W stored FP32, every reduction FP64; the matrix
// values come exactly from the FP64 table of the packed format.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <memory>
#include <thread>

#include "bh_kernels.hpp"
#include "bh_spmv.hpp"
#include "bh_tridiag.hpp"
#include "bh_engine.hpp"

constexpr int LZ_WG = 512;         // threads per workgroup of the Lanczos kernels (8 waves)
constexpr int LZ_GRID_MAX = 1024;  // workgroups that write partial sums
constexpr int LZ_LMAX = 16;        // locked (and frozen) vectors per lane with selective orthogonalisation
constexpr int LZ_NVMAX = 16;       // vectors per register-tiled launch (locks formed at once, Gram and combine tiles)
constexpr int LZ_KMAX = 64;        // wanted eigenvalues per solve
constexpr int LZ_CMAX = 2 * LZ_KMAX + 8;  // candidates of the FP64 finish (frozen vectors + the last pass's)
constexpr int LZ_SO_KMAX = 6;      // multiplicities: larger k runs with full reorthogonalisation (frozen + locked
                                   // vectors of a selectively orthogonalised lane share its LZ_LMAX slots)
constexpr int LZ_PASS_MAX = 64;    // deflation passes per solve
constexpr int LZ_GRAM_VALS = 2 * 2 * LZ_NVMAX;  // partial sums per lane of one Gram-tile launch (NI 2 x NV 16, or 4 x 8)
constexpr int LZ_RED_WG = 1024;    // reduction workgroup: 16 waves, one value per wave
constexpr int LZ_WS_A = 5, LZ_WS_B = 6, LZ_WS_C = 7;  // context workspace slots of the engine's scratch (0-4: basis,
                                                       // W, locked vectors, start vector, current vectors)
constexpr int LZ_GRAPH_STEPS = 8;  // Lanczos steps per captured HIP graph
constexpr int LZ_K2_GRID = 1024;   // K2 workgroups at most (rows of partial sums for R2; 512 and 2048 measured slower)
constexpr int LZ_K1_TILES = 4;         // row groups per K1 workgroup on the row-lane engine
constexpr int LZ_TUNE_REPS = 3;        // timed launches per K1 variant in the autotuner
constexpr int LZ_CHECK_MIN = 4;        // fewest steps between two analyses of a lane
constexpr int64_t LZ_SPEC_MIN_WORK = 4ll << 20;  // D B from which analyses overlap speculative steps
constexpr int64_t LZ_PART_BYTES = 64ll << 20;  // partial sums of the staged K1 (more workgroups than grid_rl)
constexpr int LZ_K1_GRID_MAX = 32768;  // K1 workgroups at most (rows of partial sums)
constexpr int LZ_REFINE_STEPS = 3;         // Davidson steps of the FP64 refinement of a rejected mixed lane
constexpr double LZ_REFINE_CLAMP = 0.1;    // refinement: |diag(H) - theta| held at least this x max(|U|, |J|)
constexpr double LZ_ACCEPT_LEVEL = 1e-6;  // mixed: bound on ||r||^2/gap of the levels above E0, relative

struct LzState {
    double* alpha;   // [slots][B]
    double* beta;    // [slots+1][B]
    double* scale;   // [slots+1][B]  1 / ||stored vector||
    double* J;       // [B]
    double* U;       // [B]
    double* sums;    // [(1 + 2 LZ_LMAX) B] R1 results: alpha, <y_l, w>, <y_l, v_j>
    double* tau;     // [2][LZ_LMAX][B] estimates of |<y_l, v_j>| and |<y_l, v_{j-1}>| (locked y_l)
    double* ltheta;  // [LZ_LMAX][B] Ritz values of the locked vectors
    double* anorm;   // [B] estimate of ||H_b||
    double* dso;     // [3][LZ_LMAX][B] <y_l, v_j> and <y_l, v_{j-1}> (normalised vectors) for the one-read
                     //     steps, and the step's coefficients c_l
    int* active;     // [B]
    int* nlock;      // [B]
    int* step;       // [B] Lanczos step of each lane (lanes are admitted and retired independently)
    int* so;         // [2] orthogonalisation against the locked vectors in this step / the previous one:
                     //     0 none, 1 exact (K1 forms the dot products), 2 one-read (from the carried products)
    double* part;    // [LZ_GRID_MAX][nv] per-workgroup partials
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
    XT* V;         // [slots][D][B] Krylov basis; lane b's vector t at V[t][r][b]
    XT* Xc;        // [D][B] current Lanczos vector of every lane (the operand of the gathers)
    int64_t vs;    // D * B
    XT* W;         // [D][B] w - sigma_b x_b in storage precision (sigma_b = alpha_{j-1}, see lz_sigma)
    XT* Y;         // [LZ_LMAX][D][B] locked Ritz vectors (unit norm)
    int maxit;     // slots - 1: a lane stops at this step until it is restarted or retired
    LzState st;
};

// Shift of the stored operator image: K1 stores w - sigma_b x_b (x_b = s_j v_j) with sigma_b = alpha_{j-1} (0 at
// j = 0), which is of the order of the next beta where w itself is of the order of |alpha_j|, so that the
// image loses no more to its storage rounding than v_{j+1} does; K2 and the exact-step dot products add
// sigma_b x_b back (alpha_j itself is summed from the unshifted w).  Any sigma would do, as long as K1, K2 and
// the dot products of one step read the same one.
template <int B>
__device__ __forceinline__ double lz_sigma(const LzState& st, int b, int j) {
    return j > 0 ? st.alpha[(j - 1) * B + b] : 0.0;
}

// Per-workgroup sum of one value per vector b (b = lane % B), stored at part[wg * nv + q * B + b].
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

// Per-workgroup sums of NQ values per vector b: v[q] of every lane is summed over the lanes of its b and
// stored at part[wg * nv + q * B + b].  Up to LZ_PUT_Q values share one LDS round (LZ_PUT_LDS doubles).
constexpr int LZ_PUT_LDS = 4096;

template <int B, int NQ>
__device__ __forceinline__ void lz_block_put_multi(const double (&v)[NQ], double* lds, double* part, int nv) {
    constexpr int QC = (LZ_PUT_LDS / (LZ_WG / 64 * B)) < 1 ? 1 : (LZ_PUT_LDS / (LZ_WG / 64 * B));
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q0 = 0; q0 < NQ; q0 += QC) {
#pragma unroll
        for (int q = q0; q < NQ && q < q0 + QC; ++q) {
            double t = v[q];
#pragma unroll
            for (int o = B; o < 64; o <<= 1) t += __shfl_xor(t, o, 64);
            if (lane < B) lds[(wid * QC + (q - q0)) * B + lane] = t;
        }
        __syncthreads();
        const int nq = (NQ - q0) < QC ? (NQ - q0) : QC;
        for (int i = threadIdx.x; i < nq * B; i += LZ_WG) {
            const int q = i / B, bb = i % B;
            double t = 0.0;
            for (int w = 0; w < LZ_WG / 64; ++w) t += lds[(w * QC + q) * B + bb];
            part[static_cast<size_t>(blockIdx.x) * nv + (q0 + q) * B + bb] = t;
        }
        __syncthreads();
    }
}

// Per-workgroup sums of V values per thread, thread t holding vectors b = (t % (B / V)) * V + v, stored at
// part[wg * nv + q * B + b] (the V = 1 case is lz_block_put).
template <int B, int V>
__device__ __forceinline__ void lz_block_put_v(const double (&x)[V], int q, double* lds, double* part, int nv) {
    constexpr int TPR = B / V;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        double t = x[v];
#pragma unroll
        for (int o = TPR; o < 64; o <<= 1) t += __shfl_xor(t, o, 64);
        if (lane < TPR) lds[wid * B + lane * V + v] = t;
    }
    __syncthreads();
    if (threadIdx.x < B) {
        double s = 0.0;
        for (int w = 0; w < LZ_WG / 64; ++w) s += lds[w * B + threadIdx.x];
        part[static_cast<size_t>(blockIdx.x) * nv + q * B + threadIdx.x] = s;
    }
    __syncthreads();
}

// Per-lane scalars of a K1 step, staged in LDS at the start of the kernel (registers stay free for the
// gathers): s_j, J_b s_j, U_b, beta_j s_{j-1}, the offset of v_{j-1} (slot j - 1; slot 0 when j = 0) and
// whether the lane takes the step.
template <int B>
struct K1Shared {
    double sj[B], jbs[B], ub[B], bj[B], sg[B];
    int64_t om[B];
    int act[B];
};

// K1 epilogue shared by both operator engines: a thread holds the V vectors b0 .. b0 + V - 1 of its rows.
// w = J_b T x_b + U_b V x_b - beta_j v_{j-1} (x_b = s_j v_j, the stored vector times its scale) is written
// to W; the thread accumulates <x_b, w> for alpha.
template <typename XT, int B, int V, int FMT>
struct K1Epi {
    const LzArgs<XT>& a;
    const K1Shared<B>& ls;
    int b0 = 0;
    int64_t om0 = 0;    // the common v_{j-1} offset when `same`
    bool any = false;   // some lane of the thread takes this step
    bool same = true;   // every lane that takes the step reads v_{j-1} from one slot (one vector load)
    double pa[V];
    __device__ __forceinline__ void operator()(int64_t r, int, const double (&acc)[V]) {
        if (!any) return;  // the row-block engine calls every thread
        const double d = a.diag32 ? static_cast<double>(a.diag32[r]) : a.diag64[r];
        const int64_t idx = r * B + b0;
        const BhVec<XT, V> xv = bh_vload<V>(a.Xc + idx);
        BhVec<XT, V> xm;
        if (same) {
            xm = bh_vload<V>(a.V + om0 + idx);
        } else {
#pragma unroll
            for (int v = 0; v < V; ++v) xm.v[v] = a.V[ls.om[b0 + v] + idx + v];
        }
        BhVec<XT, V> wv;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const int b = b0 + v;
            const double hx = (FMT == BH_FMT_PACKED) ? -acc[v] : acc[v];
            const double xr = to_d(xv.v[v]) * ls.sj[b];
            const double w = ls.jbs[b] * hx + ls.ub[b] * d * xr - ls.bj[b] * to_d(xm.v[v]);
            wv.v[v] = static_cast<XT>(w - ls.sg[b] * xr);
            if (ls.act[b]) pa[v] += xr * w;
        }
        *reinterpret_cast<BhVec<XT, V>*>(a.W + idx) = wv;
    }
};

// fills the LDS lane table (threads < B) and the thread's view of its V lanes; ends with a barrier
template <typename XT, int B, int V, int FMT>
__device__ __forceinline__ void k1_setup(const LzArgs<XT>& a, K1Shared<B>& ls, K1Epi<XT, B, V, FMT>& e) {
    if (threadIdx.x < B) {
        const int b = threadIdx.x;
        const int j = a.st.step[b];
        const double sj = a.st.scale[j * B + b];
        ls.sj[b] = sj;
        ls.jbs[b] = a.st.J[b] * sj;
        ls.ub[b] = a.st.U[b];
        ls.bj[b] = j > 0 ? a.st.beta[j * B + b] * a.st.scale[(j - 1) * B + b] : 0.0;
        ls.sg[b] = lz_sigma<B>(a.st, b, j);
        ls.om[b] = static_cast<int64_t>(j > 0 ? j - 1 : 0) * a.vs;
        ls.act[b] = a.st.active[b] != 0 && j < a.maxit;
    }
    __syncthreads();
    e.b0 = (threadIdx.x % (B / V)) * V;  // fixed per thread: LZ_WG is a multiple of B
    int64_t om0 = -1;
#pragma unroll
    for (int v = 0; v < V; ++v) {
        e.pa[v] = 0.0;
        if (ls.act[e.b0 + v]) {
            e.any = true;
            if (om0 < 0) om0 = ls.om[e.b0 + v];
            e.same = e.same && ls.om[e.b0 + v] == om0;
        }
    }
    // lanes that do not take the step read (and discard) the active lanes' v_{j-1} slot
    e.om0 = om0 < 0 ? 0 : om0;
}

// K1 on the LDS-staged row-block engine (B = 1 in the automatic choice)
template <typename XT, int B, int FMT>
__global__ void __launch_bounds__(LZ_WG) k_lz_a_rb(LzArgs<XT> a, const int32_t* __restrict__ rb, int64_t nblocks, int nv) {
    extern __shared__ double smem[];
    double* tab = smem;
    double* prod = smem + (FMT == BH_FMT_PACKED ? (1 << a.cb) : 0);
    __shared__ double lds[LZ_WG / 64 * B];
    if constexpr (FMT == BH_FMT_PACKED) {
        for (int i = threadIdx.x; i < (1 << a.cb); i += LZ_WG) tab[i] = a.table[i];
        __syncthreads();
    }
    __shared__ K1Shared<B> ls;
    K1Epi<XT, B, 1, FMT> e{a, ls};
    k1_setup(a, ls, e);
    const int64_t per = (nblocks + gridDim.x - 1) / gridDim.x;
    const int64_t L = bh_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t0 = min(L * per, nblocks), t1 = min(t0 + per, nblocks);
    bh_rowblock_run<XT, B, FMT, LZ_WG>(t0, t1, rb, a.row_ptr, a.words, a.col64, a.val64, tab, a.cb, a.Xc, prod, e);
    lz_block_put_v<B, 1>(e.pa, 0, lds, a.st.part, nv);
}

// K1 on the row-lane engine (B / V threads per row, contiguous row ranges per workgroup)
template <typename XT, int B, int V, int FMT, int U>
__device__ __forceinline__ void k1_rowlane(const LzArgs<XT>& a, int64_t rows_per_wg, int nv) {
    extern __shared__ double tab[];
    __shared__ double lds[LZ_WG / 64 * B];
    if constexpr (FMT == BH_FMT_PACKED) {
        for (int i = threadIdx.x; i < (1 << a.cb); i += LZ_WG) tab[i] = a.table[i];
    }
    __shared__ K1Shared<B> ls;
    K1Epi<XT, B, V, FMT> e{a, ls};
    k1_setup(a, ls, e);
    const int64_t L = bh_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t r0 = min(L * rows_per_wg, a.D), r1 = min(r0 + rows_per_wg, a.D);
    if (e.any)
        bh_rowlane_run<XT, B, V, FMT, LZ_WG, U>(r0, r1, a.row_ptr, a.words, a.col64, a.val64, tab, a.cb, a.Xc, e);
    lz_block_put_v<B, V>(e.pa, 0, lds, a.st.part, nv);
}

// V <= 4, U = 4: at most 80 VGPRs (six waves per SIMD)
template <typename XT, int B, int V, int FMT>
__global__ void __launch_bounds__(LZ_WG) __attribute__((amdgpu_waves_per_eu(6))) k_lz_a_rl(LzArgs<XT> a, int64_t rows_per_wg,
                                                                                            int nv) {
    k1_rowlane<XT, B, V, FMT, 4>(a, rows_per_wg, nv);
}

// wide variants: V = 8 (two 16-byte gathers per entry and thread) or U = 8, 16 entries of a row in flight
// (long rows at small B)
template <typename XT, int B, int V, int FMT, int U>
__global__ void __launch_bounds__(LZ_WG) k_lz_a_rlw(LzArgs<XT> a, int64_t rows_per_wg, int nv) {
    k1_rowlane<XT, B, V, FMT, U>(a, rows_per_wg, nv);
}

// Dot products of an exact orthogonalisation step (so = 1): partials of <y_l, w> and <y_l, s_j v_j> for the
// locked vectors l < nlock, at part[wg * nv + (1 + l) * B + b] and part[wg * nv + (1 + LM + l) * B + b] over
// the same rows of partials as K1 (which writes q = 0).  Other steps leave at once.
template <typename XT, int B, int LM>
__global__ void __launch_bounds__(LZ_WG) k_lz_sodots(LzArgs<XT> a, int64_t rows_per_wg, int nv) {
    if (a.st.so[0] != 1) return;  // uniform
    __shared__ double lds[LZ_PUT_LDS];
    const int b = threadIdx.x % B;
    const int j = a.st.step[b];
    const bool act = a.st.active[b] != 0 && j < a.maxit;
    const int nl = act ? min(a.st.nlock[b], LM) : 0;
    const double sj = a.st.scale[j * B + b], sg = lz_sigma<B>(a.st, b, j);
    double acc[2 * LM];
#pragma unroll
    for (int q = 0; q < 2 * LM; ++q) acc[q] = 0.0;
    const int64_t L = bh_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t r0 = min(L * rows_per_wg, a.D), r1 = min(r0 + rows_per_wg, a.D);
    if (nl > 0) {
        for (int64_t r = r0 + threadIdx.x / B; r < r1; r += LZ_WG / B) {
            const int64_t idx = r * B + b;
            const double xr = to_d(a.Xc[idx]) * sj, w = to_d(a.W[idx]) + sg * xr;
            // the locked vectors in groups of four, the group's loads issued together (a vector l >= nl reads
            // slot nl - 1 again and adds zeros)
#pragma unroll
            for (int l0 = 0; l0 < LM; l0 += 4) {
                if (l0 >= nl) break;
                XT y[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) y[q] = a.Y[static_cast<int64_t>(min(l0 + q, nl - 1)) * a.vs + idx];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const double yd = l0 + q < nl ? to_d(y[q]) : 0.0;
                    acc[l0 + q] += yd * w;
                    acc[LM + l0 + q] += yd * xr;
                }
            }
        }
    }
    lz_block_put_multi<B, 2 * LM>(acc, lds, a.st.part + B, nv);
}

// K2: v_{j+1} = w - alpha v_j - sum_l c_l y_l (stored in slot j+1 and, replacing v_j, in the current-vector
// buffer; each thread reads its own elements before writing them), partials of ||v_{j+1}||^2 and of
// e_l = <y_l, w - alpha v_j>, from which R2 forms <y_l, v_{j+1}> = (e_l - c_l) / beta_{j+1} (the locked vectors
// are orthonormal; the cross terms c_m <y_l, y_m> are products of two small quantities).  Block 0 records
// alpha_j and the c_l.  The coefficient c_l = <y_l, w> - alpha_j <y_l, v_j> comes from
// K1's dot products in an exact step (so = 1) and, in a one-read step (so = 2), from the products K2 formed in
// the two previous steps:  <y_l, w> = <H y_l, v_j> - beta_j <y_l, v_{j-1}> ~ theta_l <y_l, v_j> - beta_j
// <y_l, v_{j-1}>, dropping <r_l, v_j> for the locked vector's residual r_l (in exact arithmetic r_l lies along
// one earlier Lanczos vector, so the term is at the level of the basis' loss of orthogonality).  A one-read
// step reads each locked vector once instead of twice (K1 skips them).
// Eight locked vectors: held to 64 VGPRs (eight waves per SIMD, the 1024-workgroup grid in one round; 20 bytes
// of spills per lane) -- 131 us per launch at the chain m=12 sector against 138 us at 66 VGPRs and seven waves.
template <typename XT, int B, int LM>
__global__ void __launch_bounds__(LZ_WG) __attribute__((amdgpu_waves_per_eu(LM == 8 ? 8 : 1))) k_lz_b(LzArgs<XT> a) {
    constexpr int LL = LM > 0 ? LM : 1;
    constexpr int NV2 = (1 + LM) * B;
    __shared__ double lds[LZ_PUT_LDS];
    const int b = threadIdx.x % B;
    const int j = a.st.step[b];
    XT* vj = a.Xc;
    XT* vn = a.V + static_cast<int64_t>(j + 1) * a.vs;
    const bool act = a.st.active[b] != 0 && j < a.maxit;
    const double sj = a.st.scale[j * B + b];
    const double al = a.st.sums[b], sg = lz_sigma<B>(a.st, b, j);
    const int mode = a.st.so[0];
    const int nl = (act && LM > 0 && mode != 0) ? min(a.st.nlock[b], LM) : 0;
    if (blockIdx.x == 0 && threadIdx.x < B && act) a.st.alpha[j * B + b] = al;
    // acc[0]: ||v_{j+1}||^2;  acc[1 + l]: e_l = <y_l, w - alpha v_j>, from which the reduction forms
    // <y_l, v_{j+1}> = (e_l - c_l) / beta_{j+1} (the locked vectors are orthonormal)
    double acc[1 + LM];
#pragma unroll
    for (int q = 0; q <= LM; ++q) acc[q] = 0.0;
    // c_l of every (l, lane), one thread each, staged in LDS and read from there where it is used (registers
    // stay free for the accumulators; workgroup 0 also records them)
    __shared__ double cls[LL * B];
    if constexpr (LM > 0) {
        for (int q = threadIdx.x; q < LM * B; q += LZ_WG) {
            const int bb = q % B, l = q / B, jj = a.st.step[bb];
            const bool aa = a.st.active[bb] != 0 && jj < a.maxit;
            const int nll = (aa && mode != 0) ? min(a.st.nlock[bb], LM) : 0;
            double c = 0.0;
            if (l < nll) {
                const double all = a.st.sums[bb];
                if (mode == 1) {
                    c = a.st.sums[(1 + l) * B + bb] - all * a.st.sums[(1 + LM + l) * B + bb];
                } else {
                    const double bj = jj > 0 ? a.st.beta[jj * B + bb] : 0.0;
                    c = (a.st.ltheta[l * B + bb] - all) * a.st.dso[l * B + bb] - bj * a.st.dso[(LZ_LMAX + l) * B + bb];
                }
            }
            cls[q] = c;
            if (blockIdx.x == 0 && aa) a.st.dso[(2 * LZ_LMAX + l) * B + bb] = c;
        }
        __syncthreads();
    }
    constexpr int ROWS = LZ_WG / B;
    // rows per thread in flight: one with sixteen locked vectors (their accumulators and a group of loads in
    // flight already take the registers; two rows per thread were slower at the chain m=12 sector)
    constexpr int UN = LM > 4 ? 1 : (LM > 0 ? 2 : 4);
    const int64_t ntile = (a.D + ROWS - 1) / ROWS;
    const int64_t per = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t L = bh_xcd_tile(blockIdx.x, gridDim.x);
    const int64_t t0 = min(L * per, ntile), t1 = min(t0 + per, ntile);
    if (act) {
        for (int64_t t = t0; t < t1; t += UN) {
            double w[UN], v[UN];
            int64_t idx[UN];
            bool ok[UN];
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const int64_t r = (t + u) * ROWS + threadIdx.x / B;
                ok[u] = (t + u < t1) && r < a.D;
                idx[u] = ok[u] ? r * B + b : 0;  // a valid element either way: the loads below are unconditional
            }
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                const double wl = to_d(a.W[idx[u]]);
                const double vl = to_d(vj[idx[u]]);
                w[u] = ok[u] ? wl : 0.0;
                v[u] = ok[u] ? vl : 0.0;
            }
#pragma unroll
            for (int u = 0; u < UN; ++u) w[u] -= (al - sg) * v[u] * sj;
            if constexpr (LM > 0) {
                double w0[UN];
#pragma unroll
                for (int u = 0; u < UN; ++u) w0[u] = w[u];
                // the locked vectors in groups of LG, the group's loads issued together (a vector l >= nl of a
                // group reads slot nl - 1 again, a cache hit, and is discarded); the subtractions run over l in order
                constexpr int LG = LL < 4 ? LL : 4;
#pragma unroll
                for (int l0 = 0; l0 < LL; l0 += LG) {
                    if (l0 >= nl) break;
                    XT y[LG][UN];
#pragma unroll
                    for (int q = 0; q < LG; ++q) {
                        const int64_t off = static_cast<int64_t>(min(l0 + q, nl - 1)) * a.vs;
#pragma unroll
                        for (int u = 0; u < UN; ++u) y[q][u] = a.Y[off + idx[u]];
                    }
#pragma unroll
                    for (int q = 0; q < LG; ++q) {
#pragma unroll
                        for (int u = 0; u < UN; ++u) {
                            // every vector of the group is consumed (a value selected to zero for l >= nl adds
                            // nothing), so that no load is sunk into a branch
                            const double yd = (ok[u] && l0 + q < nl) ? to_d(y[q][u]) : 0.0;
                            acc[1 + l0 + q] += yd * w0[u];
                            w[u] -= cls[(l0 + q) * B + b] * yd;
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < UN; ++u) {
                if (!ok[u]) continue;
                const XT wr = static_cast<XT>(w[u]);
                vn[idx[u]] = wr;
                vj[idx[u]] = wr;
                const double wd = to_d(wr);
                acc[0] += wd * wd;
            }
        }
    }
    if (LM > 0 && mode != 0) {  // uniform
        lz_block_put_multi<B, 1 + LM>(acc, lds, a.st.part, NV2);
    } else {
        double a0[1] = {acc[0]};
        lz_block_put_multi<B, 1>(a0, lds, a.st.part, NV2);
    }
}

// Full reorthogonalisation (small sectors), classical Gram-Schmidt pass over the whole basis: workgroup t
// computes dd[t][b] = scale_t <V[t][.][b], v_{j+1}[.][b]> for the lanes whose step j has t <= j; workgroups
// t >= maxit do the same for the lane's frozen vectors Y[t - maxit] (t - maxit < nlock[b]; unit norm), so that
// a deflation pass stays orthogonal to the eigenvectors the earlier passes froze.
template <typename XT, int B>
__global__ void __launch_bounds__(LZ_WG) k_lz_fr_dots(LzArgs<XT> a, double* __restrict__ dd) {
    __shared__ double lds[LZ_WG / 64 * B];
    const int t = blockIdx.x;
    const int b = threadIdx.x % B;
    const int j = a.st.step[b];
    const bool frozen = t >= a.maxit;
    const bool act = a.st.active[b] != 0 && j < a.maxit && (frozen ? t - a.maxit < a.st.nlock[b] : t <= j);
    double acc = 0.0;
    if (act) {
        const XT* vt = frozen ? a.Y + static_cast<int64_t>(t - a.maxit) * a.vs : a.V + static_cast<int64_t>(t) * a.vs;
        const XT* vn = a.V + static_cast<int64_t>(j + 1) * a.vs;
        for (int64_t r = threadIdx.x / B; r < a.D; r += LZ_WG / B) acc += to_d(vt[r * B + b]) * to_d(vn[r * B + b]);
        if (!frozen) acc *= a.st.scale[t * B + b];
    }
    lz_block_put<B>(acc, 0, lds, dd, B);
}

// v_{j+1} -= sum_{t <= j} dd[t] scale_t V[t] + sum_{l < nlock} dd[maxit + l] Y[l]; the last pass also refreshes
// the current vector and writes the partials of ||v_{j+1}||^2 (as K2 does)
template <typename XT, int B>
__global__ void __launch_bounds__(LZ_WG) k_lz_fr_axpy(LzArgs<XT> a, const double* __restrict__ dd, int last) {
    __shared__ double lds[LZ_WG / 64 * B];
    const int b = threadIdx.x % B;
    const int j = a.st.step[b];
    const int nfz = a.st.nlock[b];
    const bool act = a.st.active[b] != 0 && j < a.maxit;
    constexpr int ROWS = LZ_WG / B;
    const int64_t ntile = (a.D + ROWS - 1) / ROWS;
    const int64_t per = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t t0 = min(static_cast<int64_t>(blockIdx.x) * per, ntile), t1 = min(t0 + per, ntile);
    double pn = 0.0;
    if (act) {
        XT* vn = a.V + static_cast<int64_t>(j + 1) * a.vs;
        for (int64_t tt = t0; tt < t1; ++tt) {
            const int64_t r = tt * ROWS + threadIdx.x / B;
            if (r >= a.D) continue;
            const int64_t idx = r * B + b;
            double w = to_d(vn[idx]);
            for (int t = 0; t <= j; ++t) w -= dd[t * B + b] * a.st.scale[t * B + b] * to_d(a.V[static_cast<int64_t>(t) * a.vs + idx]);
            for (int l = 0; l < nfz; ++l) w -= dd[(a.maxit + l) * B + b] * to_d(a.Y[static_cast<int64_t>(l) * a.vs + idx]);
            const XT wr = static_cast<XT>(w);
            vn[idx] = wr;
            if (last) {
                a.Xc[idx] = wr;
                const double wd = to_d(wr);
                pn += wd * wd;
            }
        }
    }
    if (last) lz_block_put<B>(pn, 0, lds, a.st.part, B);
}

// V[0][r][b] = r0[r] in storage precision (start vector shared by all b)
template <typename XT, int B>
__global__ void k_lz_init(XT* __restrict__ V0, const double* __restrict__ r0, int64_t D) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx < D * B) V0[idx] = static_cast<XT>(r0[idx / B]);
}

// Plain reduction of per-workgroup partials: out[v] = sum_wg part[wg * nv + v], one wave per value, lanes
// striding over the workgroups in a fixed order (deterministic); values spread over the launch's workgroups.
__global__ void __launch_bounds__(LZ_RED_WG) k_lz_reduce(const double* __restrict__ part, int nwg, int nv,
                                                         double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int waves = LZ_RED_WG / 64;
    for (int v = blockIdx.x * waves + (threadIdx.x >> 6); v < nv; v += gridDim.x * waves) {
        double s = 0.0;
        for (int w = lane; w < nwg; w += 64) s += part[static_cast<size_t>(w) * nv + v];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) out[v] = s;
    }
}

// R2: beta_{j+1} = ||v_{j+1}||, 1/beta, each active lane's step advanced; the carried products
// <y_l, v_{j+1}> = (e_l - c_l) / beta_{j+1} and the Parlett-Scott estimates of the locked vectors; the next
// step's orthogonalisation flag.  K2's partial sums are first reduced by k_lz_reduce_wide (one workgroup per
// value, fixed order) into `pre`; this one workgroup stages them in LDS, then updates the locked-vector
// quantities with a thread per (vector, lane), then each lane's beta and step counter with a thread per lane
// (the step counters are read by all of phase 2a before any is written).
constexpr int LZ_BETA_MAXV = (1 + LZ_LMAX) * 64;

template <int B>
__global__ void __launch_bounds__(LZ_RED_WG) k_lz_beta(LzState st, const double* __restrict__ pre, int nv, int maxit,
                                                       double eps, int so_always) {
    __shared__ double sums[LZ_BETA_MAXV];
    __shared__ int want;
    const int so_cur = st.so[0], so_prev = st.so[1];
    const int nvv = so_cur ? nv : B;
    if (threadIdx.x == 0) want = 0;
    for (int v = threadIdx.x; v < nvv; v += LZ_RED_WG) sums[v] = pre[v];
    __syncthreads();
    // phase 2a: one thread per (locked vector l, lane b), all in parallel
    for (int q = threadIdx.x; q < LZ_LMAX * B; q += LZ_RED_WG) {
        const int b = q % B, l = q / B;
        const int j = st.step[b];
        if (!st.active[b] || j >= maxit || l >= min(st.nlock[b], LZ_LMAX)) continue;
        const double nrm = std::sqrt(sums[b]);
        const double al = st.sums[b], bj = st.beta[j * B + b];
        const size_t i0 = static_cast<size_t>(l) * B + b, i1 = static_cast<size_t>(LZ_LMAX + l) * B + b;
        // Parlett-Scott estimate of |<y_l, v_{j+1}>| (reset after an orthogonalising step)
        const double tj = st.tau[i0], tjm = st.tau[i1];
        const double tn = so_cur ? eps
                                 : (nrm > 0.0 ? (std::fabs(st.ltheta[i0] - al) * tj + bj * tjm + eps * st.anorm[b]) / nrm
                                              : 1.0);
        st.tau[i1] = tj;
        st.tau[i0] = tn;
        if (tn > std::sqrt(eps)) atomicOr(&want, 1);
        if (so_cur && (1 + l) * B + b < nv && nrm > 0.0) {
            st.dso[i1] = st.dso[i0];
            st.dso[i0] = (sums[(1 + l) * B + b] - st.dso[(2 * LZ_LMAX + l) * B + b]) / nrm;
        }
    }
    __syncthreads();  // every thread of 2a has read the step counters
    // phase 2b: one thread per lane
    if (threadIdx.x < B) {
        const int b = threadIdx.x;
        const int j = st.step[b];
        if (st.active[b] && j < maxit) {
            const double nrm = std::sqrt(sums[b]);
            st.beta[(j + 1) * B + b] = nrm;
            st.scale[(j + 1) * B + b] = nrm > 0.0 ? 1.0 / nrm : 0.0;
            st.step[b] = j + 1;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        // an exact step that opened a pair is followed by a second exact step; then one-read steps (default)
        // or, with the estimate-driven schedule, exact pairs when the estimate asks
        st.so[0] = (so_cur == 1 && so_prev != 1) ? 1 : (so_always ? 2 : (want ? 1 : 0));
        st.so[1] = so_cur;
    }
}

// R1 over many K1 workgroups: one workgroup per value (blockIdx.x = v), each thread summing a strided
// subset of the partials in a fixed order, then a fixed tree over the workgroup.  Only the alpha sums
// (v < nmin) unless the step is exact (then also the locked-vector sums).
template <int B>
__global__ void __launch_bounds__(LZ_RED_WG) k_lz_reduce_wide(LzState st, const double* __restrict__ part, int nwg,
                                                              int nv, int nmin, int so_any, double* __restrict__ out) {
    __shared__ double red[LZ_RED_WG / 64];
    const int v = blockIdx.x;
    const int nvv = (so_any ? st.so[0] != 0 : st.so[0] == 1) ? nv : nmin;
    if (v >= nvv) return;  // uniform over the workgroup
    double s = 0.0;
    for (int w = threadIdx.x; w < nwg; w += LZ_RED_WG) s += part[static_cast<size_t>(w) * nv + v];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < LZ_RED_WG / 64; ++w) t += red[w];
        out[v] = t;
    }
}

// admission of lane `slot`: V[0][r][slot] = Xc[r][slot] = r0[r] in storage precision
template <typename XT, int B>
__global__ void k_lz_admit(XT* __restrict__ V0, XT* __restrict__ Xc, const double* __restrict__ r0, int64_t D, int slot) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r < D) {
        const XT x = static_cast<XT>(r0[r]);
        V0[r * B + slot] = x;
        Xc[r * B + slot] = x;
    }
}

// Thick restart of the lanes with nin[b] > 0, in place: V[c][r][b] = sum_{t < nin[b]} V[t][r][b] C[(t nv + c) B + b]
// for c < nout[b], and the last of them also becomes the lane's current vector.  Every thread reads all of
// its (r, b) column before writing any of it, and no two threads share a column, so the in-place update
// is safe.
constexpr int LZ_KEEP_MAX = 32;  // restart outputs per lane (kept Ritz vectors + the residual vector)

template <typename XT, int B>
__global__ void __launch_bounds__(256) k_lz_restart(XT* V, XT* __restrict__ Xc, int64_t vs, const int* __restrict__ nin,
                                                    const int* __restrict__ nout, const double* __restrict__ C, int nv) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    const int nj = nin[b], no = nout[b];
    if (nj <= 0 || no <= 0) return;
    double acc[LZ_KEEP_MAX];
#pragma unroll
    for (int i = 0; i < LZ_KEEP_MAX; ++i) acc[i] = 0.0;
    for (int t = 0; t < nj; ++t) {
        const double v = to_d(V[static_cast<int64_t>(t) * vs + idx]);
        const double* ct = C + (static_cast<int64_t>(t) * nv) * B + b;
#pragma unroll
        for (int i = 0; i < LZ_KEEP_MAX; ++i)
            if (i < no) acc[i] += v * ct[i * B];
    }
#pragma unroll
    for (int i = 0; i < LZ_KEEP_MAX; ++i)
        if (i < no) V[static_cast<int64_t>(i) * vs + idx] = static_cast<XT>(acc[i]);
    double last = 0.0;
#pragma unroll
    for (int i = 0; i < LZ_KEEP_MAX; ++i)
        if (i == no - 1) last = acc[i];
    Xc[idx] = static_cast<XT>(last);
}

// Ritz vectors: out[i][r][b] = sum_{t < nj[b]} V[t][r][b] * C[(t * nv + i) * B + b]   (C includes scale_t),
// for the group of outputs i0 <= i < i0 + ng (ng <= LZ_RZ_G; the host loops over groups).  Each run stops
// at its own step count: Krylov slots past it were never written for that run.  The coefficients of a run
// of slots are staged in LDS once per workgroup (they are shared by all rows), and each thread carries
// LZ_RZ_R elements, so that the Krylov basis is streamed with several independent loads in flight; the
// accumulators have compile-time indices only, so they stay in registers.
constexpr int LZ_RZ_R = 4;       // elements per thread (strided by the workgroup size)
constexpr int LZ_RZ_G = 8;       // outputs per launch
constexpr int LZ_RZ_LDS = 4096;  // coefficients staged per pass (32 KB)
constexpr int LZ_RZ_T = 4;       // Krylov slots loaded together

template <typename XT, typename OT, int B>
__global__ void __launch_bounds__(256) k_lz_ritz(const XT* __restrict__ V, int64_t vs, const int* __restrict__ njb, int njmax,
                                                 const double* __restrict__ C, int nv, int i0, int ng, OT* __restrict__ out,
                                                 int64_t os) {
    __shared__ double cs[LZ_RZ_LDS];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * (256 * LZ_RZ_R) + threadIdx.x;
    const int b = static_cast<int>(base % B);  // the same lane for all of this thread's elements (256 % B == 0)
    const int nj = njb[b];
    double acc[LZ_RZ_R][LZ_RZ_G];
#pragma unroll
    for (int u = 0; u < LZ_RZ_R; ++u)
#pragma unroll
        for (int i = 0; i < LZ_RZ_G; ++i) acc[u][i] = 0.0;
    const int per_t = LZ_RZ_G * B;  // staged coefficients per slot: the group's outputs for every lane
    const int tc = max(1, LZ_RZ_LDS / per_t);
    for (int t0 = 0; t0 < njmax; t0 += tc) {
        const int t1 = min(t0 + tc, njmax);
        __syncthreads();
        for (int q = threadIdx.x; q < (t1 - t0) * per_t; q += 256) {
            const int tt = q / per_t, rem = q % per_t, i = rem / B, bb = rem % B;
            cs[q] = i < ng ? C[(static_cast<int64_t>(t0 + tt) * nv + i0 + i) * B + bb] : 0.0;
        }
        __syncthreads();
        const int te = min(t1, nj);
        int t = t0;
        // LZ_RZ_T slots per iteration: all their loads are issued before the multiply-adds (the sums still
        // run over t in order)
        for (; t + LZ_RZ_T <= te; t += LZ_RZ_T) {
            XT v[LZ_RZ_T][LZ_RZ_R];
#pragma unroll
            for (int q = 0; q < LZ_RZ_T; ++q)
#pragma unroll
                for (int u = 0; u < LZ_RZ_R; ++u) {
                    // unconditional loads of a valid element (selected to zero past the end), so that all of them
                    // are in flight together
                    const int64_t idx = base + u * 256;
                    const XT x = V[static_cast<int64_t>(t + q) * vs + min(idx, vs - 1)];
                    v[q][u] = idx < vs ? x : XT(0);
                }
#pragma unroll
            for (int q = 0; q < LZ_RZ_T; ++q) {
                const double* ct = cs + (t + q - t0) * per_t + b;
#pragma unroll
                for (int i = 0; i < LZ_RZ_G; ++i) {
                    const double c = ct[i * B];
#pragma unroll
                    for (int u = 0; u < LZ_RZ_R; ++u) acc[u][i] += to_d(v[q][u]) * c;
                }
            }
        }
        for (; t < te; ++t) {
            double v[LZ_RZ_R];
#pragma unroll
            for (int u = 0; u < LZ_RZ_R; ++u) {
                const int64_t idx = base + u * 256;
                const XT x = V[static_cast<int64_t>(t) * vs + min(idx, vs - 1)];
                v[u] = idx < vs ? to_d(x) : 0.0;
            }
            const double* ct = cs + (t - t0) * per_t + b;
#pragma unroll
            for (int i = 0; i < LZ_RZ_G; ++i) {
                const double c = ct[i * B];
#pragma unroll
                for (int u = 0; u < LZ_RZ_R; ++u) acc[u][i] += v[u] * c;
            }
        }
    }
#pragma unroll
    for (int u = 0; u < LZ_RZ_R; ++u) {
        const int64_t idx = base + u * 256;
#pragma unroll
        for (int i = 0; i < LZ_RZ_G; ++i)
            if (i < ng && idx < vs) out[static_cast<int64_t>(i0 + i) * os + idx] = static_cast<OT>(acc[u][i]);
    }
}

// Y[dst[i][b]][r][b] = src[i][r][b] for the new locked vectors (dst -1: none)
template <typename XT, int B>
__global__ void __launch_bounds__(256) k_lz_put_locked(const XT* __restrict__ src, XT* __restrict__ Y, int64_t vs, int nv,
                                                       const int* __restrict__ dst) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    for (int i = 0; i < nv; ++i) {
        const int d = dst[i * B + b];
        if (d >= 0) Y[static_cast<int64_t>(d) * vs + idx] = src[static_cast<int64_t>(i) * vs + idx];
    }
}

// Xc[r][b] = V[slot[b]][r][b] for the lanes rolled back to step slot[b] (slot -1: lane untouched)
template <typename XT, int B>
__global__ void __launch_bounds__(256) k_lz_restore_current(const XT* __restrict__ V, XT* __restrict__ Xc, int64_t vs,
                                                            const int* __restrict__ slot) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= vs) return;
    const int d = slot[idx % B];
    if (d >= 0) Xc[idx] = V[static_cast<int64_t>(d) * vs + idx];
}

// per-(i, b) partial sums over rows of [nv][D][B] arrays (reduced by k_lz_reduce mode 0):
//   mode 0: value 2i -> <y,y>, 2i+1 -> <y,Hy>;  mode 1: value i -> ||Hy - theta y||^2;  mode 2: i -> <y,y>
template <typename YT, int B>
__global__ void __launch_bounds__(LZ_WG) k_lz_dots(const YT* __restrict__ Y, const double* __restrict__ HY, int64_t D, int nv,
                                                   int mode, const double* __restrict__ theta, double* part) {
    __shared__ double lds[LZ_WG / 64 * B];
    const int b = threadIdx.x % B;
    constexpr int ROWS = LZ_WG / B;
    const int64_t vs = D * B;
    const int64_t ntile = (D + ROWS - 1) / ROWS;
    const int64_t per = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t t0 = min(static_cast<int64_t>(blockIdx.x) * per, ntile), t1 = min(t0 + per, ntile);
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
}

// Rayleigh-Ritz of the finished candidates, row i of the Gram matrices: per lane b, partials of
// value q < nv: <Y_i, Y_q> and value nv + q: <Y_i, HY_q> (reduced by k_lz_reduce mode 0)
// Rows i0 .. i0 + NI - 1 of the Gram matrices of the finish, per lane: G(i, q) = <y_i, y_q> and
// A(i, q) = <y_i, H y_q> for q < nv <= NV, as per-workgroup partials of the values ((i - i0) 2 + w) NV + q
// (w = 0: G, 1: A); each thread reads its NI row elements once and every column element once.
template <int B, int NV, int NI>
__global__ void __launch_bounds__(LZ_WG) k_lz_gram_rows(const double* __restrict__ Y, const double* __restrict__ HY, int64_t D,
                                                        int nv, int i0, double* part) {
    __shared__ double lds[LZ_PUT_LDS];
    const int b = threadIdx.x % B;
    constexpr int ROWS = LZ_WG / B;
    const int64_t vs = D * B;
    const int64_t ntile = (D + ROWS - 1) / ROWS;
    const int64_t per = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t t0 = min(static_cast<int64_t>(blockIdx.x) * per, ntile), t1 = min(t0 + per, ntile);
    double g[NI * 2 * NV];
#pragma unroll
    for (int q = 0; q < NI * 2 * NV; ++q) g[q] = 0.0;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t r = t * ROWS + threadIdx.x / B;
        if (r >= D) continue;
        const int64_t idx = r * B + b;
        double yi[NI];
#pragma unroll
        for (int ii = 0; ii < NI; ++ii) yi[ii] = i0 + ii < nv ? Y[static_cast<int64_t>(i0 + ii) * vs + idx] : 0.0;
#pragma unroll
        for (int q = 0; q < NV; ++q)
            if (q < nv) {
                const double yq = Y[static_cast<int64_t>(q) * vs + idx], hq = HY[static_cast<int64_t>(q) * vs + idx];
#pragma unroll
                for (int ii = 0; ii < NI; ++ii) {
                    g[(ii * 2) * NV + q] += yi[ii] * yq;
                    g[(ii * 2 + 1) * NV + q] += yi[ii] * hq;
                }
            }
    }
    lz_block_put_multi<B, NI * 2 * NV>(g, lds, part, NI * 2 * NV * B);
}

// in place, per lane b: X[c][r][b] = sum_{i < nin} X[i][r][b] Z[(i nout + c) B + b] for c < nout (nout <= nin;
// each thread reads its (r, b) column of X before writing it)
template <int B>
__global__ void __launch_bounds__(256) k_lz_combine(double* X, int64_t vs, const double* __restrict__ Z, int nin, int nout) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    double x[LZ_NVMAX];
#pragma unroll
    for (int i = 0; i < LZ_NVMAX; ++i) x[i] = i < nin ? X[static_cast<int64_t>(i) * vs + idx] : 0.0;
    for (int c = 0; c < nout; ++c) {
        double acc = 0.0;
#pragma unroll
        for (int i = 0; i < LZ_NVMAX; ++i)
            if (i < nin) acc += x[i] * Z[(static_cast<int64_t>(i) * nout + c) * B + b];
        X[static_cast<int64_t>(c) * vs + idx] = acc;
    }
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

// Davidson correction of the FP64 refinement, lanes with on[b]: T[i][r][b] = (HY - theta Y)[i][r][b] / (U_b V_r - theta_ib),
// the denominator held at least dl[b] away from zero (a diagonal entry close to theta is a direction the diagonal
// does not resolve); T = 0 on the other lanes
template <int B>
__global__ void __launch_bounds__(256) k_lz_dav_t(const double* __restrict__ Y, const double* __restrict__ HY, double* __restrict__ T,
                                                  int64_t D, int nv, const double* __restrict__ theta, const double* __restrict__ U,
                                                  const double* __restrict__ dl, const float* __restrict__ diag,
                                                  const int* __restrict__ on) {
    const int64_t vs = D * B;
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs * nv) return;
    const int64_t i = idx / vs, e = idx - i * vs;
    const int b = static_cast<int>(e % B);
    double t = 0.0;
    if (on[b]) {
        const double th = theta[i * B + b], d = dl[b];
        double den = U[b] * static_cast<double>(diag[e / B]) - th;
        if (fabs(den) < d) den = den < 0.0 ? -d : d;
        t = (HY[idx] - th * Y[idx]) / den;
    }
    T[idx] = t;
}

// out-of-place form of k_lz_combine for more than LZ_NVMAX inputs: T[c][r][b] = sum_{i < nin} X[i][r][b] Z[(i nout + c) B + b]
template <int B>
__global__ void __launch_bounds__(256) k_lz_combine_wide(const double* __restrict__ X, double* __restrict__ T, int64_t vs,
                                                         const double* __restrict__ Z, int nin, int nout) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    for (int c = 0; c < nout; ++c) {
        double acc = 0.0;
        for (int i = 0; i < nin; ++i) acc += X[static_cast<int64_t>(i) * vs + idx] * Z[(static_cast<int64_t>(i) * nout + c) * B + b];
        T[static_cast<int64_t>(c) * vs + idx] = acc;
    }
}

// Gram-matrix block of the finish for more than LZ_NVMAX candidates: rows i0 .. i0 + NI - 1 and columns
// q0 .. q0 + NV - 1 of G = Y^T Y and A = Y^T H Y per lane, as per-workgroup partials of the values
// ((i - i0) 2 + w) NV + (q - q0) (w = 0: G, 1: A)
template <int B, int NV, int NI>
__global__ void __launch_bounds__(LZ_WG) k_lz_gram_blk(const double* __restrict__ Y, const double* __restrict__ HY, int64_t D,
                                                       int nv, int i0, int q0, double* part) {
    __shared__ double lds[LZ_PUT_LDS];
    const int b = threadIdx.x % B;
    constexpr int ROWS = LZ_WG / B;
    const int64_t vs = D * B;
    const int64_t ntile = (D + ROWS - 1) / ROWS;
    const int64_t per = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t t0 = min(static_cast<int64_t>(blockIdx.x) * per, ntile), t1 = min(t0 + per, ntile);
    double g[NI * 2 * NV];
#pragma unroll
    for (int q = 0; q < NI * 2 * NV; ++q) g[q] = 0.0;
    for (int64_t t = t0; t < t1; ++t) {
        const int64_t r = t * ROWS + threadIdx.x / B;
        if (r >= D) continue;
        const int64_t idx = r * B + b;
        double yi[NI];
#pragma unroll
        for (int ii = 0; ii < NI; ++ii) yi[ii] = i0 + ii < nv ? Y[static_cast<int64_t>(i0 + ii) * vs + idx] : 0.0;
#pragma unroll
        for (int q = 0; q < NV; ++q)
            if (q0 + q < nv) {
                const double yq = Y[static_cast<int64_t>(q0 + q) * vs + idx], hq = HY[static_cast<int64_t>(q0 + q) * vs + idx];
#pragma unroll
                for (int ii = 0; ii < NI; ++ii) {
                    g[(ii * 2) * NV + q] += yi[ii] * yq;
                    g[(ii * 2 + 1) * NV + q] += yi[ii] * hq;
                }
            }
    }
    lz_block_put_multi<B, NI * 2 * NV>(g, lds, part, NI * 2 * NV * B);
}

// ---- deflation passes ---------------------------------------------------------------------------------------
// A lane that has converged its pass freezes the levels it found (their Ritz vectors go to Y after the vectors
// frozen before) and restarts from a fresh start vector orthogonal to every frozen vector; the kernels below
// form that vector and keep frozen sets orthonormal.  cnt[b] < 0 leaves lane b untouched.

// Z[r][b] = x_{off + i(r) + 1} / (2^31 - 1) - 0.5 for lane b, where x_i = 16807^i x_0 mod (2^31 - 1): the
// Park-Miller sequence of the start vector continued past its first `off` values (jumped to by modular powers);
// i(r) = r, or lex_of_pos[r] when given (the caller-chosen orders, as for the start vector)
template <int B>
__global__ void __launch_bounds__(256) k_lz_pm_fill(double* __restrict__ Z, int64_t D, int b, uint64_t x0, uint64_t off,
                                                    const int32_t* __restrict__ lex_of_pos) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= D) return;
    const uint64_t p = 2147483647ull;
    const uint64_t ir = lex_of_pos ? static_cast<uint64_t>(lex_of_pos[r]) : static_cast<uint64_t>(r);
    uint64_t e = off + ir + 1, base = 16807ull, pw = 1;
    while (e) {
        if (e & 1) pw = (pw * base) % p;
        base = (base * base) % p;
        e >>= 1;
    }
    const uint64_t x = (pw * x0) % p;
    Z[r * B + b] = static_cast<double>(x) / static_cast<double>(p) - 0.5;
}

// partials of <Y_l, Z> over Y's slots l < cnt[b] (values l B + b of rows of width lmax B)
template <typename XT, int B>
__global__ void __launch_bounds__(LZ_WG) k_lz_ydots(const XT* __restrict__ Y, const double* __restrict__ Z, int64_t D,
                                                    const int* __restrict__ cnt, int lmax, double* part) {
    __shared__ double lds[LZ_WG / 64 * B];
    const int b = threadIdx.x % B;
    constexpr int ROWS = LZ_WG / B;
    const int64_t vs = D * B;
    const int64_t ntile = (D + ROWS - 1) / ROWS;
    const int64_t per = (ntile + gridDim.x - 1) / gridDim.x;
    const int64_t t0 = min(static_cast<int64_t>(blockIdx.x) * per, ntile), t1 = min(t0 + per, ntile);
    const int nl = cnt[b];
    for (int l = 0; l < lmax; ++l) {
        double s = 0.0;
        if (l < nl) {
            for (int64_t t = t0; t < t1; ++t) {
                const int64_t r = t * ROWS + threadIdx.x / B;
                if (r < D) s += to_d(Y[static_cast<int64_t>(l) * vs + r * B + b]) * Z[r * B + b];
            }
        }
        lz_block_put<B>(s, l, lds, part, lmax * B);
    }
}

// Z -= sum_{l < cnt[b]} c[l B + b] Y_l
template <typename XT, int B>
__global__ void __launch_bounds__(256) k_lz_yaxpy(const XT* __restrict__ Y, double* __restrict__ Z, int64_t vs,
                                                  const int* __restrict__ cnt, const double* __restrict__ c) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    const int nl = cnt[b];
    if (nl <= 0) return;
    double z = Z[idx];
    for (int l = 0; l < nl; ++l) z -= c[l * B + b] * to_d(Y[static_cast<int64_t>(l) * vs + idx]);
    Z[idx] = z;
}

// base[slot[b]][r][b] = Z[r][b] f[b] in storage precision (and the lane's current vector when xc is given)
template <typename XT, int B>
__global__ void __launch_bounds__(256) k_lz_defl_store(const double* __restrict__ Z, XT* base, int64_t vs,
                                                       const int* __restrict__ slot, const double* __restrict__ f, XT* xc) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    const int s = slot[b];
    if (s < 0) return;
    const XT x = static_cast<XT>(Z[idx] * f[b]);
    base[static_cast<int64_t>(s) * vs + idx] = x;
    if (xc) xc[idx] = x;
}

// Y[s][r][b] = Y[src[s B + b]][r][b] for s < ns (src -1: slot untouched; src[s] >= s and ascending per lane, so the
// moves in ascending s never read a slot an earlier move wrote)
template <typename XT, int B>
__global__ void __launch_bounds__(256) k_lz_compact(XT* Y, int64_t vs, const int* __restrict__ src, int ns) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int b = static_cast<int>(idx % B);
    for (int q = 0; q < ns; ++q) {
        const int a = src[q * B + b];
        if (a >= 0 && a != q) Y[static_cast<int64_t>(q) * vs + idx] = Y[static_cast<int64_t>(a) * vs + idx];
    }
}

// Yf[i][r][b] = Y[i][r][b] for i < nf[b] (the frozen vectors among the finish's candidates, in FP64)
template <typename XT, int B>
__global__ void __launch_bounds__(256) k_lz_copy_frozen(const XT* __restrict__ Y, double* __restrict__ Yf, int64_t vs,
                                                        const int* __restrict__ nf, int nfmax) {
    const int64_t idx = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= vs) return;
    const int n = nf[idx % B];
    for (int i = 0; i < nfmax && i < n; ++i) Yf[static_cast<int64_t>(i) * vs + idx] = to_d(Y[static_cast<int64_t>(i) * vs + idx]);
}

// ---- host driver: the Lanczos engine --------------------------------------------------------------------
//
// An engine holds B lanes over one Hamiltonian.  A lane runs one solve (J_b, U_b); lanes are admitted
// and retired independently (each keeps its own step counter and Krylov column), so a serving loop can
// refill a lane as soon as its solve finishes while the others keep iterating.

namespace {

int lz_debug() {
    static int d = [] {
        const char* e = std::getenv("BHX_DEBUG");
        return e ? std::atoi(e) : 0;
    }();
    return d;
}

// Orthogonalisation against the locked vectors: in every step (default), or only in the steps the
// Parlett-Scott estimate asks for (BHX_SO=auto).  With FP32-stored vectors and the flag shared by all
// lanes of an engine, the estimate-driven schedule fires in most steps of a batch anyway and leaves the
// basis only semi-orthogonal (sqrt(eps32)) to converged levels, which costs accuracy in the excited
// levels; every-step orthogonalisation keeps it at eps32.
int lz_so_always() {
    static int d = [] {
        const char* e = std::getenv("BHX_SO");
        return (e && std::strcmp(e, "auto") == 0) ? 0 : 1;
    }();
    return d;
}

// Davidson steps of the FP64 refinement that a mixed lane failing its acceptance test gets in the finish
// before it is reported (and rerun in FP64): BHX_REFINE=<n>, 0 turns the refinement off.
int lz_refine_steps() {
    static int d = [] {
        const char* e = std::getenv("BHX_REFINE");
        return e ? std::max(0, std::atoi(e)) : LZ_REFINE_STEPS;
    }();
    return d;
}

// Steps enqueued at an analysis before the host analyses (the device runs them meanwhile): one captured
// graph's worth by default, BHX_SPEC=0 turns the overlap off.
int lz_spec_steps() {
    static int d = [] {
        const char* e = std::getenv("BHX_SPEC");
        return (e && std::strcmp(e, "0") == 0) ? 0 : LZ_GRAPH_STEPS;
    }();
    return d;
}

// start vector of the solver (Park-Miller sequence x <- 16807 x mod (2^31 - 1), seed 0 -> 1, values
// x / (2^31 - 1) - 0.5), shared by every lane: position q takes value number q + 1 of the sequence, or, in the
// caller-chosen orders (the Hamiltonian keeps lex_of_pos), value number lex(q) + 1 for the lexicographic rank
// lex(q) of its state
int lz_start_vector(bh_ctx* ctx, const bh_ham* h, uint64_t seed, std::vector<double>& r) {
    const int64_t D = h->D;
    r.assign(static_cast<size_t>(D), 0.0);
    const uint64_t p = 2147483647ull;
    uint64_t x = seed ? (seed & p) : 1;
    for (int64_t i = 0; i < D; ++i) {
        x = (16807ull * x) % p;
        r[i] = static_cast<double>(x) / static_cast<double>(p) - 0.5;
    }
    if (h->lex_of_pos) {
        std::vector<int32_t> lop(static_cast<size_t>(D));
        BH_HIP(ctx, hipMemcpyAsync(lop.data(), h->lex_of_pos, lop.size() * sizeof(int32_t), hipMemcpyDeviceToHost,
                                   ctx->stream));
        BH_HIP(ctx, hipStreamSynchronize(ctx->stream));
        std::vector<double> q(static_cast<size_t>(D));
        for (int64_t i = 0; i < D; ++i) q[i] = r[lop[i]];
        r.swap(q);
    }
    return BH_OK;
}

struct RunHost {
    std::vector<double> alpha, beta;     // alpha[t], beta[t] (beta[0] = 0)
    std::vector<double> lock_theta;      // locked Ritz values
    std::vector<double> lock_bound;      // their bounds when they were (last) formed
    bool done = false;
    int iters = 0;
    int status = BH_ENOCONV;
    std::vector<double> theta;           // wanted levels and extra candidates (Ritz values of T)
    std::vector<std::vector<double>> s;  // their eigenvectors of T
    double tnorm = 0.0;
    double tnorm0 = 0.0;                 // ||T|| at the first analysis (divergence guard)
    bool exhausted = false;              // the run ended on an invariant Krylov space
    // convergence history for the check schedule: max over the unconverged wanted levels of
    // log(bound / threshold) at the previous analysis, and the lane's total steps then (-1: none)
    double prev_gap = 0.0;
    int prev_tot = -1;
    // deflation passes (eigenvalues with multiplicity): passes completed before the current one, and the
    // vectors they froze (Y slots 0 .. nfz - 1 of the lane) with their Ritz values
    int pass = 0;
    int nfz = 0;
    std::vector<double> fz_theta;
    int refined = 0;  // Davidson steps of the FP64 refinement in the finish
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

// Krylov slots of a lane: the largest Lanczos matrix it builds before a thick restart (ncv, at most D
// and at most max_iter) plus one slot for the next Lanczos vector
static int lz_slots(int64_t D, const bh_lanczos_opts& o) {
    const int maxsteps = o.max_iter > 0 ? o.max_iter : 10000;
    int ncv = o.ncv > 0 ? o.ncv : 512;
    ncv = std::max(ncv, 2 * o.k + 8);
    int64_t cap = std::min<int64_t>(ncv, D);
    cap = std::min<int64_t>(cap, maxsteps);
    return static_cast<int>(std::max<int64_t>(cap, 1)) + 1;
}

// Sectors up to this dimension run FP64 Lanczos with full reorthogonalisation: their solves are bound by
// launch latency, not bandwidth, so FP32 storage buys nothing there, and reorthogonalising against the whole
// basis (j D B words per step) is cheap.
bool bh_lz_fullre(int64_t D, const bh_lanczos_opts& o) {
    const int64_t lim = o.fullre_d > 0 ? o.fullre_d : (o.fullre_d < 0 ? 0 : 16384);
    return D <= lim || (!o.distinct && o.k > LZ_SO_KMAX);
}

// Y slots of a lane: locked and frozen vectors under selective orthogonalisation; frozen vectors only (up to
// 2k + 8: a pass freezes only values among the lowest k) under full reorthogonalisation
static int lz_ymax(int64_t D, const bh_lanczos_opts& o) {
    if (!bh_lz_fullre(D, o)) return LZ_LMAX;
    return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(D, std::max(LZ_LMAX, 2 * o.k + 8))));
}

template <typename XT, int B>
class LzEngine final : public LzEngineBase {
    static constexpr bool MIXED = std::is_same<XT, float>::value;
    // Both precisions read the packed words: the FP64 table gives the exact matrix values (4 bytes per entry
    // instead of the 12 of an FP64 CSR), and the diagonal V (small integers) is exact in FP32.
    static constexpr int FMT = BH_FMT_PACKED;

   public:
    LzEngine(bh_ctx* c, bh_ham* ham, const bh_lanczos_opts& opts, hipStream_t stream, bool use_ctx_ws)
        : ctx(c), h(ham), s(stream), o(opts), ws(use_ctx_ws) {}

    int setup() {
        D = h->D;
        vs = D * B;
        k = o.k;
        maxsteps = o.max_iter > 0 ? o.max_iter : 10000;
        slots = lz_slots(D, o);
        cap = slots - 1;
        fullre = bh_lz_fullre(D, o);
        mult = !o.distinct;
        ymax = lz_ymax(D, o);
        {
            // Step graphs are off under rocprofv3 (it exports ROCPROF_OUTPUT_PATH to the application): its kernel
            // tracing of graph launches ended two runs with a host SIGSEGV inside hipGraphLaunch.  BHX_GRAPH=0/1
            // forces them off/on.
            const char* e = std::getenv("BHX_GRAPH");
            use_graphs = e ? std::strcmp(e, "0") != 0 : std::getenv("ROCPROF_OUTPUT_PATH") == nullptr;
        }
        chunk = o.chunk > 0 ? o.chunk : 16;
        tol = o.tol > 0 ? o.tol : 1e-10;
        dedup = o.dedup_rel > 0 ? o.dedup_rel : 1e-9;
        eps_store = MIXED ? 5.96e-8 : 1.11e-16;
        eps23 = std::pow(2.220446049250313e-16, 2.0 / 3.0);
        const size_t nV = static_cast<size_t>(slots) * vs, nY = static_cast<size_t>(ymax) * vs;
        if (ws) {
            void* q = nullptr;
            BH_TRY(bh_ws_get(ctx, 0, nV * sizeof(XT), &q));
            V = static_cast<XT*>(q);
            BH_TRY(bh_ws_get(ctx, 1, static_cast<size_t>(vs) * sizeof(double), &q));
            W = static_cast<double*>(q);
            BH_TRY(bh_ws_get(ctx, 2, nY * sizeof(XT), &q));
            Y = static_cast<XT*>(q);
            BH_TRY(bh_ws_get(ctx, 3, static_cast<size_t>(D) * sizeof(double), &q));
            r0 = static_cast<double*>(q);
            BH_TRY(bh_ws_get(ctx, 4, static_cast<size_t>(vs) * sizeof(XT), &q));
            Xc = static_cast<XT*>(q);
        } else {
            BH_TRY(dalloc(ctx, V_own, nV));
            BH_TRY(dalloc(ctx, W_own, static_cast<size_t>(vs)));
            BH_TRY(dalloc(ctx, Y_own, nY));
            BH_TRY(dalloc(ctx, r0_own, static_cast<size_t>(D)));
            BH_TRY(dalloc(ctx, Xc_own, static_cast<size_t>(vs)));
            V = V_own.p;
            W = W_own.p;
            Y = Y_own.p;
            r0 = r0_own.p;
            Xc = Xc_own.p;
        }
        n_alpha = static_cast<size_t>(slots) * B;
        n_beta = static_cast<size_t>(slots + 1) * B;
        const size_t n_sums = static_cast<size_t>(1 + 2 * LZ_LMAX) * B;
        const size_t n_tau = static_cast<size_t>(2 * LZ_LMAX) * B, n_lth = static_cast<size_t>(LZ_LMAX) * B;
        const size_t nd = n_alpha + 2 * n_beta + 2 * B + n_sums + n_tau + n_lth + B + 3 * n_lth;
        BH_TRY(dalloc(ctx, st_d, nd));
        BH_TRY(dalloc(ctx, st_i, static_cast<size_t>(3 * B + 2)));
        const int nvmax = std::max({static_cast<int>(n_sums), 2 * LZ_NVMAX * B, ymax * B});
        // K1 on the row-lane engine: small row tiles (LZ_K1_TILES row groups of LZ_WG / B rows), many
        // workgroups in flight; its partial sums need a row per workgroup
        rpw = std::max<int64_t>(static_cast<int64_t>(LZ_WG / B) * LZ_K1_TILES, (D + LZ_K1_GRID_MAX - 1) / LZ_K1_GRID_MAX);
        grid_rl = static_cast<int>((D + rpw - 1) / rpw);
        // K1 on the staged row-block engine: a workgroup per block up to LZ_K1_GRID_MAX workgroups, within
        // LZ_PART_BYTES of partial sums
        staged_ok = static_cast<int64_t>(h->max_row) * B <= BH_STAGE_PRODUCTS_HOST;
        if (staged_ok) BH_TRY(bh_ham_rowblocks(h, B, &rblk, &nrblk));
        const int64_t rb_cap = std::max<int64_t>(grid_rl, LZ_PART_BYTES / (static_cast<int64_t>(nvmax) * 8));
        grid_rb = static_cast<int>(std::min<int64_t>({std::max<int64_t>(nrblk, 1), LZ_K1_GRID_MAX, rb_cap}));
        // partial-sum rows: K1 and R2 use up to max(grid) rows of nvmax values; the launches over grid_k2
        // workgroups of the finish and the deflation write up to LZ_GRAM_VALS * B values per row (Gram tiles)
        {
            const int64_t ntb = (D + LZ_WG / B - 1) / (LZ_WG / B);
            const size_t g2 = static_cast<size_t>(std::min<int64_t>(ntb, LZ_K2_GRID));
            const size_t n1 = static_cast<size_t>(std::max({LZ_GRID_MAX, LZ_K2_GRID, grid_rl, grid_rb})) * nvmax;
            const size_t n2 = g2 * static_cast<size_t>(std::max(LZ_GRAM_VALS, ymax)) * B;
            BH_TRY(dalloc(ctx, part, std::max(n1, n2)));
        }
        BH_TRY(dalloc(ctx, dnorm, static_cast<size_t>(2 * std::max(LZ_CMAX, ymax) * B)));
        BH_TRY(dalloc(ctx, sums2, static_cast<size_t>(LZ_BETA_MAXV)));
        BH_TRY(dalloc(ctx, dnj, B));
        BH_TRY(dalloc(ctx, lock_map, static_cast<size_t>(std::max(LZ_CMAX, ymax)) * B));
        BH_TRY(dalloc(ctx, dno, B));
        if (fullre) BH_TRY(dalloc(ctx, dd, static_cast<size_t>(slots + ymax) * B));
        st.alpha = st_d.p;
        st.beta = st.alpha + n_alpha;
        st.scale = st.beta + n_beta;
        st.J = st.scale + n_beta;
        st.U = st.J + B;
        st.sums = st.U + B;
        st.tau = st.sums + n_sums;
        st.ltheta = st.tau + n_tau;
        st.anorm = st.ltheta + n_lth;
        st.dso = st.anorm + B;
        st.active = st_i.p;
        st.nlock = st_i.p + B;
        st.step = st_i.p + 2 * B;
        st.so = st_i.p + 3 * B;
        st.part = part.p;
        BH_HIP(ctx, hipMemsetAsync(st_d.p, 0, nd * sizeof(double), s));
        BH_HIP(ctx, hipMemsetAsync(st_i.p, 0, st_i.bytes(), s));
        // start vector and the inverse norm of its stored copy (cached in the context workspace per
        // (D, seed, storage precision) when the workspace is used)
        const uint64_t layout = h->lex_of_pos ? h->layout_id : 0;
        if (!ws || ctx->r0_D != D || ctx->r0_seed != o.seed || ctx->r0_mixed != static_cast<int>(MIXED) ||
            ctx->r0_layout != layout) {
            std::vector<double> r;
            BH_TRY(lz_start_vector(ctx, h, o.seed, r));
            BH_HIP(ctx, hipMemcpyAsync(r0, r.data(), D * 8, hipMemcpyHostToDevice, s));
            double nrm = 0.0;
            for (int64_t i = 0; i < D; ++i) {
                const double v = static_cast<double>(static_cast<XT>(r[i]));
                nrm += v * v;
            }
            scale0 = nrm > 0 ? 1.0 / std::sqrt(nrm) : 0.0;
            BH_HIP(ctx, hipStreamSynchronize(s));
            if (ws) {
                ctx->r0_D = D;
                ctx->r0_seed = o.seed;
                ctx->r0_mixed = static_cast<int>(MIXED);
                ctx->r0_layout = layout;
                ctx->r0_scale = scale0;
            }
        } else {
            scale0 = ctx->r0_scale;
        }
        a = LzArgs<XT>{D,
                       h->row_ptr,
                       h->words,
                       nullptr,
                       nullptr,
                       h->diag,
                       nullptr,
                       h->table,
                       h->code_bits,
                       V,
                       Xc,
                       vs,
                       reinterpret_cast<XT*>(W),  // the FP64 buffer holds the narrower W of the steps
                       Y,
                       cap,
                       st};
        shm = FMT == BH_FMT_PACKED ? (static_cast<size_t>(1) << h->code_bits) * sizeof(double) : 0;
        // K1 variant before the tuner has run: staged row blocks, or one thread per (row, b)
        shm_rb = shm + BH_STAGE_PRODUCTS * sizeof(double);
        set_k1(bh_engine_choice(B, h) == 1 && staged_ok ? 0 : 1);
        const int64_t ntile_b = (D + LZ_WG / B - 1) / (LZ_WG / B);
        grid_k2 = static_cast<int>(std::min<int64_t>(ntile_b, LZ_K2_GRID));
        rh.assign(B, RunHost());
        nextc.assign(B, 0);
        act.assign(B, 0);
        waiting.assign(B, 0);
        nlock.assign(B, 0);
        jl.assign(B, 0);
        jan.assign(B, 0);
        tot_an.assign(B, 0);
        tot.assign(B, 0);
        kept.assign(B, 0);
        Jb.assign(B, 0.0);
        Ub.assign(B, 0.0);
        hlo.assign(B, 0.0);
        hhi.assign(B, 0.0);
        for (auto& R : rh) R.done = true;
        bytes = st_d.bytes() + st_i.bytes() + part.bytes() + dnorm.bytes() +
                (ws ? 0 : V_own.bytes() + W_own.bytes() + Y_own.bytes() + r0_own.bytes() + Xc_own.bytes());
        return BH_OK;
    }

    int lanes() const override { return B; }
    void set_deflation_barrier(bool on) override { defl_barrier = on; }
    int export_tridiag(int steps, double* alpha, double* beta, double* basis) override {
        if (steps < 1 || steps > cap) return bh_set_error(ctx, BH_EINVAL, "tridiag: steps beyond the Krylov slots");
        std::vector<double> ha2(static_cast<size_t>(steps) * B), hb2(static_cast<size_t>(steps + 1) * B),
            hs2(static_cast<size_t>(steps) * B);
        BH_HIP(ctx, hipMemcpyAsync(ha2.data(), st.alpha, ha2.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hb2.data(), st.beta, hb2.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hs2.data(), st.scale, hs2.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (int t = 0; t < steps; ++t) alpha[t] = ha2[static_cast<size_t>(t) * B];
        beta[0] = 0.0;
        for (int t = 1; t <= steps; ++t) beta[t] = hb2[static_cast<size_t>(t) * B];
        if (basis) {
            std::vector<XT> col(static_cast<size_t>(vs));
            for (int t = 0; t < steps; ++t) {
                BH_HIP(ctx, hipMemcpyAsync(col.data(), V + static_cast<int64_t>(t) * vs, vs * sizeof(XT), hipMemcpyDeviceToHost, s));
                BH_HIP(ctx, hipStreamSynchronize(s));
                const double sc = hs2[static_cast<size_t>(t) * B];
                for (int64_t r = 0; r < D; ++r) basis[static_cast<size_t>(t) * D + r] = static_cast<double>(col[r * B]) * sc;
            }
        }
        return BH_OK;
    }
    int active() const override {
        int c = 0;
        for (int b = 0; b < B; ++b) c += act[b];
        return c;
    }
    int steps_of(int slot) const override { return tot[slot]; }
    size_t device_bytes() const override { return bytes; }
    bool fp64() const override { return !MIXED; }
    // Steps until the next lane reaches its own analysis point.  Each lane follows its own schedule
    // (first at max(2k+8, 20) steps, then every max(chunk, j/8); always when its Krylov slots are full
    // and when its step budget is spent), so the steps at which a lane is analysed, locked, restarted
    // and retired do not depend on the other lanes.
    int suggest_chunk() const override {
        int st = -1;
        for (int b = 0; b < B; ++b)
            if (act[b]) {
                // with speculative steps a lane may already be past its next analysis: check at once
                const int d = std::max(speculate() ? 0 : 1, next_event(b) - jl[b]);
                st = st < 0 ? d : std::min(st, d);
            }
        return st < 0 ? chunk : st;
    }

    int admit(int slot, double J, double U) override {
        if (slot < 0 || slot >= B || act[slot] || waiting[slot])
            return bh_set_error(ctx, BH_EINVAL, "engine: lane busy or out of range");
        rh[slot] = RunHost();
        act[slot] = 1;
        waiting[slot] = 0;
        nlock[slot] = 0;
        jl[slot] = 0;
        tot[slot] = 0;
        kept[slot] = 0;
        nextc[slot] = std::max(2 * k + 8, 20);
        Jb[slot] = J;
        Ub[slot] = U;
        const double va = U * h->vmin, vb = U * h->vmax;
        hlo[slot] = std::fmin(va, vb) - std::fabs(J) * h->tmax;
        hhi[slot] = std::fmax(va, vb) + std::fabs(J) * h->tmax;
        an0[slot] = std::fmax(std::fabs(hlo[slot]), std::fabs(hhi[slot]));
        const int one = 1, zero = 0;
        const double dz = 0.0;
        BH_HIP(ctx, hipMemcpyAsync(st.J + slot, &Jb[slot], 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.U + slot, &Ub[slot], 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.scale + slot, &scale0, 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.beta + slot, &dz, 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.anorm + slot, &an0[slot], 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.nlock + slot, &zero, 4, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.step + slot, &zero, 4, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.active + slot, &one, 4, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_lz_admit<XT, B>), dim3(static_cast<unsigned>((D + 255) / 256)), dim3(256), 0, s, V, Xc, r0, D,
                           slot);
        BH_HIP(ctx, hipGetLastError());
        BH_HIP(ctx, hipStreamSynchronize(s));  // the sources of the copies are host locals/members
        return BH_OK;
    }

    // Work may still be queued on the engine's stream (the speculative steps a check enqueues before it
    // retires lanes): the stream is drained before the graph executables and the buffers those launches
    // use are released.
    ~LzEngine() override {
        (void)hipStreamSynchronize(s);
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
    }

    // Steps are enqueued as replays of captured HIP graphs of LZ_GRAPH_STEPS and of 1 step when kernel
    // timing is off (BHX_GRAPH=0 disables them): every quantity a step reads -- each lane's step counter, its
    // scalars, the orthogonalisation flag -- lives in device memory, so one captured step is valid for any
    // step of any lane.  With timing on, the launches are issued one by one so that each carries its events.
    int run(int steps) override {
        if (!tuned) BH_TRY(tune_k1());
        const int LM = step_lm();
        const int nsteps = steps;
        if (use_graphs && !ctx->timing) {
            while (steps >= LZ_GRAPH_STEPS) {
                BH_TRY(replay(LM, LZ_GRAPH_STEPS));
                steps -= LZ_GRAPH_STEPS;
            }
            for (; steps > 0; --steps) BH_TRY(replay(LM, 1));
        } else {
            for (int t = 0; t < steps; ++t) step(LM);
        }
        BH_HIP(ctx, hipGetLastError());
        for (int b = 0; b < B; ++b)
            if (act[b]) {
                const int d = std::min(nsteps, cap - jl[b]);
                jl[b] += d;
                tot[b] += d;
            }
        return BH_OK;
    }

    int check(std::vector<int>& finished) override {
        finished.clear();
        int jmax = 0;
        for (int b = 0; b < B; ++b)
            if (act[b]) jmax = std::max(jmax, jl[b]);
        if (jmax == 0) return BH_OK;
        ha.resize(static_cast<size_t>(jmax) * B);
        hb.resize(static_cast<size_t>(jmax + 1) * B);
        const auto tw0 = std::chrono::steady_clock::now();
        BH_HIP(ctx, hipMemcpyAsync(ha.data(), st.alpha, ha.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hb.data(), st.beta, hb.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        const auto tw1 = std::chrono::steady_clock::now();
        // The device is idle from here on.  Steps for the active lanes are enqueued before the analysis so
        // that the device works while the host analyses.  Each lane is analysed at its own scheduled step
        // counts up to the synced ones (the coefficients are all there), so its analysis points do not
        // depend on when the checks happen.  A lane that a decision changes at an earlier step than the one
        // it has reached -- a lock -- is rolled back to that step first (its step counter and current
        // vector; the steps after it only wrote later basis slots); a lane that retires keeps its first
        // basis vectors untouched; a lane whose basis is full runs no speculative step.
        jsync = jl;
        totsync = tot;
        if (speculate()) BH_TRY(run(lz_spec_steps()));
        struct HostClock {  // adds the host time of this analysis to the engine's profile on return
            LzEngine* e;
            std::chrono::steady_clock::time_point t;
            ~HostClock() { e->prof.host_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count(); }
        } hc{this, tw1};
        prof.wait_ms += std::chrono::duration<double, std::milli>(tw1 - tw0).count();
        prof.checks += 1;
        std::vector<std::vector<int>> new_locks(B), lock_slot(B);  // Ritz index, Y slot (-1: append)
        std::vector<Analysis> ans(B);
        std::vector<int> done_now(B, 0), restart_now(B, 0), deflate_now(B, 0);
        // analysis of lane b at step j (total steps tj); true when it retires or restarts the lane
        auto analyse_at = [&](int b, int j, int tj) -> bool {
            RunHost& R = rh[b];
            nextc[b] = j + std::max(chunk, j / 8);
            R.alpha.assign(j, 0.0);
            R.beta.assign(j + 1, 0.0);
            for (int t = 0; t < j; ++t) R.alpha[t] = ha[static_cast<size_t>(t) * B + b];
            for (int t = 0; t <= j; ++t) R.beta[t] = hb[static_cast<size_t>(t) * B + b];
            // An invariant Krylov space shows as a beta at the rounding level of the stored vectors.  It
            // may have happened anywhere since the last check; the steps after it restart from rounding
            // noise and would duplicate every level, so T is cut at the first such beta.  The couplings
            // inside the block kept by a thick restart may be exactly zero and are not searched.
            int jr = j;
            {
                double amax = 0.0;
                for (int t = 0; t < j; ++t) amax = std::fmax(amax, std::fabs(R.alpha[t]) + R.beta[t]);
                const double thr = std::fmax(1e-12, 100.0 * eps_store) * std::fmax(amax, 1e-300);
                for (int t = kept[b] + 1; t <= j; ++t)
                    if (R.beta[t] <= thr) {
                        jr = t;
                        break;
                    }
            }
            const double bnext = R.beta[jr];
            const bool full = j >= cap && jr == j;
            Analysis& an = ans[b];
            const int p = std::min(jr, full ? std::max(2 * k + 4, LZ_KEEP_MAX + 8) : 2 * k + 4);
            analyze(R, jr, bnext, p, an);
            // Breakdown guard: Ritz values are Rayleigh quotients of H and must lie in its Gershgorin
            // interval; values outside it (or non-finite coefficients) mean the Lanczos vectors lost
            // orthogonality beyond repair.  The lane stops unconverged (mixed runs are rerun in FP64).
            bool broken = !std::isfinite(an.tnorm) || !std::isfinite(bnext) || (R.tnorm0 > 0 && an.tnorm > 1e3 * R.tnorm0);
            if (!broken && !an.theta.empty()) {
                const double slack = 1e-6 * std::fmax(std::fmax(std::fabs(hlo[b]), std::fabs(hhi[b])), 1.0);
                if (an.theta.front() < hlo[b] - slack || an.theta.back() > hhi[b] + slack) broken = true;
            }
            if (broken) {
                if (lz_debug())
                    std::fprintf(stderr, "[lz] broken lane=%d j=%d tot=%d tnorm=%.3e tnorm0=%.3e bnext=%.3e lo=%.6f hi=%.6f th0=%.6f thn=%.6f\n",
                                 b, j, tj, an.tnorm, R.tnorm0, bnext, hlo[b], hhi[b], an.theta.empty() ? 0.0 : an.theta.front(),
                                 an.theta.empty() ? 0.0 : an.theta.back());
                R.done = true;
                R.iters = tj;
                R.status = BH_ENOCONV;
                R.theta.clear();
                R.s.clear();
                done_now[b] = 1;
                return true;
            }
            if (R.tnorm0 == 0.0) R.tnorm0 = an.tnorm;
            const bool exhausted = jr < j || bnext <= std::fmax(1e-12, 100.0 * eps_store) * std::fmax(an.tnorm, 1e-300) ||
                                   j >= D;
            if (lz_debug()) {
                std::fprintf(stderr, "[lz] j=%d tot=%d lane=%d tnorm=%.3e bnext=%.3e locks=%d:", j, tj, b, an.tnorm, bnext,
                             nlock[b]);
                for (size_t i = 0; i < an.theta.size() && i < 8; ++i)
                    std::fprintf(stderr, " (%.10f %.1e)", an.theta[i], an.bound[i]);
                std::fprintf(stderr, "\n");
            }
            // FP32-stored vectors: Ritz residuals cannot fall below the storage rounding, ~eps32 ||T||
            const double floor = MIXED ? 4.0 * eps_store * an.tnorm : 0.0;
            // walk distinct levels from the bottom.  With FP32-stored vectors, copies of a degenerate level
            // split at the storage rounding (~eps32 ||T||), so they are merged at that scale here and
            // separated again, if they are distinct, by the FP64 finish.
            const double tdedup = MIXED ? std::fmax(dedup, 50.0 * eps_store) : dedup;
            // lev[q]: first Ritz value of the q-th distinct level; lbound[q]: the smallest bound among the
            // level's copies (a copy that keeps converging carries the level)
            std::vector<int> lev;
            std::vector<double> lbound;
            for (int i = 0; i < static_cast<int>(an.theta.size()); ++i) {
                if (!lev.empty() && std::fabs(an.theta[i] - an.theta[lev.back()]) <= tdedup * an.tnorm) {
                    lbound.back() = std::fmin(lbound.back(), an.bound[i]);
                    continue;
                }
                if (static_cast<int>(lev.size()) == k) break;
                lev.push_back(i);
                lbound.push_back(an.bound[i]);
            }
            // extra candidates finished in FP64 beside the k wanted levels
            std::vector<int> cand = lev;
            for (int i = lev.empty() ? 0 : lev.back() + 1;
                 i < static_cast<int>(an.theta.size()) && static_cast<int>(cand.size()) < std::min(k + 3, LZ_CMAX - R.nfz); ++i) {
                if (std::fabs(an.theta[i] - an.theta[cand.back()]) <= tdedup * an.tnorm) continue;
                cand.push_back(i);
            }
            bool conv = static_cast<int>(lev.size()) == k || exhausted;
            std::vector<double> cthr(lev.size());
            for (size_t q = 0; q < lev.size(); ++q) {
                cthr[q] = std::fmax(tol * std::fmax(eps23, std::fabs(an.theta[lev[q]])), floor);
                if (!mult && lbound[q] >= cthr[q] && !exhausted) conv = false;
            }
            // Eigenvalues with their multiplicity.  One pass (one start vector) meets each eigenspace once, so
            // a pass that has converged freezes the levels it found and the lane starts again from a fresh
            // vector kept orthogonal to every frozen vector: that pass meets the further copies of degenerate
            // levels.  The lowest k values with multiplicity are the lowest k of the frozen values and the
            // current pass's levels; kappa is the k-th of them.  A pass is complete when its levels at or below
            // kappa have converged and its lowest level has (in a later pass that level tells whether the
            // complement of the frozen vectors holds anything at or below kappa; when it lies clearly above
            // kappa it needs only to be told apart from kappa).  A pass that adds nothing at or below kappa
            // ends the solve.
            double kappa = INFINITY;
            const double kmarg = tdedup * an.tnorm;
            int nmerged = static_cast<int>(lev.size());
            if (mult) {
                std::vector<double> mv(R.fz_theta);
                for (int i : lev) mv.push_back(an.theta[i]);
                std::sort(mv.begin(), mv.end());
                nmerged = static_cast<int>(mv.size());
                if (nmerged >= k) kappa = mv[k - 1];
                // every Ritz value of this pass at or below kappa is among the analysed ones
                const bool covered = static_cast<int>(lev.size()) == k ||
                                     (!an.theta.empty() && an.theta.back() > kappa + kmarg) ||
                                     static_cast<int>(an.theta.size()) >= jr;
                conv = (nmerged >= k && covered) || exhausted;
                if (lev.empty()) conv = conv && exhausted;
                for (size_t q = 0; q < lev.size(); ++q) {
                    const double th = an.theta[lev[q]];
                    const bool below = th <= kappa + kmarg;
                    double thr = cthr[q];
                    if (!below && q == 0 && R.pass > 0) thr = std::fmax(thr, 1e-2 * (th - kappa));
                    if ((below || q == 0) && lbound[q] >= thr && !exhausted) conv = false;
                }
            }
            bool capped = false;
            if (mult && conv && tj < maxsteps) {
                std::vector<int> fz;  // this pass's levels among the lowest k
                for (int i : lev)
                    if (an.theta[i] <= kappa + kmarg) fz.push_back(i);
                // frozen values now above kappa are no longer among the lowest k: their vectors are dropped (a
                // later pass may meet them again, above kappa)
                std::vector<int> keep;
                for (int l = 0; l < R.nfz; ++l)
                    if (R.fz_theta[l] <= kappa + kmarg) keep.push_back(l);
                const int nnew = static_cast<int>(keep.size() + fz.size());
                if (!fz.empty() && nnew < D && nnew <= ymax && R.pass + 1 < LZ_PASS_MAX) {
                    Deflation& P = dplan[b];
                    P.theta.clear();
                    P.s.clear();
                    for (int i : fz) {
                        P.theta.push_back(an.theta[i]);
                        P.s.push_back(an.s[i]);
                    }
                    P.keep = keep;
                    deflate_now[b] = 1;
                    return true;
                }
                // out of frozen slots or passes: the copies found so far are reported, the solve is not certified
                capped = !fz.empty() && nnew < D;
            }
            if (conv || tj >= maxsteps) {
                R.done = true;
                R.exhausted = exhausted;
                R.iters = tj;
                R.status = conv && !capped && (mult ? (nmerged >= k || exhausted) : static_cast<int>(lev.size()) == k)
                               ? BH_OK
                               : BH_ENOCONV;
                if (lz_debug())
                    std::fprintf(stderr, "[lz] done lane=%d tot=%d conv=%d capped=%d nmerged=%d exhausted=%d kappa=%.10f\n", b,
                                 tj, conv ? 1 : 0, capped ? 1 : 0, nmerged, exhausted ? 1 : 0, kappa);
                R.theta.clear();
                R.s.clear();
                for (int i : cand) {
                    R.theta.push_back(an.theta[i]);
                    R.s.push_back(an.s[i]);
                }
                R.tnorm = an.tnorm;
                done_now[b] = 1;
                return true;
            }
            // Next analysis: where the slowest unconverged wanted level is predicted to meet its threshold,
            // extrapolating log(bound / threshold) linearly in steps from the lane's previous analysis
            // (Lanczos bounds fall at least geometrically once a level has emerged, so the prediction errs
            // late).  Never later than the default interval, at least LZ_CHECK_MIN steps ahead.
            if (static_cast<int>(lev.size()) == k) {
                double gap = 0.0;
                for (size_t q = 0; q < lev.size(); ++q)
                    if (lbound[q] >= cthr[q]) gap = std::fmax(gap, std::log(lbound[q] / cthr[q]));
                if (R.prev_tot >= 0 && tj > R.prev_tot && R.prev_gap > gap) {
                    const double rate = (R.prev_gap - gap) / (tj - R.prev_tot);
                    const double d = std::ceil(1.05 * gap / rate) + 1.0;
                    if (d < nextc[b] - j) nextc[b] = j + std::max(LZ_CHECK_MIN, static_cast<int>(d));
                }
                R.prev_gap = gap;
                R.prev_tot = tj;
            } else {
                R.prev_tot = -1;
            }
            if (full) {
                // thick restart: keep the lowest distinct Ritz pairs (one per cluster of copies)
                std::vector<int> keep;
                const int pk = std::min(LZ_KEEP_MAX - 1, std::max(1, j / 2));
                for (int i = 0; i < static_cast<int>(an.theta.size()) && static_cast<int>(keep.size()) < pk; ++i) {
                    if (!keep.empty() && std::fabs(an.theta[i] - an.theta[keep.back()]) <= tdedup * an.tnorm) continue;
                    keep.push_back(i);
                }
                Restart& P = plan[b];
                P.j = j;
                P.bnext = bnext;
                P.theta.clear();
                P.s.clear();
                for (int i : keep) {
                    P.theta.push_back(an.theta[i]);
                    P.s.push_back(an.s[i]);
                }
                restart_now[b] = 1;
                return true;
            }
            // selective orthogonalisation: lock converged wanted-region Ritz vectors (Parlett and Scott:
            // bound below sqrt(eps) ||T||); not needed when every step reorthogonalises fully
            if (fullre) return false;
            const double lock_thr = std::sqrt(eps_store) * an.tnorm;
            for (size_t q = 0; q < lev.size(); ++q) {
                const int i = lev[q];
                if (an.bound[i] > lock_thr) continue;
                if (nlock[b] + static_cast<int>(new_locks[b].size()) >= LZ_LMAX) break;
                // a level already locked keeps its Ritz value only up to its bound: match within the
                // lock threshold, so that one level is never locked twice (near-parallel locked vectors
                // would break the Gram-Schmidt step of the orthogonalisation)
                bool have = false;
                for (size_t l = R.nfz; l < R.lock_theta.size(); ++l)
                    if (std::fabs(R.lock_theta[l] - an.theta[i]) <= lock_thr) have = true;
                for (int q : new_locks[b])
                    if (std::fabs(an.theta[q] - an.theta[i]) <= lock_thr) have = true;
                if (have) continue;
                new_locks[b].push_back(i);
                lock_slot[b].push_back(-1);
            }
            // Refresh (FP64 storage): a locked vector is off by about bound/gap from its eigenvector.  When a
            // close level surfaces later, the Krylov space kept orthogonal to the old vector rebuilds a second
            // copy of the locked level from that remainder; a locked level whose Ritz bound has since dropped a
            // thousandfold gets its vector re-formed from the current Ritz vector.  FP32-stored runs lock far
            // earlier (sqrt(eps32) ||T||) and measured slower with refreshes, so they keep their first vector.
            for (int l = R.nfz; !MIXED && l < static_cast<int>(R.lock_theta.size()); ++l) {
                int best = -1;
                for (int i = 0; i < static_cast<int>(an.theta.size()); ++i)
                    if (std::fabs(an.theta[i] - R.lock_theta[l]) <= lock_thr && (best < 0 || an.bound[i] < an.bound[best]))
                        best = i;
                if (best < 0 || !(an.bound[best] <= 1e-3 * R.lock_bound[l])) continue;
                bool taken = false;
                for (int q : new_locks[b]) taken = taken || q == best;
                if (taken) continue;
                new_locks[b].push_back(best);
                lock_slot[b].push_back(l);
            }
            return false;
        };
        auto analyse_run = [&](int b) {
            for (;;) {
                const int e = std::min({nextc[b], cap, jsync[b] + std::max(0, maxsteps - totsync[b])});
                if (e > jsync[b]) return;
                const int tj = totsync[b] - (jsync[b] - e);
                jan[b] = e;
                tot_an[b] = tj;
                if (analyse_at(b, e, tj) || !new_locks[b].empty()) return;
            }
        };
        {
            const auto ta0 = std::chrono::steady_clock::now();
            struct AnaClock {
                LzEngine* e;
                std::chrono::steady_clock::time_point t;
                ~AnaClock() {
                    e->prof.analysis_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
                }
            } ac{this, ta0};
            std::vector<int> todo;  // lanes with an analysis due
            for (int b = 0; b < B; ++b)
                if (act[b] && std::min({nextc[b], cap, jsync[b] + std::max(0, maxsteps - totsync[b])}) <= jsync[b])
                    todo.push_back(b);
            const int nth = std::max(1, std::min(static_cast<int>(todo.size()),
                                                 static_cast<int>(std::thread::hardware_concurrency())));
            if (nth <= 1) {
                for (int b : todo) analyse_run(b);
            } else {
                std::vector<std::thread> pool;
                for (int t = 0; t < nth; ++t)
                    pool.emplace_back([&, t]() {
                        for (size_t q = t; q < todo.size(); q += nth) analyse_run(todo[q]);
                    });
                for (auto& th : pool) th.join();
            }
        }
        bool any_lock = false, any_done = false, any_restart = false, any_deflate = false;
        for (int b = 0; b < B; ++b) {
            if (done_now[b]) {
                act[b] = 0;
                any_done = true;
                finished.push_back(b);
            }
            if (!new_locks[b].empty()) any_lock = true;
            if (restart_now[b]) any_restart = true;
            if (deflate_now[b]) any_deflate = true;
        }
        // roll lanes back to the step of their decision
        bool rolled = false;
        std::vector<int> src_slot(B, -1);  // current vector of lane b restored from this slot
        for (int b = 0; b < B; ++b) {
            const bool decided = done_now[b] || !new_locks[b].empty() || restart_now[b] || deflate_now[b];
            if (!decided || jan[b] >= jl[b]) continue;
            jl[b] = jan[b];
            tot[b] = tot_an[b];
            if (done_now[b]) continue;  // retired: only the host counters
            src_slot[b] = jl[b];
            rolled = true;
        }
        if (rolled) {
            step_h = jl;
            BH_HIP(ctx, hipMemcpyAsync(st.step, step_h.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipMemcpyAsync(lock_map.p, src_slot.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((k_lz_restore_current<XT, B>), dim3(static_cast<unsigned>((vs + 255) / 256)), dim3(256), 0, s, V,
                               Xc, vs, lock_map.p);
            BH_HIP(ctx, hipStreamSynchronize(s));
        }
        if (any_lock) BH_TRY(lock(new_locks, lock_slot, ans));
        if (any_restart) BH_TRY(restart(restart_now));
        if (any_deflate) {
            if (!defl_barrier) {
                BH_TRY(deflate(deflate_now));
            } else {
                // the lane waits (inactive, not finished) until every unfinished lane has completed its pass
                for (int b = 0; b < B; ++b)
                    if (deflate_now[b]) {
                        waiting[b] = 1;
                        act[b] = 0;
                    }
                any_done = true;  // the device's active flags are refreshed below
            }
        }
        // deflation barrier: once no lane iterates, the waiting lanes start their next pass together (one
        // formation of their frozen vectors, one read of the Krylov basis)
        if (defl_barrier) {
            bool any_act = false, any_wait = false;
            for (int b = 0; b < B; ++b) {
                any_act = any_act || act[b];
                any_wait = any_wait || waiting[b];
            }
            if (!any_act && any_wait) {
                std::vector<int> flag(waiting);
                BH_TRY(deflate(flag));
                for (int b = 0; b < B; ++b)
                    if (flag[b]) {
                        waiting[b] = 0;
                        act[b] = 1;
                    }
                any_done = true;
            }
        }
        if (any_done) {
            BH_HIP(ctx, hipMemcpyAsync(st.active, act.data(), B * 4, hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
        }
        return BH_OK;
    }

    int finish(const std::vector<int>& lanes, std::vector<LaneResult>& out, bool want_vectors) override {
        out.assign(lanes.size(), LaneResult());
        int jmax = 0;
        for (int b : lanes) jmax = std::max(jmax, jl[b]);
        std::vector<double> hsc(static_cast<size_t>(jmax + 1) * B);
        BH_HIP(ctx, hipMemcpyAsync(hsc.data(), st.scale, hsc.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        // candidates of lane b: its frozen vectors (Y slots 0 .. nf[b] - 1, earlier passes) first, then the
        // Ritz vectors of its last pass
        std::vector<int> nf(B, 0), nc(B, 0), nt(B, 0);
        int nvf = 0, nfmax = 0;
        for (int b : lanes) {
            nf[b] = mult ? std::min(rh[b].nfz, LZ_CMAX) : 0;
            nc[b] = std::min(static_cast<int>(rh[b].theta.size()), LZ_CMAX - nf[b]);
            nt[b] = nf[b] + nc[b];
            nvf = std::max(nvf, nt[b]);
            nfmax = std::max(nfmax, nf[b]);
        }
        std::vector<double> theta_f(static_cast<size_t>(std::max(nvf, 1)) * B, NAN), resid_f(theta_f.size(), NAN);
        std::vector<double> dots;
        std::vector<double> hy;
        if (nvf > 0) {
            int njmax = 0;
            std::vector<int> njb(B, 0);
            for (int b : lanes)
                if (nc[b] > 0) {
                    njb[b] = static_cast<int>(rh[b].s[0].size());
                    njmax = std::max(njmax, njb[b]);
                }
            std::vector<double> C(static_cast<size_t>(std::max(njmax, 1)) * nvf * B, 0.0);
            for (int b : lanes)
                for (int i = 0; i < nc[b]; ++i)
                    for (int t = 0; t < static_cast<int>(rh[b].s[i].size()); ++t)
                        C[(static_cast<size_t>(t) * nvf + nf[b] + i) * B + b] = rh[b].s[i][t] * hsc[static_cast<size_t>(t) * B + b];
            DevArray<double> Yf, HY;  // engines without the context workspace own their scratch
            double *Yfp = nullptr, *HYp = nullptr;
            // mixed lanes get slots for the corrections of the FP64 refinement of their lowest k + 1 candidates
            // (dropped when the memory is not there: such a lane is then rerun in FP64 as without the refinement)
            int pref = MIXED && lz_refine_steps() > 0 ? std::min(nvf, k + 1) : 0;
            {
                int rc0 = scratch(LZ_WS_A, Yf, static_cast<size_t>(nvf + pref) * vs, &Yfp);
                if (rc0 == BH_OK) rc0 = scratch(LZ_WS_B, HY, static_cast<size_t>(nvf + pref) * vs, &HYp);
                if (rc0 != BH_OK && pref > 0) {
                    pref = 0;
                    rc0 = scratch(LZ_WS_A, Yf, static_cast<size_t>(nvf) * vs, &Yfp);
                    if (rc0 == BH_OK) rc0 = scratch(LZ_WS_B, HY, static_cast<size_t>(nvf) * vs, &HYp);
                }
                BH_TRY(rc0);
            }
            BH_TRY(form_vectors(njb, nvf, C, Yfp, vs));
            if (nfmax > 0) {
                BH_HIP(ctx, hipMemcpyAsync(dnj.p, nf.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL((k_lz_copy_frozen<XT, B>), dim3(static_cast<unsigned>((vs + 255) / 256)), dim3(256), 0, s, Y,
                                   Yfp, vs, dnj.p, nfmax);
                BH_HIP(ctx, hipStreamSynchronize(s));
            }
            // exact FP64 operator on each candidate (packed words, FP64 table, FP64 vectors)
            DevArray<double> dJ, dU;
            BH_TRY(dJ.alloc(ctx, B));
            BH_TRY(dU.alloc(ctx, B));
            BH_HIP(ctx, hipMemcpyAsync(dJ.p, Jb.data(), B * 8, hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipMemcpyAsync(dU.p, Ub.data(), B * 8, hipMemcpyHostToDevice, s));
            for (int i = 0; i < nvf; ++i) {
                KTimer kt(ctx, "lanczos_finish_spmv", s);
                SpmvArgs<double> sa{D, h->row_ptr, h->words, nullptr, nullptr, h->diag, nullptr, h->table, h->code_bits,
                                    dJ.p, dU.p, Yfp + static_cast<int64_t>(i) * vs, HYp + static_cast<int64_t>(i) * vs, 1};
                BH_TRY(bh_apply(ctx, h, &sa, nullptr, BH_FMT_PACKED, B, s));
            }
            BH_HIP(ctx, hipGetLastError());
            BH_TRY(rayleigh_ritz(lanes, nvf, nt, Yfp, HYp));
            dots.assign(static_cast<size_t>(2 * nvf) * B, 0.0);
            BH_TRY(measure(nvf, Yfp, HYp, nullptr, dots, theta_f, resid_f));
            // lanes whose FP64 Rayleigh-Ritz values fail the acceptance test get up to lz_refine_steps() Davidson
            // steps in FP64 (refine); a lane that still fails is reported as before and rerun in FP64
            for (int it = 0; pref > 0 && it < lz_refine_steps(); ++it) {
                std::vector<int> on(B, 0);
                bool any = false;
                for (size_t q = 0; q < lanes.size(); ++q) {
                    const int b = lanes[q];
                    if (nt[b] == 0 || rh[b].status != BH_OK) continue;
                    std::vector<std::pair<double, int>> lv;
                    LaneResult L;
                    select_levels(b, nt[b], theta_f, resid_f, L, lv);
                    // a lane short of k levels is left to the rerun: the corrections of the levels it has do
                    // not supply a missing copy
                    if (static_cast<int>(lv.size()) < k && !rh[b].exhausted) continue;
                    if (!bh_lane_accepted(L.status, L.evals.data(), L.errest.data(), k, o.accept_rel)) {
                        on[b] = 1;
                        any = true;
                    }
                }
                if (!any) break;
                BH_TRY(refine(on, nvf, pref, Yfp, HYp, theta_f));
                prof.refine_steps += 1;
                std::vector<double> d2(dots.size()), t2(theta_f.size()), r2(resid_f.size());
                BH_TRY(measure(nvf, Yfp, HYp, &on, d2, t2, r2));
                for (int b = 0; b < B; ++b)
                    if (on[b]) {
                        for (int i = 0; i < nvf; ++i) {
                            theta_f[i * B + b] = t2[i * B + b];
                            resid_f[i * B + b] = r2[i * B + b];
                            dots[(2 * i) * B + b] = d2[(2 * i) * B + b];
                            dots[(2 * i + 1) * B + b] = d2[(2 * i + 1) * B + b];
                        }
                        rh[b].refined = it + 1;
                        prof.refined_lanes += 1;
                    }
            }
            if (want_vectors) {
                hy.resize(static_cast<size_t>(nvf) * vs);
                BH_HIP(ctx, hipMemcpyAsync(hy.data(), Yfp, hy.size() * 8, hipMemcpyDeviceToHost, s));
                BH_HIP(ctx, hipStreamSynchronize(s));
            }
        }
        for (size_t q = 0; q < lanes.size(); ++q) {
            const int b = lanes[q];
            RunHost& R = rh[b];
            LaneResult& L = out[q];
            std::vector<std::pair<double, int>> lv;
            select_levels(b, nt[b], theta_f, resid_f, L, lv);
            if (lz_debug()) {
                std::fprintf(stderr, "[lz] finish lane=%d run_status=%d status=%d refined=%d levels=%d frozen=%d pass=%d:", b,
                             R.status, L.status, R.refined, static_cast<int>(lv.size()), R.nfz, R.pass);
                for (int i = 0; i < nt[b]; ++i) std::fprintf(stderr, " (%.10f r=%.1e)", theta_f[i * B + b], resid_f[i * B + b]);
                std::fprintf(stderr, "\n");
            }
            if (want_vectors) L.vecs.assign(static_cast<size_t>(k) * D, 0.0);
            for (int i = 0; i < k && i < static_cast<int>(lv.size()); ++i) {
                if (want_vectors && !hy.empty()) {
                    const int iv = lv[i].second;
                    const double yy = dots[(2 * iv) * B + b];
                    const double f = yy > 0 ? 1.0 / std::sqrt(yy) : 0.0;
                    for (int64_t r = 0; r < D; ++r)
                        L.vecs[static_cast<size_t>(i) * D + r] = hy[static_cast<size_t>(iv) * vs + r * B + b] * f;
                }
            }
        }
        return BH_OK;
    }

   private:
    // Levels of lane b from its nl candidates (FP64 Rayleigh quotients theta_f, residual norms resid_f):
    // L.status, evals, resid, errest and the chosen candidates lv (value, slot).
    void select_levels(int b, int nl, const std::vector<double>& theta_f, const std::vector<double>& resid_f, LaneResult& L,
                       std::vector<std::pair<double, int>>& lv) {
            const RunHost& R = rh[b];
            L.iters = R.iters;
            L.status = R.status;
            L.evals.assign(k, NAN);
            L.resid.assign(k, NAN);
            L.errest.assign(k, NAN);
            std::vector<std::pair<double, int>> lv0;
            lv.clear();
            for (int i = 0; i < nl; ++i) lv0.push_back({theta_f[i * B + b], i});
            std::sort(lv0.begin(), lv0.end());
            const double mtol = dedup * std::fmax(R.tnorm, 1e-300);
            auto rs_of = [&](const std::pair<double, int>& p) { return resid_f[p.second * B + b]; };
            bool dropped = false;
            if (mult) {
                // With multiplicity.  The Rayleigh-Ritz vectors are orthonormal, so m pairs whose residuals are
                // small certify m eigenvalues, copies of one level included (each within its residual).  A pair
                // whose residual is large is kept as an uncertified value, unless its interval holds the value
                // of a certified pair: then it is a direction left over from two approximations of one
                // eigenvector and certifies nothing.
                const double rfloor = MIXED ? 100.0 * eps_store * std::fmax(R.tnorm, 1e-300) : 0.0;
                auto certified = [&](const std::pair<double, int>& p) {
                    return rs_of(p) <= std::fmax(1e3 * tol * std::fmax(eps23, std::fabs(p.first)), rfloor);
                };
                std::vector<double> good;
                for (const auto& qq : lv0)
                    if (std::isfinite(qq.first) && certified(qq)) good.push_back(qq.first);
                for (const auto& qq : lv0) {
                    if (!std::isfinite(qq.first)) continue;
                    if (!certified(qq)) {
                        bool inside = false;
                        for (double g : good) inside = inside || std::fabs(g - qq.first) <= rs_of(qq);
                        if (inside) {
                            if (static_cast<int>(lv.size()) < k) dropped = true;
                            continue;
                        }
                    }
                    lv.push_back(qq);
                }
            } else {
                // Distinct levels.  Rayleigh quotients of one level agree to ~||r||^2/gap and are merged.  Each
                // finished pair certifies an eigenvalue within its residual of its Rayleigh quotient (Weinstein);
                // a pair whose interval holds the value of a pair with a smaller residual certifies nothing new
                // (it can be an unconverged copy of that level, which finite-precision Lanczos produces) and is
                // dropped.
                for (const auto& qq : lv0) {
                    if (!std::isfinite(qq.first)) continue;
                    if (!lv.empty() && std::fabs(qq.first - lv.back().first) <= mtol) {
                        if (rs_of(qq) < rs_of(lv.back())) lv.back() = qq;
                        continue;
                    }
                    if (!lv.empty()) {
                        const double sep = std::fabs(qq.first - lv.back().first);
                        const double rq = rs_of(qq), rb = rs_of(lv.back());
                        if (rq >= sep && rq >= rb) {
                            if (static_cast<int>(lv.size()) < k) dropped = true;
                            continue;
                        }
                        if (rb >= sep && rb > rq) {
                            if (static_cast<int>(lv.size()) <= k) dropped = true;
                            lv.back() = qq;
                            continue;
                        }
                    }
                    lv.push_back(qq);
                }
            }
            // A mixed-precision solve that had to drop one among its wanted levels is reported unconverged, so
            // that it is rerun in FP64
            if (MIXED && dropped && L.status == BH_OK) L.status = BH_ENOCONV;
            // fewer than k levels is an answer only when the Krylov space was exhausted
            if (L.status == BH_OK && static_cast<int>(lv.size()) < k && !R.exhausted) L.status = BH_ENOCONV;
            for (int i = 0; i < k && i < static_cast<int>(lv.size()); ++i) {
                const double ev = lv[i].first;
                const double rs = rs_of(lv[i]);
                double gap = INFINITY;  // to the nearest other finished value of the lane outside ev's cluster
                for (const auto& qq : lv0)
                    if (std::isfinite(qq.first) && std::fabs(qq.first - ev) > mtol) gap = std::fmin(gap, std::fabs(qq.first - ev));
                L.evals[i] = ev;
                L.resid[i] = rs;
                L.errest[i] = std::isfinite(gap) ? rs * rs / gap : rs;
            }
    }
    // FP64 Rayleigh quotients and residual norms of slots 0 .. nvf - 1 (LZ_NVMAX vectors per launch): dots
    // (<y,y>, <y,Hy>), theta and resid, for every lane
    int measure(int nvf, const double* Yfp, const double* HYp, const std::vector<int>* /*unused*/, std::vector<double>& dots,
                std::vector<double>& theta_f, std::vector<double>& resid_f) {
        std::vector<double> rr(static_cast<size_t>(nvf) * B);
        for (int i0 = 0; i0 < nvf; i0 += LZ_NVMAX) {
            const int nq = std::min(LZ_NVMAX, nvf - i0);
            hipLaunchKernelGGL((k_lz_dots<double, B>), dim3(grid_k2), dim3(LZ_WG), 0, s, Yfp + static_cast<int64_t>(i0) * vs,
                               HYp + static_cast<int64_t>(i0) * vs, D, nq, 0, nullptr, part.p);
            reduce_plain(grid_k2, 2 * nq * B, dnorm.p);
            BH_HIP(ctx, hipMemcpyAsync(dots.data() + static_cast<size_t>(2 * i0) * B, dnorm.p, static_cast<size_t>(2 * nq) * B * 8,
                                       hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
        }
        for (int i = 0; i < nvf; ++i)
            for (int b = 0; b < B; ++b) {
                const double yy = dots[(2 * i) * B + b], yhy = dots[(2 * i + 1) * B + b];
                theta_f[i * B + b] = yy > 0 ? yhy / yy : NAN;
            }
        DevArray<double> dth;
        BH_TRY(dth.alloc(ctx, static_cast<size_t>(nvf) * B));
        std::vector<double> th0(theta_f.begin(), theta_f.begin() + static_cast<size_t>(nvf) * B);
        for (auto& v : th0)
            if (!std::isfinite(v)) v = 0.0;
        BH_HIP(ctx, hipMemcpyAsync(dth.p, th0.data(), th0.size() * 8, hipMemcpyHostToDevice, s));
        for (int i0 = 0; i0 < nvf; i0 += LZ_NVMAX) {
            const int nq = std::min(LZ_NVMAX, nvf - i0);
            hipLaunchKernelGGL((k_lz_dots<double, B>), dim3(grid_k2), dim3(LZ_WG), 0, s, Yfp + static_cast<int64_t>(i0) * vs,
                               HYp + static_cast<int64_t>(i0) * vs, D, nq, 1, dth.p + static_cast<size_t>(i0) * B, part.p);
            reduce_plain(grid_k2, nq * B, dnorm.p);
            BH_HIP(ctx, hipMemcpyAsync(rr.data() + static_cast<size_t>(i0) * B, dnorm.p, static_cast<size_t>(nq) * B * 8,
                                       hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));  // dth is freed on return
        }
        for (int i = 0; i < nvf; ++i)
            for (int b = 0; b < B; ++b) {
                const double yy = dots[(2 * i) * B + b];
                resid_f[i * B + b] = yy > 0 ? std::sqrt(rr[i * B + b] / yy) : NAN;
            }
        return BH_OK;
    }
    // One Davidson step in FP64 for the lanes with on[b] = 1, over their nvf candidates in slots 0 .. nvf - 1 of
    // Y (and HY = H Y), lowest first: the corrections t_i = (H y_i - theta_i y_i) / (diag(H) - theta_i) of the
    // lowest np, each normalised, go to slots nvf .. nvf + np - 1 with their images under the exact FP64
    // operator, and a Rayleigh-Ritz over the nvf + np vectors leaves the lowest nvf Ritz vectors in slots
    // 0 .. nvf - 1 (the other lanes keep theirs).
    // The candidates are combinations of FP32-stored Krylov vectors, so their residuals stop near the FP32
    // rounding of H y; where that fails the certificate (small J/U: tiny |E0| or close levels), the diagonal
    // U V dominates the operator and the preconditioned corrections remove most of the residual.
    int refine(const std::vector<int>& on, int nvf, int np, double* Yp, double* HYp, const std::vector<double>& theta_f) {
        double* Tp = Yp + static_cast<int64_t>(nvf) * vs;
        double* HTp = HYp + static_cast<int64_t>(nvf) * vs;
        std::vector<double> th0(theta_f.begin(), theta_f.begin() + static_cast<size_t>(np) * B), dl(B);
        for (auto& v : th0)
            if (!std::isfinite(v)) v = 0.0;
        for (int b = 0; b < B; ++b) dl[b] = LZ_REFINE_CLAMP * std::fmax(std::fmax(std::fabs(Ub[b]), std::fabs(Jb[b])), 1e-300);
        DevArray<double> dth, ddl, dU, dJ, df;
        DevArray<int> don;
        BH_TRY(dth.alloc(ctx, th0.size()));
        BH_TRY(ddl.alloc(ctx, B));
        BH_TRY(dU.alloc(ctx, B));
        BH_TRY(dJ.alloc(ctx, B));
        BH_TRY(df.alloc(ctx, static_cast<size_t>(np) * B));
        BH_TRY(don.alloc(ctx, B));
        BH_HIP(ctx, hipMemcpyAsync(dth.p, th0.data(), th0.size() * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(ddl.p, dl.data(), B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dU.p, Ub.data(), B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dJ.p, Jb.data(), B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(don.p, on.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
        {
            KTimer kt(ctx, "lanczos_refine", s);
            const int64_t n = static_cast<int64_t>(np) * vs;
            hipLaunchKernelGGL((k_lz_dav_t<B>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, Yp, HYp, Tp, D, np,
                               dth.p, dU.p, ddl.p, h->diag, don.p);
        }
        // unit corrections (a zero correction stays zero)
        std::vector<double> nrm(static_cast<size_t>(np) * B), f(nrm.size());
        for (int i0 = 0; i0 < np; i0 += LZ_NVMAX) {
            const int nq = std::min(LZ_NVMAX, np - i0);
            hipLaunchKernelGGL((k_lz_dots<double, B>), dim3(grid_k2), dim3(LZ_WG), 0, s, Tp + static_cast<int64_t>(i0) * vs, nullptr,
                               D, nq, 2, nullptr, part.p);
            reduce_plain(grid_k2, nq * B, dnorm.p);
            BH_HIP(ctx, hipMemcpyAsync(nrm.data() + static_cast<size_t>(i0) * B, dnorm.p, static_cast<size_t>(nq) * B * 8,
                                       hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
        }
        for (size_t i = 0; i < nrm.size(); ++i) f[i] = nrm[i] > 0 && std::isfinite(nrm[i]) ? 1.0 / std::sqrt(nrm[i]) : 0.0;
        BH_HIP(ctx, hipMemcpyAsync(df.p, f.data(), f.size() * 8, hipMemcpyHostToDevice, s));
        {
            const int64_t n = static_cast<int64_t>(np) * vs;
            hipLaunchKernelGGL((k_lz_scale<double, B>), dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, Tp, vs, np,
                               df.p);
        }
        for (int i = 0; i < np; ++i) {
            KTimer kt(ctx, "lanczos_finish_spmv", s);
            SpmvArgs<double> sa{D, h->row_ptr, h->words, nullptr, nullptr, h->diag, nullptr, h->table, h->code_bits,
                                dJ.p, dU.p, Tp + static_cast<int64_t>(i) * vs, HTp + static_cast<int64_t>(i) * vs, 1};
            BH_TRY(bh_apply(ctx, h, &sa, nullptr, BH_FMT_PACKED, B, s));
        }
        BH_HIP(ctx, hipGetLastError());
        std::vector<int> lr, n2(B, 0);
        for (int b = 0; b < B; ++b)
            if (on[b]) {
                lr.push_back(b);
                n2[b] = nvf + np;
            }
        BH_TRY(rayleigh_ritz(lr, nvf + np, n2, Yp, HYp, nvf));
        BH_HIP(ctx, hipStreamSynchronize(s));  // the small arrays above are freed on return
        return BH_OK;
    }

   private:
    struct Graph {
        hipGraphExec_t exec = nullptr;
        int lm = -1, steps = 0;
    };
    // launches the captured graph of `ns` steps for orthogonalisation variant LM, capturing it on first use
    int replay(int LM, int ns) {
        Graph* g = nullptr;
        for (auto& q : graphs)
            if (q.exec && q.lm == LM && q.steps == ns) g = &q;
        if (!g) {
            const auto tc0 = std::chrono::steady_clock::now();
            struct CaptureClock {
                LzEngine* e;
                std::chrono::steady_clock::time_point t;
                ~CaptureClock() {
                    e->prof.capture_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
                }
            } cc{this, tc0};
            hipGraph_t graph = nullptr;
            BH_HIP(ctx, hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int t = 0; t < ns; ++t) step(LM);
            BH_HIP(ctx, hipStreamEndCapture(s, &graph));
            Graph ng;
            ng.lm = LM;
            ng.steps = ns;
            const hipError_t e = hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            BH_HIP(ctx, e);
            graphs.push_back(ng);
            g = &graphs.back();
        }
        BH_HIP(ctx, hipGraphLaunch(g->exec, s));
        return BH_OK;
    }
    // Rayleigh-Ritz over each lane's finished candidates (FP64 vectors Y, HY = H Y): with G = Y^T Y and
    // A = Y^T H Y, the pencil (A, G) is solved on the host after dropping the numerically dependent
    // directions of G (copies of one level that finite-precision Lanczos produces), and Y, HY are replaced
    // in place by the Ritz vectors of the span and their images.  A candidate that mixes two eigenvectors
    // is separated when the other one is in the span.  The lowest nout (<= nvf, default nvf) Ritz vectors are
    // kept; lanes not listed keep their first nout vectors unchanged (identity coefficients, exact).
    int rayleigh_ritz(const std::vector<int>& lanes, int nvf, const std::vector<int>& nt, double* Yp, double* HYp,
                      int nout = -1) {
        if (nout < 0 || nout > nvf) nout = nvf;
        DevArray<double> dg, dz;
        std::vector<double> g(static_cast<size_t>(nvf) * 2 * nvf * B);
        if (nvf <= LZ_NVMAX) {
            // rows i0 .. i0 + NI - 1 of [G | A] per launch, NV columns each (NV >= nvf)
            const bool small = nvf <= 8;
            const int NI = small ? 4 : 2, NV = small ? 8 : LZ_NVMAX;
            const int nblk = (nvf + NI - 1) / NI;
            const size_t per_blk = static_cast<size_t>(NI) * 2 * NV * B;
            BH_TRY(dg.alloc(ctx, nblk * per_blk));
            for (int q = 0; q < nblk; ++q) {
                if (small)
                    hipLaunchKernelGGL((k_lz_gram_rows<B, 8, 4>), dim3(grid_k2), dim3(LZ_WG), 0, s, Yp, HYp, D, nvf, q * NI, part.p);
                else
                    hipLaunchKernelGGL((k_lz_gram_rows<B, LZ_NVMAX, 2>), dim3(grid_k2), dim3(LZ_WG), 0, s, Yp, HYp, D, nvf, q * NI,
                                       part.p);
                reduce_plain(grid_k2, static_cast<int>(per_blk), dg.p + q * per_blk);
            }
            std::vector<double> gg(nblk * per_blk);
            BH_HIP(ctx, hipMemcpyAsync(gg.data(), dg.p, gg.size() * 8, hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
            for (int i = 0; i < nvf; ++i)
                for (int w = 0; w < 2; ++w)
                    for (int q = 0; q < nvf; ++q)
                        for (int b = 0; b < B; ++b)
                            g[(static_cast<size_t>(i) * 2 * nvf + w * nvf + q) * B + b] =
                                gg[(i / NI) * per_blk + ((static_cast<size_t>(i % NI) * 2 + w) * NV + q) * B + b];
        } else {
            // blocks of 2 rows x LZ_NVMAX columns
            constexpr int NI = 2, NV = LZ_NVMAX;
            const int nri = (nvf + NI - 1) / NI, nqi = (nvf + NV - 1) / NV;
            const size_t per_blk = static_cast<size_t>(NI) * 2 * NV * B;
            BH_TRY(dg.alloc(ctx, static_cast<size_t>(nri) * nqi * per_blk));
            for (int ri = 0; ri < nri; ++ri)
                for (int qi = 0; qi < nqi; ++qi) {
                    hipLaunchKernelGGL((k_lz_gram_blk<B, NV, NI>), dim3(grid_k2), dim3(LZ_WG), 0, s, Yp, HYp, D, nvf, ri * NI,
                                       qi * NV, part.p);
                    reduce_plain(grid_k2, static_cast<int>(per_blk), dg.p + (static_cast<size_t>(ri) * nqi + qi) * per_blk);
                }
            std::vector<double> gg(static_cast<size_t>(nri) * nqi * per_blk);
            BH_HIP(ctx, hipMemcpyAsync(gg.data(), dg.p, gg.size() * 8, hipMemcpyDeviceToHost, s));
            BH_HIP(ctx, hipStreamSynchronize(s));
            for (int i = 0; i < nvf; ++i)
                for (int w = 0; w < 2; ++w)
                    for (int q = 0; q < nvf; ++q)
                        for (int b = 0; b < B; ++b)
                            g[(static_cast<size_t>(i) * 2 * nvf + w * nvf + q) * B + b] =
                                gg[((static_cast<size_t>(i / NI) * nqi + q / NV) * per_blk) +
                                   ((static_cast<size_t>(i % NI) * 2 + w) * NV + q % NV) * B + b];
        }
        std::vector<double> Z(static_cast<size_t>(nvf) * nout * B, 0.0);
        {
            std::vector<char> listed(B, 0);
            for (int b : lanes) listed[b] = 1;
            for (int b = 0; b < B; ++b)
                if (!listed[b])
                    for (int i = 0; i < nout; ++i) Z[(static_cast<size_t>(i) * nout + i) * B + b] = 1.0;
        }
        for (int b : lanes) {
            const int n = std::min(nt[b], nvf);
            if (n == 0) continue;
            auto G = [&](int i, int q) { return g[(static_cast<size_t>(i) * 2 * nvf + q) * B + b]; };
            auto A = [&](int i, int q) { return g[(static_cast<size_t>(i) * 2 * nvf + nvf + q) * B + b]; };
            std::vector<double> Gm(n * n), gw(n), Ug(n * n);
            for (int i = 0; i < n; ++i)
                for (int q = 0; q < n; ++q) Gm[i * n + q] = 0.5 * (G(i, q) + G(q, i));
            bhtri::jacobi_eig(n, Gm.data(), gw.data(), Ug.data());
            const double gmax = gw[n - 1];
            std::vector<int> keep;
            for (int c = 0; c < n; ++c)
                if (gw[c] > 1e-10 * gmax && gw[c] > 0.0) keep.push_back(c);
            const int nk = static_cast<int>(keep.size());
            if (nk == 0) continue;
            // S = U_keep diag(g^-1/2) (n x nk); Ahat = S^T A S
            std::vector<double> S(n * nk), Ah(nk * nk), th(nk), Wv(nk * nk);
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < nk; ++c) S[i * nk + c] = Ug[i * n + keep[c]] / std::sqrt(gw[keep[c]]);
            for (int c = 0; c < nk; ++c)
                for (int d = 0; d < nk; ++d) {
                    double v = 0.0;
                    for (int i = 0; i < n; ++i)
                        for (int q = 0; q < n; ++q) v += S[i * nk + c] * 0.5 * (A(i, q) + A(q, i)) * S[q * nk + d];
                    Ah[c * nk + d] = v;
                }
            bhtri::jacobi_eig(nk, Ah.data(), th.data(), Wv.data());
            for (int i = 0; i < n; ++i)
                for (int c = 0; c < nk && c < nout; ++c) {
                    double v = 0.0;
                    for (int d = 0; d < nk; ++d) v += S[i * nk + d] * Wv[d * nk + c];
                    Z[(static_cast<size_t>(i) * nout + c) * B + b] = v;
                }
        }
        BH_TRY(dz.alloc(ctx, Z.size()));
        BH_HIP(ctx, hipMemcpyAsync(dz.p, Z.data(), Z.size() * 8, hipMemcpyHostToDevice, s));
        const unsigned grid = static_cast<unsigned>((vs + 255) / 256);
        if (nvf <= LZ_NVMAX) {
            hipLaunchKernelGGL((k_lz_combine<B>), dim3(grid), dim3(256), 0, s, Yp, vs, dz.p, nvf, nout);
            hipLaunchKernelGGL((k_lz_combine<B>), dim3(grid), dim3(256), 0, s, HYp, vs, dz.p, nvf, nout);
        } else {
            DevArray<double> tmp;
            double* tmpp = nullptr;
            BH_TRY(scratch(LZ_WS_C, tmp, static_cast<size_t>(nout) * vs, &tmpp));
            for (double* X : {Yp, HYp}) {
                hipLaunchKernelGGL((k_lz_combine_wide<B>), dim3(grid), dim3(256), 0, s, X, tmpp, vs, dz.p, nvf, nout);
                BH_HIP(ctx, hipMemcpyAsync(X, tmpp, static_cast<size_t>(nout) * vs * 8, hipMemcpyDeviceToDevice, s));
            }
            BH_HIP(ctx, hipStreamSynchronize(s));  // tmp is freed on return
        }
        BH_HIP(ctx, hipGetLastError());
        BH_HIP(ctx, hipStreamSynchronize(s));  // Z is freed on return
        return BH_OK;
    }
    struct Restart {
        int j = 0;
        double bnext = 0.0;
        std::vector<double> theta;
        std::vector<std::vector<double>> s;
    };

    // next step count at which lane b must be analysed
    int next_event(int b) const {
        int e = std::min(nextc[b], cap);
        e = std::min<int64_t>(e, static_cast<int64_t>(jl[b]) + std::max(0, maxsteps - tot[b]));
        return e;
    }
    // Speculative steps pay when they keep the device busy for a good part of a host analysis (a few
    // milliseconds): on for engines of at least LZ_SPEC_MIN_WORK vector elements (D B), whose steps take
    // hundreds of microseconds; smaller engines analyse at the scheduled steps only.  The rule depends on
    // the sizes alone, so the sequence of analyses and its results are reproducible.
    bool speculate() const { return lz_spec_steps() > 0 && vs >= LZ_SPEC_MIN_WORK; }
    // orthogonalisation variant of the next steps: the smallest of 0, 4, 8, 16 locked vectors that covers every
    // iterating lane's current set (a lane's set shrinks back to its frozen vectors at a deflation); none under
    // full reorthogonalisation, whose steps project out the frozen vectors themselves
    int step_lm() const {
        if (fullre) return 0;
        int mx = 0;
        for (int b = 0; b < B; ++b)
            if (act[b]) mx = std::max(mx, nlock[b]);
        return mx == 0 ? 0 : (mx <= 4 ? 4 : (mx <= 8 ? 8 : 16));
    }
    // K1 variants: 0 = staged row blocks; V + 16 u (V = 1, 2, 4, 8 vectors per thread, u = 0, 1, 2 for
    // U = 4, 8, 16 entries of a row in flight) = row-lane engine
    static int k1_V(int var) { return var & 15; }
    static int k1_U(int var) { return 4 << (var >> 4); }
    // rows of a K1 workgroup on the row-lane engine: LZ_K1_TILES passes of LZ_WG V / B rows, more when the
    // grid would exceed LZ_K1_GRID_MAX
    int64_t k1_rows(int var) const {
        return std::max<int64_t>(static_cast<int64_t>(LZ_WG * k1_V(var) / B) * LZ_K1_TILES,
                                 (D + LZ_K1_GRID_MAX - 1) / LZ_K1_GRID_MAX);
    }
    void set_k1(int var) {
        k1v = var;
        grid_k1 = var == 0 ? grid_rb : static_cast<int>((D + k1_rows(var) - 1) / k1_rows(var));
        rpw_so = (D + grid_k1 - 1) / grid_k1;
    }
    void launch_k1_var(int var, int nv) {
        if (var == 0) {
            hipLaunchKernelGGL((k_lz_a_rb<XT, B, FMT>), dim3(grid_rb), dim3(LZ_WG), shm_rb, s, a, rblk, nrblk, nv);
            return;
        }
        const int64_t rows = k1_rows(var);
        const dim3 g(static_cast<unsigned>((D + rows - 1) / rows));
#define K1W(VV, UU)                                                                                    \
    if (var == VV + 16 * (UU == 4 ? 0 : (UU == 8 ? 1 : 2))) {                                          \
        hipLaunchKernelGGL((k_lz_a_rlw<XT, B, VV, FMT, UU>), g, dim3(LZ_WG), shm, s, a, rows, nv);     \
        return;                                                                                        \
    }
        if constexpr (B >= 8 && sizeof(XT) == 4) K1W(8, 4)
        if constexpr (B >= 4 && B <= 32 && sizeof(XT) == 4) K1W(4, 8)
        if constexpr (B <= 16) {
            K1W(1, 8) K1W(1, 16)
            if constexpr (B >= 2) {
                K1W(2, 8) K1W(2, 16)
            }
        }
#undef K1W
        if constexpr (B >= 4 && sizeof(XT) * 4 <= 16) {
            if (var == 4) {
                hipLaunchKernelGGL((k_lz_a_rl<XT, B, 4, FMT>), g, dim3(LZ_WG), shm, s, a, rows, nv);
                return;
            }
        }
        if constexpr (B >= 2) {
            if (var == 2) {
                hipLaunchKernelGGL((k_lz_a_rl<XT, B, 2, FMT>), g, dim3(LZ_WG), shm, s, a, rows, nv);
                return;
            }
        }
        hipLaunchKernelGGL((k_lz_a_rl<XT, B, 1, FMT>), g, dim3(LZ_WG), shm, s, a, rows, nv);
    }
    void launch_k1(int LM) {
        const int nv = (1 + 2 * LM) * B;
        {
            KTimer kt(ctx, "lanczos_spmv", s);
            launch_k1_var(k1v, nv);
        }
        if (LM > 0) {
            KTimer kt(ctx, "lanczos_sodots", s);
            if (LM == 4) hipLaunchKernelGGL((k_lz_sodots<XT, B, 4>), dim3(grid_k1), dim3(LZ_WG), 0, s, a, rpw_so, nv);
            else if (LM == 8) hipLaunchKernelGGL((k_lz_sodots<XT, B, 8>), dim3(grid_k1), dim3(LZ_WG), 0, s, a, rpw_so, nv);
            else hipLaunchKernelGGL((k_lz_sodots<XT, B, 16>), dim3(grid_k1), dim3(LZ_WG), 0, s, a, rpw_so, nv);
        }
    }
    // Autotuner of K1, run once per (sector, precision, batch width) on the engine's first steps (the
    // result is kept in the sector's handle): every variant that fits is timed with HIP events over a
    // warm-up and LZ_TUNE_REPS launches on the admitted lanes (K1 writes only W and the partial sums, which
    // the real step overwrites), and the fastest is kept.  The variants differ in rows per wavefront
    // (64 V / B on the row-lane engine), in the entries of a row in flight per thread and in the staging of
    // the gathers; each (row, b) sums its entries in
    // storage order in all of them (the row-lane variants give identical sums, the staged engine rounds each
    // product before adding it).  BHX_TUNE=0 keeps the default variant.
    int tune_k1() {
        tuned = true;
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        const int key = bh_tune_key(0, MIXED ? BH_PREC_MIXED : BH_PREC_FP64, B);
        auto it = h->tuned.find(key);
        if (it != h->tuned.end()) {
            set_k1(it->second.variant);
            return BH_OK;
        }
        std::vector<int> cand;
        if (staged_ok && rblk) cand.push_back(0);
        cand.push_back(1);
        if (B >= 2) cand.push_back(2);
        if (B >= 4 && sizeof(XT) * 4 <= 16) cand.push_back(4);
        if (B >= 8 && sizeof(XT) == 4) cand.push_back(8);
        // more entries in flight per thread for long rows at small widths
        const double mean_row = D ? static_cast<double>(h->nnz) / static_cast<double>(D) : 0.0;
        if (B <= 16)
            for (int V = 1; V <= std::min(B, 2); V *= 2) {
                if (mean_row > 6.0) cand.push_back(V + 16);
                if (mean_row > 12.0) cand.push_back(V + 32);
            }
        if (B >= 4 && B <= 32 && sizeof(XT) == 4 && mean_row > 6.0) cand.push_back(4 + 16);
        BhTuned t;
        t.variant = k1v;
        if (bh_tune_enabled() && cand.size() > 1) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            BH_HIP(ctx, hipEventCreate(&e0));
            BH_HIP(ctx, hipEventCreate(&e1));
            double best = 0.0;
            for (int var : cand) {
                launch_k1_var(var, B);
                BH_HIP(ctx, hipEventRecord(e0, s));
                for (int r = 0; r < LZ_TUNE_REPS; ++r) launch_k1_var(var, B);
                BH_HIP(ctx, hipEventRecord(e1, s));
                BH_HIP(ctx, hipEventSynchronize(e1));
                float ms = 0.f;
                BH_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
                const double us = 1e3 * ms / LZ_TUNE_REPS;
                t.cand.push_back({var, us});
                if (best == 0.0 || us < best) {
                    best = us;
                    t.variant = var;
                }
            }
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
        }
        h->tuned[key] = t;
        set_k1(t.variant);
        return BH_OK;
    }
    void reduce_plain(int nwg, int nv, double* out) {
        KTimer kt(ctx, "lanczos_reduce", s);
        const int waves = LZ_RED_WG / 64;
        const int g = std::max(1, std::min((nv + waves - 1) / waves, 64));
        hipLaunchKernelGGL(k_lz_reduce, dim3(g), dim3(LZ_RED_WG), 0, s, part.p, nwg, nv, out);
    }
    void step(int LM) {
        launch_k1(LM);
        {
            KTimer kt(ctx, "lanczos_reduce", s);
            const int nv = (1 + 2 * LM) * B;
            hipLaunchKernelGGL((k_lz_reduce_wide<B>), dim3(nv), dim3(LZ_RED_WG), 0, s, st, part.p, grid_k1, nv, B, 0, st.sums);
        }
        {
            KTimer kt(ctx, "lanczos_update", s);
            if (LM == 0) hipLaunchKernelGGL((k_lz_b<XT, B, 0>), dim3(grid_k2), dim3(LZ_WG), 0, s, a);
            else if (LM == 4) hipLaunchKernelGGL((k_lz_b<XT, B, 4>), dim3(grid_k2), dim3(LZ_WG), 0, s, a);
            else if (LM == 8) hipLaunchKernelGGL((k_lz_b<XT, B, 8>), dim3(grid_k2), dim3(LZ_WG), 0, s, a);
            else hipLaunchKernelGGL((k_lz_b<XT, B, 16>), dim3(grid_k2), dim3(LZ_WG), 0, s, a);
        }
        if (fullre) {
            // two classical Gram-Schmidt passes against every stored Lanczos vector of the lane
            KTimer kt(ctx, "lanczos_reorth", s);
            for (int pass = 0; pass < 2; ++pass) {
                hipLaunchKernelGGL((k_lz_fr_dots<XT, B>), dim3(static_cast<unsigned>(cap + (mult ? ymax : 0))), dim3(LZ_WG), 0,
                                   s, a, dd.p);
                hipLaunchKernelGGL((k_lz_fr_axpy<XT, B>), dim3(grid_k2), dim3(LZ_WG), 0, s, a, dd.p, pass);
            }
        }
        {
            KTimer kt(ctx, "lanczos_reduce", s);
            const int nv2 = (1 + LM) * B;
            hipLaunchKernelGGL((k_lz_reduce_wide<B>), dim3(nv2), dim3(LZ_RED_WG), 0, s, st, part.p, grid_k2, nv2, B, 1,
                               sums2.p);
            hipLaunchKernelGGL((k_lz_beta<B>), dim3(1), dim3(LZ_RED_WG), 0, s, st, sums2.p, nv2, cap, eps_store,
                               lz_so_always());
        }
    }
    // Scratch of the finish, the deflations and the locks: a grow-only slot of the context workspace when the
    // engine uses it (batch solves: no allocation per call, and no hipFree, which waits for the device), else
    // the caller's own array (server engines).  LZ_WS_A: the finish's candidates / a deflation's Ritz vectors;
    // LZ_WS_B: the finish's H Y; LZ_WS_C: the Rayleigh-Ritz combination / a lock's Ritz vectors.
    template <typename T>
    int scratch(int slot, DevArray<T>& own, size_t n, T** out) {
        if (ws) {
            void* q = nullptr;
            BH_TRY(bh_ws_get(ctx, slot, n * sizeof(T), &q));
            *out = static_cast<T*>(q);
            return BH_OK;
        }
        BH_TRY(own.alloc(ctx, n));
        *out = own.p;
        return BH_OK;
    }
    // C holds rows t < max(njb) of [(t * nv + i) * B + b]
    template <typename OT>
    int form_vectors(const std::vector<int>& njb, int nv, const std::vector<double>& C, OT* out, int64_t os) {
        int njmax = 0;
        for (int b = 0; b < B; ++b) njmax = std::max(njmax, njb[b]);
        if (C.size() < static_cast<size_t>(njmax) * nv * B) return bh_set_error(ctx, BH_EINVAL, "ritz: coefficient rows");
        BH_TRY(dC.alloc(ctx, C.size()));
        BH_HIP(ctx, hipMemcpyAsync(dC.p, C.data(), C.size() * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dnj.p, njb.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
        if (lz_debug()) std::fprintf(stderr, "[lz] ritz njmax=%d outputs=%d out_bytes=%d\n", njmax, nv, static_cast<int>(sizeof(OT)));
        {
            KTimer kt(ctx, "lanczos_ritz", s);
            const int64_t per_wg = 256 * LZ_RZ_R;
            for (int i0 = 0; i0 < nv; i0 += LZ_RZ_G)
                hipLaunchKernelGGL((k_lz_ritz<XT, OT, B>), dim3(static_cast<unsigned>((vs + per_wg - 1) / per_wg)), dim3(256), 0,
                                   s, V, vs, dnj.p, njmax, dC.p, nv, i0, std::min(LZ_RZ_G, nv - i0), out, os);
        }
        BH_HIP(ctx, hipGetLastError());
        BH_HIP(ctx, hipStreamSynchronize(s));
        return BH_OK;
    }
    // restarts the flagged lanes (plans made by the analysis): kept Ritz vectors in the tridiagonal basis of
    // arrow_to_tridiag in slots 0..p-1, the normalised next Lanczos vector in slot p, T rewritten to match.
    // Rows 0..p of alpha, beta and scale are written back for every lane, so they are read from the device
    // here, after the speculative steps: another lane early in a pass (just admitted or deflated) has written
    // some of those rows since the analysis copied them, and a stale copy would overwrite its recurrence.
    int restart(const std::vector<int>& flag) {
        int jmax = 0, nv = 0, pmax = 0;
        for (int b = 0; b < B; ++b)
            if (flag[b]) {
                jmax = std::max(jmax, plan[b].j);
                nv = std::max(nv, static_cast<int>(plan[b].theta.size()) + 1);
                pmax = std::max(pmax, static_cast<int>(plan[b].theta.size()));
            }
        std::vector<double> hsc(static_cast<size_t>(std::max(jmax, pmax) + 1) * B);
        std::vector<double> fa(static_cast<size_t>(std::max(pmax, 1)) * B), fb(static_cast<size_t>(pmax + 1) * B);
        std::vector<int> hstep(B);
        BH_HIP(ctx, hipMemcpyAsync(hsc.data(), st.scale, hsc.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hstep.data(), st.step, B * 4, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(fa.data(), st.alpha, fa.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(fb.data(), st.beta, fb.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        std::vector<double> C(static_cast<size_t>(jmax + 1) * nv * B, 0.0);
        std::vector<int> nin(B, 0), nout(B, 0);
        for (int b = 0; b < B; ++b) {
            if (!flag[b]) continue;
            const Restart& P = plan[b];
            const int p = static_cast<int>(P.theta.size()), j = P.j;
            std::vector<double> sig(p), Q(static_cast<size_t>(p) * p), ta(p), tb(p);
            double ns = 0.0;
            for (int i = 0; i < p; ++i) {
                sig[i] = P.bnext * P.s[i][j - 1];
                ns += sig[i] * sig[i];
            }
            ns = std::sqrt(ns);
            bhtri::arrow_to_tridiag(p, P.theta.data(), sig.data(), Q.data(), ta.data(), tb.data());
            for (int t = 0; t < j; ++t) {
                const double sc = hsc[static_cast<size_t>(t) * B + b];
                for (int c = 0; c < p; ++c) {
                    double v = 0.0;
                    for (int i = 0; i < p; ++i) v += P.s[i][t] * Q[i + static_cast<size_t>(p) * c];
                    C[(static_cast<size_t>(t) * nv + c) * B + b] = v * sc;
                }
            }
            C[(static_cast<size_t>(j) * nv + p) * B + b] = hsc[static_cast<size_t>(j) * B + b];
            nin[b] = j + 1;
            nout[b] = p + 1;
            // T of the restarted lane, in the fresh copies of rows 0..p
            for (int c = 0; c < p; ++c) fa[static_cast<size_t>(c) * B + b] = ta[c];
            fb[b] = 0.0;
            for (int c = 1; c < p; ++c) fb[static_cast<size_t>(c) * B + b] = tb[c];
            fb[static_cast<size_t>(p) * B + b] = ns;
            for (int c = 0; c <= p; ++c) hsc[static_cast<size_t>(c) * B + b] = 1.0;
            hstep[b] = p;
            jl[b] = p;
            kept[b] = p;
            nextc[b] = p + chunk;
            if (lz_debug()) std::fprintf(stderr, "[lz] restart lane=%d j=%d tot=%d keep=%d sigma=%.3e\n", b, j, tot[b], p, ns);
        }
        BH_TRY(dC.alloc(ctx, C.size()));
        BH_HIP(ctx, hipMemcpyAsync(dC.p, C.data(), C.size() * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dnj.p, nin.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dno.p, nout.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
        {
            KTimer kt(ctx, "lanczos_restart", s);
            hipLaunchKernelGGL((k_lz_restart<XT, B>), dim3(static_cast<unsigned>((vs + 255) / 256)), dim3(256), 0, s, V, Xc, vs,
                               dnj.p, dno.p, dC.p, nv);
        }
        BH_HIP(ctx, hipGetLastError());
        BH_HIP(ctx, hipMemcpyAsync(st.alpha, fa.data(), static_cast<size_t>(pmax) * B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.beta, fb.data(), static_cast<size_t>(pmax + 1) * B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.scale, hsc.data(), static_cast<size_t>(pmax + 1) * B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.step, hstep.data(), B * 4, hipMemcpyHostToDevice, s));
        // the first two steps after a restart orthogonalise against the locked vectors
        const int so2[2] = {1, 0};
        BH_HIP(ctx, hipMemcpyAsync(st.so, so2, 2 * 4, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        return BH_OK;
    }
    // forms the Ritz vectors new_locks[b] and stores them in Y: appended (lock_slot -1) or replacing slot
    int lock(const std::vector<std::vector<int>>& new_locks, const std::vector<std::vector<int>>& lock_slot,
             const std::vector<Analysis>& ans) {
        const auto tl0 = std::chrono::steady_clock::now();
        struct Clock {
            double& acc;
            std::chrono::steady_clock::time_point t;
            ~Clock() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count(); }
        } clk{prof.lock_ms, tl0};
        int nv = 0, jmax = 0;
        for (int b = 0; b < B; ++b)
            if (!new_locks[b].empty()) {
                nv = std::max(nv, static_cast<int>(new_locks[b].size()));
                jmax = std::max(jmax, jl[b]);
            }
        nv = std::min(nv, LZ_NVMAX);
        std::vector<double> hsc(static_cast<size_t>(jmax) * B);
        BH_HIP(ctx, hipMemcpyAsync(hsc.data(), st.scale, hsc.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        std::vector<double> C(static_cast<size_t>(jmax) * nv * B, 0.0);
        std::vector<int> njb(B, 0);
        for (int b = 0; b < B; ++b) {
            if (new_locks[b].empty()) continue;
            njb[b] = static_cast<int>(ans[b].s[new_locks[b][0]].size());  // the analysed step count
            for (int i = 0; i < nv && i < static_cast<int>(new_locks[b].size()); ++i) {
                const std::vector<double>& sv = ans[b].s[new_locks[b][i]];
                for (int t = 0; t < static_cast<int>(sv.size()) && t < njb[b]; ++t)
                    C[(static_cast<size_t>(t) * nv + i) * B + b] = sv[t] * hsc[static_cast<size_t>(t) * B + b];
            }
        }
        DevArray<XT> tmp;
        XT* tmpp = nullptr;
        BH_TRY(scratch(LZ_WS_C, tmp, static_cast<size_t>(nv) * vs, &tmpp));
        BH_TRY(form_vectors(njb, nv, C, tmpp, vs));
        // normalise: ||y||^2 per (i, b)
        hipLaunchKernelGGL((k_lz_dots<XT, B>), dim3(grid_k2), dim3(LZ_WG), 0, s, tmpp, nullptr, D, nv, 2, nullptr, part.p);
        reduce_plain(grid_k2, nv * B, dnorm.p);
        std::vector<double> nrm(static_cast<size_t>(nv) * B);
        BH_HIP(ctx, hipMemcpyAsync(nrm.data(), dnorm.p, nrm.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (auto& v : nrm) v = v > 0 ? 1.0 / std::sqrt(v) : 0.0;
        BH_HIP(ctx, hipMemcpyAsync(dnorm.p, nrm.data(), nrm.size() * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_lz_scale<XT, B>), dim3(static_cast<unsigned>((vs * nv + 255) / 256)), dim3(256), 0, s, tmpp, vs,
                           nv, dnorm.p);
        // vector i of lane b goes to Y slot nlock[b] + i (strided copy of one interleaved column); its Ritz
        // value and a fresh orthogonality estimate go to the device state
        std::vector<double> lth(static_cast<size_t>(LZ_LMAX) * B), tau(static_cast<size_t>(2 * LZ_LMAX) * B), anv(B);
        BH_HIP(ctx, hipMemcpyAsync(lth.data(), st.ltheta, lth.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(tau.data(), st.tau, tau.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(anv.data(), st.anorm, anv.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        std::vector<int> dst(static_cast<size_t>(nv) * B, -1);  // Y slot of vector i of lane b
        for (int b = 0; b < B; ++b) {
            int appended = 0;
            for (int i = 0; i < static_cast<int>(new_locks[b].size()) && i < nv; ++i) {
                const int target = lock_slot[b][i];
                if (target < 0 && nlock[b] + appended >= LZ_LMAX) continue;
                const int slot = target >= 0 ? target : nlock[b] + appended++;
                dst[static_cast<size_t>(i) * B + b] = slot;
                const double th = ans[b].theta[new_locks[b][i]];
                const double bd = ans[b].bound[new_locks[b][i]];
                if (target < 0) {
                    rh[b].lock_theta.push_back(th);
                    rh[b].lock_bound.push_back(bd);
                } else {
                    rh[b].lock_theta[slot] = th;
                    rh[b].lock_bound[slot] = bd;
                }
                lth[static_cast<size_t>(slot) * B + b] = th;
                tau[static_cast<size_t>(slot) * B + b] = eps_store;
                tau[static_cast<size_t>(LZ_LMAX + slot) * B + b] = eps_store;
            }
            if (!new_locks[b].empty()) anv[b] = ans[b].tnorm;
            nlock[b] += appended;
        }
        // vector i of lane b goes to Y slot dst[i][b] (one interleaved column each), in one launch
        BH_HIP(ctx, hipMemcpyAsync(lock_map.p, dst.data(), dst.size() * sizeof(int), hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_lz_put_locked<XT, B>), dim3(static_cast<unsigned>((vs + 255) / 256)), dim3(256), 0, s, tmpp,
                           Y, vs, nv, lock_map.p);
        const int so2[2] = {1, 0};  // orthogonalise the next two Lanczos vectors against the locked set
        BH_HIP(ctx, hipMemcpyAsync(st.ltheta, lth.data(), lth.size() * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.tau, tau.data(), tau.size() * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.anorm, anv.data(), anv.size() * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.so, so2, 2 * 4, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.nlock, nlock.data(), B * 4, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        return BH_OK;
    }

    // Z (FP64 [D][B]) -= its components along Y slots 0 .. cnt[b] - 1 of each lane (two classical Gram-Schmidt
    // passes, FP64 dot products); lanes with cnt[b] < 0 untouched.  cnt is a host array.
    int ortho_frozen(double* Z, const std::vector<int>& cnt) {
        int lmax = 0;
        for (int b = 0; b < B; ++b) lmax = std::max(lmax, cnt[b]);
        if (lmax <= 0) return BH_OK;
        BH_HIP(ctx, hipMemcpyAsync(dnj.p, cnt.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
        for (int pass = 0; pass < 2; ++pass) {
            hipLaunchKernelGGL((k_lz_ydots<XT, B>), dim3(grid_k2), dim3(LZ_WG), 0, s, Y, Z, D, dnj.p, lmax, part.p);
            reduce_plain(grid_k2, lmax * B, dnorm.p);
            hipLaunchKernelGGL((k_lz_yaxpy<XT, B>), dim3(static_cast<unsigned>((vs + 255) / 256)), dim3(256), 0, s, Y, Z, vs,
                               dnj.p, dnorm.p);
        }
        BH_HIP(ctx, hipGetLastError());
        return BH_OK;
    }
    // 1 / ||Z[.][b]|| per lane (FP64 [D][B] or storage-precision vector), host result
    template <typename ZT>
    int inv_norms(const ZT* Z, std::vector<double>& f) {
        hipLaunchKernelGGL((k_lz_dots<ZT, B>), dim3(grid_k2), dim3(LZ_WG), 0, s, Z, nullptr, D, 1, 2, nullptr, part.p);
        reduce_plain(grid_k2, B, dnorm.p);
        f.assign(B, 0.0);
        BH_HIP(ctx, hipMemcpyAsync(f.data(), dnorm.p, B * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (auto& v : f) v = v > 0 ? 1.0 / std::sqrt(v) : 0.0;
        return BH_OK;
    }
    // Deflation restart of the flagged lanes (plans made by the analysis, lanes rolled back to the analysed
    // step): the Ritz vectors of the levels in dplan[b] are formed in FP64, orthonormalised against the
    // lane's frozen vectors and among themselves, and stored in Y after them; the lane then starts its next
    // pass at slot 0 from the Park-Miller sequence continued past the values of its earlier start vectors,
    // orthogonalised against every frozen vector.  Under selective orthogonalisation the frozen vectors are
    // the lane's locked set at the start of the pass (the pass's own locks are appended after them); under
    // full reorthogonalisation every step projects them out (k_lz_fr_dots / k_lz_fr_axpy).
    int deflate(const std::vector<int>& flag) {
        struct Clock {
            double& acc;
            std::chrono::steady_clock::time_point t;
            ~Clock() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count(); }
        } clk{prof.deflate_ms, std::chrono::steady_clock::now()};
        int nv = 0, jmax = 0;
        for (int b = 0; b < B; ++b)
            if (flag[b]) {
                nv = std::max(nv, static_cast<int>(dplan[b].theta.size()));
                jmax = std::max(jmax, static_cast<int>(dplan[b].s[0].size()));
            }
        std::vector<double> hsc(static_cast<size_t>(jmax) * B);
        BH_HIP(ctx, hipMemcpyAsync(hsc.data(), st.scale, hsc.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        std::vector<double> C(static_cast<size_t>(jmax) * nv * B, 0.0);
        std::vector<int> njb(B, 0);
        for (int b = 0; b < B; ++b) {
            if (!flag[b]) continue;
            njb[b] = static_cast<int>(dplan[b].s[0].size());
            for (int i = 0; i < static_cast<int>(dplan[b].s.size()); ++i) {
                const std::vector<double>& sv = dplan[b].s[i];
                for (int t = 0; t < njb[b]; ++t) C[(static_cast<size_t>(t) * nv + i) * B + b] = sv[t] * hsc[static_cast<size_t>(t) * B + b];
            }
        }
        DevArray<double> Zf;
        double* Zfp = nullptr;
        BH_TRY(scratch(LZ_WS_A, Zf, static_cast<size_t>(nv) * vs, &Zfp));
        BH_TRY(form_vectors(njb, nv, C, Zfp, vs));
        const unsigned g256 = static_cast<unsigned>((vs + 255) / 256);
        // drop the frozen vectors the plans do not keep: kept slots move down, in ascending order
        {
            std::vector<int> src(static_cast<size_t>(ymax) * B, -1);
            int ns = 0;
            for (int b = 0; b < B; ++b) {
                if (!flag[b]) continue;
                const std::vector<int>& kp = dplan[b].keep;
                std::vector<double> th;
                for (size_t q = 0; q < kp.size(); ++q) {
                    src[q * B + b] = kp[q];
                    th.push_back(rh[b].fz_theta[kp[q]]);
                }
                ns = std::max(ns, static_cast<int>(kp.size()));
                rh[b].fz_theta = th;
                rh[b].nfz = static_cast<int>(kp.size());
            }
            if (ns > 0) {
                BH_HIP(ctx, hipMemcpyAsync(lock_map.p, src.data(), static_cast<size_t>(ns) * B * sizeof(int), hipMemcpyHostToDevice, s));
                hipLaunchKernelGGL((k_lz_compact<XT, B>), dim3(g256), dim3(256), 0, s, Y, vs, lock_map.p, ns);
                BH_HIP(ctx, hipGetLastError());
                BH_HIP(ctx, hipStreamSynchronize(s));
            }
        }
        std::vector<int> cnt(B), slot(B);
        std::vector<double> f;
        // new frozen vector i of each lane: orthogonal to the lane's frozen vectors so far, then normalised
        for (int i = 0; i < nv; ++i) {
            for (int b = 0; b < B; ++b) {
                const bool on = flag[b] && i < static_cast<int>(dplan[b].theta.size());
                cnt[b] = on ? rh[b].nfz + i : -1;
                slot[b] = on ? rh[b].nfz + i : -1;
            }
            double* Zi = Zfp + static_cast<int64_t>(i) * vs;
            BH_TRY(ortho_frozen(Zi, cnt));
            BH_TRY(inv_norms(static_cast<const double*>(Zi), f));
            BH_HIP(ctx, hipMemcpyAsync(lock_map.p, slot.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipMemcpyAsync(dnorm.p, f.data(), B * 8, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL((k_lz_defl_store<XT, B>), dim3(g256), dim3(256), 0, s, Zi, Y, vs, lock_map.p, dnorm.p,
                               static_cast<XT*>(nullptr));
            BH_HIP(ctx, hipStreamSynchronize(s));  // the host arrays are reused by the next vector
        }
        // fresh start vectors in W (free between steps), orthogonal to every frozen vector
        for (int b = 0; b < B; ++b) {
            cnt[b] = flag[b] ? rh[b].nfz + static_cast<int>(dplan[b].theta.size()) : -1;
            slot[b] = flag[b] ? 0 : -1;
            if (!flag[b]) continue;
            const uint64_t p = 2147483647ull;
            const uint64_t x0 = o.seed ? (o.seed & p) : 1;
            hipLaunchKernelGGL((k_lz_pm_fill<B>), dim3(static_cast<unsigned>((D + 255) / 256)), dim3(256), 0, s, W, D, b, x0,
                               static_cast<uint64_t>(rh[b].pass + 1) * static_cast<uint64_t>(D),
                               h->lex_of_pos);
        }
        BH_TRY(ortho_frozen(W, cnt));
        BH_TRY(inv_norms(static_cast<const double*>(W), f));
        BH_HIP(ctx, hipMemcpyAsync(lock_map.p, slot.data(), B * sizeof(int), hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dnorm.p, f.data(), B * 8, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL((k_lz_defl_store<XT, B>), dim3(g256), dim3(256), 0, s, W, V, vs, lock_map.p, dnorm.p, Xc);
        BH_HIP(ctx, hipGetLastError());
        // the inverse norm of the stored start vector (its rounding to the storage precision included)
        std::vector<double> sc0;
        BH_TRY(inv_norms(static_cast<const XT*>(V), sc0));
        // lane state: the next pass starts at step 0 with the frozen set as its orthogonality constraint
        std::vector<double> hs(B), hbeta(B), lth(static_cast<size_t>(LZ_LMAX) * B), tau(static_cast<size_t>(2 * LZ_LMAX) * B);
        std::vector<int> hstep(B), hnl(B);
        BH_HIP(ctx, hipMemcpyAsync(hs.data(), st.scale, B * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hbeta.data(), st.beta, B * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hstep.data(), st.step, B * 4, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(hnl.data(), st.nlock, B * 4, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(lth.data(), st.ltheta, lth.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipMemcpyAsync(tau.data(), st.tau, tau.size() * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (int b = 0; b < B; ++b) {
            if (!flag[b]) continue;
            RunHost& R = rh[b];
            for (double th : dplan[b].theta) R.fz_theta.push_back(th);
            R.nfz += static_cast<int>(dplan[b].theta.size());
            R.pass += 1;
            R.alpha.clear();
            R.beta.clear();
            R.prev_tot = -1;
            R.lock_theta = R.fz_theta;
            R.lock_bound.assign(R.nfz, 0.0);
            hs[b] = sc0[b];
            hbeta[b] = 0.0;
            hstep[b] = 0;
            hnl[b] = R.nfz;
            nlock[b] = R.nfz;
            jl[b] = 0;
            kept[b] = 0;
            nextc[b] = std::max(2 * k + 8, 20);
            if (!fullre) {
                for (int l = 0; l < R.nfz && l < LZ_LMAX; ++l) {
                    lth[static_cast<size_t>(l) * B + b] = R.fz_theta[l];
                    tau[static_cast<size_t>(l) * B + b] = eps_store;
                    tau[static_cast<size_t>(LZ_LMAX + l) * B + b] = eps_store;
                }
            }
            if (lz_debug())
                std::fprintf(stderr, "[lz] deflate lane=%d tot=%d pass=%d frozen=%d\n", b, tot[b], R.pass, R.nfz);
        }
        BH_HIP(ctx, hipMemcpyAsync(st.scale, hs.data(), B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.beta, hbeta.data(), B * 8, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.step, hstep.data(), B * 4, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(st.nlock, hnl.data(), B * 4, hipMemcpyHostToDevice, s));
        if (!fullre) {
            const int so2[2] = {1, 0};  // the pass's first two steps orthogonalise exactly against the frozen set
            BH_HIP(ctx, hipMemcpyAsync(st.ltheta, lth.data(), lth.size() * 8, hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipMemcpyAsync(st.tau, tau.data(), tau.size() * 8, hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipMemcpyAsync(st.so, so2, 2 * 4, hipMemcpyHostToDevice, s));
        }
        BH_HIP(ctx, hipStreamSynchronize(s));
        return BH_OK;
    }

    bh_ctx* ctx;
    bh_ham* h;
    hipStream_t s;
    bh_lanczos_opts o;
    bool ws;
    int64_t D = 0, vs = 0;
    int k = 4, maxsteps = 10000, slots = 2, cap = 1, chunk = 16;
    bool fullre = false;  // every step reorthogonalises against the whole basis (small sectors)
    bool use_graphs = true;
    std::vector<Graph> graphs;
    double tol = 1e-10, dedup = 1e-9, eps_store = 0, eps23 = 0, scale0 = 0;
    XT* V = nullptr;
    XT* Y = nullptr;
    XT* Xc = nullptr;
    double* W = nullptr;
    double* r0 = nullptr;
    DevArray<XT> V_own, Y_own, Xc_own;
    DevArray<double> W_own, r0_own, st_d, part, sums2, dC, dnorm, dd;
    DevArray<int> st_i, dnj, dno, lock_map;
    size_t n_alpha = 0, n_beta = 0, bytes = 0;
    LzState st{};
    LzArgs<XT> a{};
    int k1v = 1, grid_rb = 1, grid_rl = 1, grid_k1 = 1, grid_k2 = 1;
    bool staged_ok = false, tuned = false;
    int64_t rpw = 1, rpw_so = 1, nrblk = 0;
    const int32_t* rblk = nullptr;
    size_t shm = 0, shm_rb = 0;
    std::vector<RunHost> rh;
    std::vector<int> act, nlock, jl, tot, kept, nextc;
    std::vector<int> jsync, totsync;  // step counts at the last analysis' device sync
    std::vector<int> jan, tot_an;     // step counts of each lane's last analysis
    std::vector<int> step_h;          // host copy of the step counters written at a rollback
    std::vector<double> Jb, Ub, hlo, hhi, ha, hb;
    double an0[B] = {};
    Restart plan[B];
    // deflation restart of a lane: the levels it freezes (Ritz values and their eigenvectors of T at the analysed
    // step) and its frozen slots that stay (ascending)
    struct Deflation {
        std::vector<double> theta;
        std::vector<std::vector<double>> s;
        std::vector<int> keep;
    };
    Deflation dplan[B];
    std::vector<int> waiting;   // lanes whose pass is complete, waiting for the deflation barrier
    bool defl_barrier = false;  // batch solves: lanes deflate together once none is iterating
    bool mult = true;  // eigenvalues with multiplicity (deflation passes); false: distinct levels
    int ymax = LZ_LMAX;
};

// an engine of B lanes (B rounded up to a power of two) in the requested precision
std::unique_ptr<LzEngineBase> bh_make_engine(bh_ctx* ctx, bh_ham* h, int prec, int B, const bh_lanczos_opts& o,
                                             hipStream_t s, bool use_ctx_ws, int* rc) {
    std::unique_ptr<LzEngineBase> e;
    *rc = BH_OK;
    if (bh_lz_fullre(h->D, o)) prec = BH_PREC_FP64;
#define MK(XT, BB)                                                                  \
    {                                                                               \
        auto* p = new LzEngine<XT, BB>(ctx, h, o, s, use_ctx_ws);                   \
        e.reset(p);                                                                 \
        *rc = p->setup();                                                           \
    }
#define MKB(XT)                   \
    if (B <= 1) MK(XT, 1)         \
    else if (B <= 2) MK(XT, 2)    \
    else if (B <= 4) MK(XT, 4)    \
    else if (B <= 8) MK(XT, 8)    \
    else if (B <= 16) MK(XT, 16)  \
    else if (B <= 32) MK(XT, 32)  \
    else MK(XT, 64)
    if (prec == BH_PREC_FP64) MKB(double) else MKB(float)
#undef MKB
#undef MK
    if (*rc != BH_OK) return nullptr;
    return e;
}

size_t bh_engine_bytes(int64_t D, int prec, int B, const bh_lanczos_opts& o) {
    int Bp = 1;
    while (Bp < B) Bp <<= 1;
    const int slots = lz_slots(D, o);
    const bool fr = bh_lz_fullre(D, o);
    const size_t es = (prec == BH_PREC_FP64 || fr) ? 8 : 4;
    const size_t vs = static_cast<size_t>(D) * Bp;
    const int64_t rpw = std::max<int64_t>(static_cast<int64_t>(LZ_WG / Bp) * LZ_K1_TILES, (D + LZ_K1_GRID_MAX - 1) / LZ_K1_GRID_MAX);
    const int ymax = lz_ymax(D, o);
    const size_t row_bytes = static_cast<size_t>(std::max(2 * LZ_NVMAX + 1 + 2 * LZ_LMAX, ymax)) * Bp * 8;
    // partial-sum rows: the row-lane and K2 grids, or the staged K1's (up to LZ_PART_BYTES)
    const size_t part_rows = std::max<size_t>(
        static_cast<size_t>(std::max<int64_t>({LZ_GRID_MAX, LZ_K2_GRID, (D + rpw - 1) / rpw})),
        std::min<size_t>(LZ_K1_GRID_MAX, static_cast<size_t>(LZ_PART_BYTES) / row_bytes));
    return (static_cast<size_t>(slots) + ymax + 1) * vs * es + vs * 8 + static_cast<size_t>(D) * 8 +
           part_rows * row_bytes +
           (static_cast<size_t>(slots) + 2) * Bp * 24 + static_cast<size_t>(3 * LZ_LMAX + 1) * Bp * 8 +
           (fr ? static_cast<size_t>(slots + ymax) * Bp * 8 : 0);
}

// Options of the FP64 rerun of a rejected mixed solve: the plain FP64 solver (selective
// orthogonalisation above fullre_d).  At D = 92,378 it reran 11 lanes in 120 ms where full
// reorthogonalisation took 1.9 s, with the same levels.
bh_lanczos_opts bh_fallback_opts(const bh_lanczos_opts& o) {
    bh_lanczos_opts f = o;
    f.prec = BH_PREC_FP64;
    return f;
}

// Options of the last resort, for FP64 reruns whose levels are not certified either: full
// reorthogonalisation whatever the sector size over a thick-restarted basis of at most 160 vectors
// (dependable on clusters of close levels, where selective orthogonalisation can let copies through).
bh_lanczos_opts bh_fallback_full_opts(const bh_lanczos_opts& o) {
    bh_lanczos_opts f = o;
    f.prec = BH_PREC_FP64;
    f.fullre_d = std::numeric_limits<int>::max();
    f.ncv = o.ncv > 0 ? std::min(o.ncv, 160) : 160;
    return f;
}

bool bh_lane_accepted(int status, const double* ev, const double* ee, int k, double accept_rel) {
    const double acc = accept_rel > 0 ? accept_rel : 1e-9;
    if (status != BH_OK || k < 1 || !std::isfinite(ev[0]) || !(ee[0] <= acc * std::fmax(std::fabs(ev[0]), 1e-300)))
        return false;
    // levels above E0: ||r||^2/gap within LZ_ACCEPT_LEVEL of the level's scale, its magnitude or the
    // width of the wanted window E_top - E0 if larger (levels near zero are not held to an absolute 0)
    double top = ev[0];
    for (int i = 1; i < k; ++i)
        if (std::isfinite(ev[i])) top = std::fmax(top, ev[i]);
    const double width = top - ev[0];
    for (int i = 1; i < k; ++i) {
        if (!std::isfinite(ev[i])) continue;
        if (!(ee[i] <= LZ_ACCEPT_LEVEL * std::fmax(std::fabs(ev[i]), width))) return false;
    }
    return true;
}

// batch solve of nreal lanes to completion on one engine (the body of bh_lanczos)
static int lanczos_batch(bh_ctx* ctx, bh_ham* h, int prec, const double* J, const double* U, int nreal,
                         const bh_lanczos_opts& o, bh_lanczos_result* res, std::vector<double>* vecs) {
    int rc = BH_OK;
    const auto ts0 = std::chrono::steady_clock::now();
    const int64_t na0 = ctx->n_alloc, nf0 = ctx->n_free;
    const double am0 = ctx->alloc_ms, fm0 = ctx->free_ms;
    auto e = bh_make_engine(ctx, h, prec, nreal, o, ctx->stream, true, &rc);
    if (!e) return rc;
    // the lanes of one batch finish together: their deflation passes start together too (one formation of the
    // frozen vectors for all of them; a lane's arithmetic does not depend on when its pass starts)
    e->set_deflation_barrier(!(std::getenv("BHX_DEFL_BARRIER") && std::strcmp(std::getenv("BHX_DEFL_BARRIER"), "0") == 0));
    for (int b = 0; b < nreal; ++b) BH_TRY(e->admit(b, J[b], U[b]));
    e->prof.setup_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts0).count();
    std::vector<int> fin, all;
    while (e->active() > 0) {
        const auto tr0 = std::chrono::steady_clock::now();
        BH_TRY(e->run(e->suggest_chunk()));
        e->prof.run_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tr0).count();
        BH_TRY(e->check(fin));
    }
    for (int b = 0; b < nreal; ++b) all.push_back(b);
    std::vector<LaneResult> out;
    const auto tf0 = std::chrono::steady_clock::now();
    BH_TRY(e->finish(all, out, vecs != nullptr));
    e->prof.finish_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tf0).count();
    const bool used_fp64 = e->fp64();
    static const bool prof_on = [] {
        const char* v = std::getenv("BHX_LZ_PROFILE");
        return v && std::strcmp(v, "1") == 0;
    }();
    if (prof_on) {
        const LzProfile pr = e->prof;
        const bool f64 = used_fp64;
        const double total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts0).count();
        const auto td0 = std::chrono::steady_clock::now();
        e.reset();
        const double teardown = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - td0).count();
        std::fprintf(stderr,
                     "[lz-profile] D=%lld lanes=%d fp64=%d checks=%d setup=%.1f ms run=%.1f ms (capture %.1f) wait=%.1f ms "
                     "host=%.1f ms (analysis %.1f, deflate %.1f, lock %.1f) finish=%.1f ms (refinement steps %d, lane-steps %d) "
                     "total=%.1f ms teardown=%.1f ms hipMalloc %lld (%.1f ms) hipFree %lld (%.1f ms)\n",
                     static_cast<long long>(h->D), nreal, f64 ? 1 : 0, pr.checks, pr.setup_ms, pr.run_ms, pr.capture_ms,
                     pr.wait_ms, pr.host_ms, pr.analysis_ms, pr.deflate_ms, pr.lock_ms, pr.finish_ms, pr.refine_steps,
                     pr.refined_lanes, total, teardown,
                     static_cast<long long>(ctx->n_alloc - na0), ctx->alloc_ms - am0,
                     static_cast<long long>(ctx->n_free - nf0), ctx->free_ms - fm0);
    }
    const int k = o.k;
    if (vecs) vecs->assign(static_cast<size_t>(nreal) * k * h->D, 0.0);
    for (int b = 0; b < nreal; ++b) {
        for (int i = 0; i < k; ++i) {
            if (res->evals) res->evals[b * k + i] = out[b].evals[i];
            if (res->resid) res->resid[b * k + i] = out[b].resid[i];
            if (res->errest) res->errest[b * k + i] = out[b].errest[i];
        }
        if (res->iters) res->iters[b] = out[b].iters;
        if (res->status) res->status[b] = out[b].status;
        if (res->used_fp64) res->used_fp64[b] = used_fp64 ? 1 : 0;
        if (vecs && !out[b].vecs.empty())
            std::copy(out[b].vecs.begin(), out[b].vecs.end(), vecs->begin() + static_cast<int64_t>(b) * k * h->D);
    }
    return BH_OK;
}

extern "C" {

void bh_lanczos_default_opts(bh_lanczos_opts* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->k = 4;
    o->prec = BH_PREC_MIXED;
    o->max_iter = 10000;
    o->chunk = 16;
    o->tol = 1e-10;
    o->dedup_rel = 1e-9;
    o->accept_rel = 1e-9;
    o->seed = 0;
    o->fallback = 1;
    o->want_vectors = 0;
    o->ncv = 0;
    o->fullre_d = 0;
    o->distinct = 0;
}

namespace {

// J = 0: H = U V is diagonal; its levels are the distinct values of U V.  Ritz vectors follow the
// Krylov convention: the start vector's component in each level's eigenspace, normalised.
int solve_diagonal(bh_ctx* ctx, bh_ham* h, const std::vector<int>& runs, const double* U, const bh_lanczos_opts& o,
                   bh_lanczos_result* res, std::vector<double>* vecs) {
    if (runs.empty()) return BH_OK;  // (the diagonal's copy to the host below took ~95 ms at D = 1.35e6)
    const int k = o.k;
    std::vector<float> d(static_cast<size_t>(h->D));
    BH_HIP(ctx, hipMemcpyAsync(d.data(), h->diag, d.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    if (h->vlevels.empty()) {
        std::vector<double> v(d.begin(), d.end());
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        h->vlevels = v;
    }
    std::vector<double> r;
    if (vecs && o.distinct) BH_TRY(lz_start_vector(ctx, h, o.seed, r));
    // with multiplicity: the k lowest diagonal entries, ties in position order, eigenvectors e_q
    std::vector<int64_t> order;
    if (!o.distinct) {
        order.resize(static_cast<size_t>(h->D));
        for (int64_t q = 0; q < h->D; ++q) order[q] = q;
        std::stable_sort(order.begin(), order.end(), [&](int64_t x, int64_t y) { return d[x] < d[y]; });
    }
    for (int b : runs) {
        if (!o.distinct) {
            // U V is diagonal: its eigenvalues are the entries U V_q (for U < 0 the largest V come first)
            std::vector<int64_t> ord(order);
            if (U[b] < 0) std::stable_sort(ord.begin(), ord.end(), [&](int64_t x, int64_t y) { return d[x] > d[y]; });
            const int nk = static_cast<int>(std::min<int64_t>(k, h->D));
            for (int i = 0; i < k; ++i) {
                const double e = i < nk ? U[b] * static_cast<double>(d[ord[i]]) : NAN;
                if (res->evals) res->evals[b * k + i] = e;
                if (res->resid) res->resid[b * k + i] = i < nk ? 0.0 : NAN;
                if (res->errest) res->errest[b * k + i] = i < nk ? 0.0 : NAN;
                if (vecs && i < nk) {
                    double* y = vecs->data() + (static_cast<size_t>(b) * k + i) * h->D;
                    std::fill(y, y + h->D, 0.0);
                    y[ord[i]] = 1.0;
                }
            }
            if (res->iters) res->iters[b] = 0;
            if (res->used_fp64) res->used_fp64[b] = o.prec == BH_PREC_FP64 ? 1 : 0;
            if (res->status) res->status[b] = nk >= k ? BH_OK : BH_ENOCONV;
            continue;
        }
        std::vector<double> lv;
        if (U[b] == 0.0) lv.push_back(0.0);
        else
            for (double v : h->vlevels) lv.push_back(U[b] * v);
        std::sort(lv.begin(), lv.end());
        for (int i = 0; i < k; ++i) {
            if (res->evals) res->evals[b * k + i] = i < static_cast<int>(lv.size()) ? lv[i] : NAN;
            if (res->resid) res->resid[b * k + i] = i < static_cast<int>(lv.size()) ? 0.0 : NAN;
            if (res->errest) res->errest[b * k + i] = i < static_cast<int>(lv.size()) ? 0.0 : NAN;
        }
        if (res->iters) res->iters[b] = 0;
        if (res->used_fp64) res->used_fp64[b] = o.prec == BH_PREC_FP64 ? 1 : 0;
        if (res->status) res->status[b] = static_cast<int>(lv.size()) >= k ? BH_OK : BH_ENOCONV;
        if (vecs) {
            for (int i = 0; i < k && i < static_cast<int>(lv.size()); ++i) {
                double* y = vecs->data() + (static_cast<size_t>(b) * k + i) * h->D;
                double nrm = 0.0;
                for (int64_t q = 0; q < h->D; ++q) {
                    const double e = U[b] * static_cast<double>(d[q]);
                    y[q] = (U[b] == 0.0 || e == lv[i]) ? r[q] : 0.0;
                    nrm += y[q] * y[q];
                }
                nrm = nrm > 0 ? 1.0 / std::sqrt(nrm) : 0.0;
                for (int64_t q = 0; q < h->D; ++q) y[q] *= nrm;
            }
        }
    }
    return BH_OK;
}

// Lanczos on a subset of the runs; results scattered into the caller's arrays
int solve_subset(bh_ctx* ctx, bh_ham* h, const std::vector<int>& runs, const double* J, const double* U,
                 const bh_lanczos_opts& o, int prec, bh_lanczos_result* res, std::vector<double>* vecs, bool add_iters) {
    const int nb = static_cast<int>(runs.size());
    if (nb == 0) return BH_OK;
    const int k = o.k;
    std::vector<double> Jb(nb), Ub(nb), ev(static_cast<size_t>(nb) * k), rs(ev.size()), ee(ev.size());
    std::vector<int> it(nb), uf(nb), st(nb);
    for (int i = 0; i < nb; ++i) {
        Jb[i] = J[runs[i]];
        Ub[i] = U[runs[i]];
    }
    bh_lanczos_opts o2 = o;
    o2.prec = prec;
    bh_lanczos_result r2{ev.data(), rs.data(), ee.data(), it.data(), uf.data(), st.data()};
    std::vector<double> v2;
    BH_TRY(lanczos_batch(ctx, h, prec, Jb.data(), Ub.data(), nb, o2, &r2, vecs ? &v2 : nullptr));
    for (int i = 0; i < nb; ++i) {
        const int b = runs[i];
        for (int q = 0; q < k; ++q) {
            if (res->evals) res->evals[b * k + q] = ev[i * k + q];
            if (res->resid) res->resid[b * k + q] = rs[i * k + q];
            if (res->errest) res->errest[b * k + q] = ee[i * k + q];
        }
        if (res->iters) res->iters[b] = (add_iters ? res->iters[b] : 0) + it[i];
        if (res->used_fp64) res->used_fp64[b] = uf[i];
        if (res->status) res->status[b] = st[i];
        if (vecs && !v2.empty())
            std::copy(v2.begin() + static_cast<int64_t>(i) * k * h->D, v2.begin() + static_cast<int64_t>(i + 1) * k * h->D,
                      vecs->begin() + static_cast<int64_t>(b) * k * h->D);
    }
    return BH_OK;
}

}  // namespace

int bh_lanczos(bh_ctx* ctx, bh_ham* h, int B, const double* J, const double* U, const bh_lanczos_opts* opts,
               bh_lanczos_result* res) {
    if (!ctx || !h || !J || !U || !opts || !res) return BH_EINVAL;
    if (B < 1 || B > 64) return bh_set_error(ctx, BH_EINVAL, "lanczos: 1 <= B <= 64");
    if (opts->k < 1 || opts->k > LZ_KMAX) return bh_set_error(ctx, BH_EINVAL, "lanczos: 1 <= k <= 64");
    if (opts->distinct && opts->k > LZ_LMAX) return bh_set_error(ctx, BH_EINVAL, "lanczos: distinct levels need k <= 16");
    if (opts->prec != BH_PREC_FP64 && opts->prec != BH_PREC_MIXED) return bh_set_error(ctx, BH_EINVAL, "lanczos: prec");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    const int k = opts->k;
    // per-run outputs always exist internally; copied to the caller's arrays at the end
    std::vector<double> ev(static_cast<size_t>(B) * k, NAN), rs(ev.size(), NAN), ee(ev.size(), NAN);
    std::vector<int> it(B, 0), uf(B, 0), st(B, BH_OK);
    bh_lanczos_result R{ev.data(), rs.data(), ee.data(), it.data(), uf.data(), st.data()};
    std::vector<double> vecs;
    std::vector<double>* vp = nullptr;
    if (opts->want_vectors) {
        vecs.assign(static_cast<size_t>(B) * k * h->D, 0.0);
        vp = &vecs;
    }
    std::vector<int> diag_runs, lz_runs;
    for (int b = 0; b < B; ++b) (J[b] == 0.0 || h->nnz == 0 ? diag_runs : lz_runs).push_back(b);
    BH_TRY(solve_diagonal(ctx, h, diag_runs, U, *opts, &R, vp));
    BH_TRY(solve_subset(ctx, h, lz_runs, J, U, *opts, opts->prec, &R, vp, false));
    // mixed: acceptance of E0 by the FP64 finish; failing solves are rerun in FP64
    if (opts->prec == BH_PREC_MIXED && opts->fallback >= 0) {
        const double acc = opts->accept_rel > 0 ? opts->accept_rel : 1e-9;
        std::vector<int> bad;
        for (int b : lz_runs)
            if (!uf[b] && !bh_lane_accepted(st[b], &ev[b * k], &ee[b * k], k, acc)) bad.push_back(b);
        const bh_lanczos_opts ofb = bh_fallback_opts(*opts);
        BH_TRY(solve_subset(ctx, h, bad, J, U, ofb, BH_PREC_FP64, &R, vp, true));
        // FP64 reruns that are not certified either (possible only with selective orthogonalisation) go
        // to full reorthogonalisation
        if (!bh_lz_fullre(h->D, ofb)) {
            std::vector<int> bad2;
            for (int b : bad)
                if (!bh_lane_accepted(st[b], &ev[b * k], &ee[b * k], k, acc)) bad2.push_back(b);
            BH_TRY(solve_subset(ctx, h, bad2, J, U, bh_fallback_full_opts(*opts), BH_PREC_FP64, &R, vp, true));
        }
    }
    for (int b = 0; b < B; ++b) {
        for (int q = 0; q < k; ++q) {
            if (res->evals) res->evals[b * k + q] = ev[b * k + q];
            if (res->resid) res->resid[b * k + q] = rs[b * k + q];
            if (res->errest) res->errest[b * k + q] = ee[b * k + q];
        }
        if (res->iters) res->iters[b] = it[b];
        if (res->used_fp64) res->used_fp64[b] = uf[b];
        if (res->status) res->status[b] = st[b];
    }
    if (vp) {
        ctx->last_vecs.swap(vecs);
        ctx->last_vecs_B = B;
        ctx->last_vecs_k = k;
        ctx->last_vecs_D = h->D;
    }
    for (int b = 0; b < B; ++b)
        if (st[b] != BH_OK) return bh_set_error(ctx, BH_ENOCONV, "lanczos: some solves did not converge within max_iter");
    return BH_OK;
}

// Replaces Op::FOLM_diag (src/operator.cpp:39-71): `nb_iter` Lanczos steps of J T + U V from the normalised
// Park-Miller start vector, FP64 with full reorthogonalisation (every new vector against the whole basis, two
// Gram-Schmidt passes), no convergence test.  The reference's loop never subtracts alpha v_i (:47-54), stores
// beta one row early (:64-67) and leaves T uninitialised after an early stop (:56-57); here the recurrence is
// the textbook one and T is cut where beta_{t+1} < 1e-6 (the reference's stopping test, :56).
int bh_lanczos_tridiag(bh_ctx* ctx, bh_ham* h, double J, double U, int nb_iter, uint64_t seed, double* alpha, double* beta,
                       int* steps, double* basis) {
    if (!ctx || !h || !alpha || !beta || !steps || nb_iter < 1) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    bh_lanczos_opts o;
    bh_lanczos_default_opts(&o);
    o.k = 1;
    o.prec = BH_PREC_FP64;
    o.fullre_d = std::numeric_limits<int>::max();
    o.max_iter = nb_iter;
    o.ncv = nb_iter;
    o.seed = seed;
    o.distinct = 1;
    int rc = BH_OK;
    auto e = bh_make_engine(ctx, h, BH_PREC_FP64, 1, o, ctx->stream, false, &rc);
    if (!e) return rc;
    BH_TRY(e->admit(0, J, U));
    const int n = static_cast<int>(std::min<int64_t>(nb_iter, h->D));
    BH_TRY(e->run(n));
    std::vector<double> a(n), b(n + 1);
    BH_TRY(e->export_tridiag(n, a.data(), b.data(), nullptr));
    int used = n;
    for (int t = 1; t < n; ++t)
        if (b[t] < 1e-6) {
            used = t;
            break;
        }
    for (int t = 0; t < used; ++t) alpha[t] = a[t];
    for (int t = 0; t < used; ++t) beta[t] = b[t];
    *steps = used;
    if (basis) BH_TRY(e->export_tridiag(used, a.data(), b.data(), basis));
    return BH_OK;
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
