// bh_spmv.hip -- application of H_b = J_b T + U_b V to B row-interleaved vectors.
//
// Replaces the Eigen CSC x dense product behind Spectra's SparseGenMatProd::perform_op
// (external/spectra/include/Spectra/MatOp/SparseGenMatProd.h:81-86 ->
// include/Eigen/src/SparseCore/SparseDenseProduct.h:86-107), a serial column scatter in the
// reference.  Here: CSR rows, G lanes per row and B vectors per CSR pass, FP64 accumulation.
// The kernel is HBM-bound: per pass it reads the matrix once (4 B per entry in the packed format,
// 12 B in FP64 CSR) plus the row pointers and the diagonal, and B * (x gather + y write).
#include <cmath>
#include <mutex>

#include "bh_kernels.hpp"
#include "bh_spmv.hpp"

template <typename XT, int B, int G, int FMT>
__global__ void __launch_bounds__(BH_WG) k_spmv(SpmvArgs<XT> a) {
    constexpr int GB = G * B;
    constexpr int SLOTS = BH_WG / GB;
    extern __shared__ double tab[];  // (1 << cb) entries, packed format only
    if constexpr (FMT == BH_FMT_PACKED) {
        for (int i = threadIdx.x; i < (1 << a.cb); i += BH_WG) tab[i] = a.table[i];
        __syncthreads();
    }
    const int lane = threadIdx.x;
    const int slot = lane / GB;
    const int g = (lane % GB) / B;
    const int b = lane % B;
    const int64_t ntile = (a.D + SLOTS - 1) / SLOTS;
    const int64_t tile = a.xcd ? bh_xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    if (tile >= ntile) return;  // whole workgroup leaves together (tile is uniform)
    const int64_t r = tile * SLOTS + slot;
    double acc = 0.0;
    if (r < a.D) acc = bh_row_partial<XT, B, G, FMT>(a.row_ptr[r], a.row_ptr[r + 1], g, b, a.words, a.col64, a.val64, tab,
                                                       a.cb, a.x);
    acc = bh_group_sum<B, G>(acc);
    if (g == 0 && r < a.D) {
        const double hx = (FMT == BH_FMT_PACKED) ? -acc : acc;
        const double d = a.diag32 ? static_cast<double>(a.diag32[r]) : a.diag64[r];
        const double xr = to_d(a.x[r * B + b]);
        a.y[r * B + b] = static_cast<XT>(a.J[b] * hx + a.U[b] * d * xr);
    }
}

template <typename XT, int B, int FMT>
__global__ void __launch_bounds__(BH_WG) k_spmv_rb(SpmvArgs<XT> a, const int32_t* __restrict__ rb, int64_t nblocks) {
    extern __shared__ double smem[];
    double* tab = smem;                                  // (1 << cb) entries (packed format)
    double* prod = smem + (FMT == BH_FMT_PACKED ? (1 << a.cb) : 0);
    if constexpr (FMT == BH_FMT_PACKED) {
        for (int i = threadIdx.x; i < (1 << a.cb); i += BH_WG) tab[i] = a.table[i];
        __syncthreads();
    }
    const int64_t t = a.xcd ? bh_xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    auto epi = [&](int64_t r, int bb, const double (&sum)[1]) {
        const double hx = (FMT == BH_FMT_PACKED) ? -sum[0] : sum[0];
        const double d = a.diag32 ? static_cast<double>(a.diag32[r]) : a.diag64[r];
        a.y[r * B + bb] = static_cast<XT>(a.J[bb] * hx + a.U[bb] * d * to_d(a.x[r * B + bb]));
    };
    bh_rowblock_run<XT, B, FMT, BH_WG>(t, t + 1, rb, a.row_ptr, a.words, a.col64, a.val64, tab, a.cb, a.x, prod, epi);
}

template <typename XT, int B, int V, int FMT>
__global__ void __launch_bounds__(BH_WG) k_spmv_rl(SpmvArgs<XT> a, int64_t rows_per_wg) {
    extern __shared__ double tab[];
    if constexpr (FMT == BH_FMT_PACKED) {
        for (int i = threadIdx.x; i < (1 << a.cb); i += BH_WG) tab[i] = a.table[i];
        __syncthreads();
    }
    const int64_t t = a.xcd ? bh_xcd_tile(blockIdx.x, gridDim.x) : blockIdx.x;
    const int64_t r0 = min(t * rows_per_wg, a.D), r1 = min(r0 + rows_per_wg, a.D);
    auto epi = [&](int64_t r, int b0, const double (&acc)[V]) {
        const double d = a.diag32 ? static_cast<double>(a.diag32[r]) : a.diag64[r];
        const BhVec<XT, V> xv = bh_vload<V>(a.x + r * B + b0);
        BhVec<XT, V> yv;
#pragma unroll
        for (int v = 0; v < V; ++v) {
            const double hx = (FMT == BH_FMT_PACKED) ? -acc[v] : acc[v];
            yv.v[v] = static_cast<XT>(a.J[b0 + v] * hx + a.U[b0 + v] * d * to_d(xv.v[v]));
        }
        *reinterpret_cast<BhVec<XT, V>*>(a.y + r * B + b0) = yv;
    };
    bh_rowlane_run<XT, B, V, FMT, BH_WG, 4>(r0, r1, a.row_ptr, a.words, a.col64, a.val64, tab, a.cb, a.x, epi);
}

// row-lane engine with V vectors per thread (V * sizeof(XT) <= 16 bytes, V <= B)
template <typename XT, int FMT>
static void launch_spmv_rl_t(const SpmvArgs<XT>& a, int B, int V, hipStream_t s) {
    const size_t shm = FMT == BH_FMT_PACKED ? (static_cast<size_t>(1) << a.cb) * sizeof(double) : 0;
    const int64_t rows_per_wg = static_cast<int64_t>(BH_WG) * V / B * 4;
    const unsigned grid = static_cast<unsigned>((a.D + rows_per_wg - 1) / rows_per_wg);
#define L(BB, VV)                                                                                              \
    if (B == BB && V == VV) {                                                                                 \
        hipLaunchKernelGGL((k_spmv_rl<XT, BB, VV, FMT>), dim3(grid), dim3(BH_WG), shm, s, a, rows_per_wg);    \
        return;                                                                                               \
    }
    L(1, 1) L(2, 1) L(4, 1) L(8, 1) L(16, 1) L(32, 1) L(64, 1)
    L(2, 2) L(4, 2) L(8, 2) L(16, 2) L(32, 2) L(64, 2)
    if constexpr (sizeof(XT) == 4) {
        L(4, 4) L(8, 4) L(16, 4) L(32, 4) L(64, 4)
    }
#undef L
}

template <typename XT, int FMT>
static void launch_spmv_rb_t(const SpmvArgs<XT>& a, int B, const int32_t* rb, int64_t nb, hipStream_t s) {
    const size_t shm = ((FMT == BH_FMT_PACKED ? (static_cast<size_t>(1) << a.cb) : 0) + BH_STAGE_PRODUCTS) * sizeof(double);
    const unsigned grid = static_cast<unsigned>(nb);
#define L(BB)                                                                                            \
    if (B == BB) {                                                                                      \
        hipLaunchKernelGGL((k_spmv_rb<XT, BB, FMT>), dim3(grid), dim3(BH_WG), shm, s, a, rb, nb);       \
        return;                                                                                         \
    }
    L(1) L(2) L(4) L(8) L(16) L(32) L(64)
#undef L
}

void bh_launch_spmv_rb(const SpmvArgs<float>& a, int B, const int32_t* rb, int64_t nb, hipStream_t s) {
    launch_spmv_rb_t<float, BH_FMT_PACKED>(a, B, rb, nb, s);
}
void bh_launch_spmv_rb(const SpmvArgs<double>& a, int B, int fmt, const int32_t* rb, int64_t nb, hipStream_t s) {
    if (fmt == BH_FMT_PACKED) launch_spmv_rb_t<double, BH_FMT_PACKED>(a, B, rb, nb, s);
    else launch_spmv_rb_t<double, BH_FMT_CSR64>(a, B, rb, nb, s);
}

// SpMV variants: 0 = LDS-staged row blocks; V = 1, 2, 4 = row-lane engine with V vectors per thread
// (64 V / B rows per wavefront); 100 + G = G lanes per row (64 / (G B) rows per wavefront)
static int spmv_launch(bh_ham* h, int var, const SpmvArgs<double>* a64, const SpmvArgs<float>* a32, int fmt, int B,
                       hipStream_t s) {
    if (var == 0) {
        const int32_t* rb = nullptr;
        int64_t nb = 0;
        BH_TRY(bh_ham_rowblocks(h, B, &rb, &nb));
        if (a64) bh_launch_spmv_rb(*a64, B, fmt, rb, nb, s);
        else bh_launch_spmv_rb(*a32, B, rb, nb, s);
    } else if (var >= 100) {
        if (a64) bh_launch_spmv(*a64, B, var - 100, fmt, s);
        else bh_launch_spmv(*a32, B, var - 100, s);
    } else if (a64) {
        if (fmt == BH_FMT_PACKED) launch_spmv_rl_t<double, BH_FMT_PACKED>(*a64, B, var, s);
        else launch_spmv_rl_t<double, BH_FMT_CSR64>(*a64, B, var, s);
    } else {
        launch_spmv_rl_t<float, BH_FMT_PACKED>(*a32, B, var, s);
    }
    return BH_OK;
}

// One operator application with the variant chosen for (h, precision, B).  The first application at a
// (precision, B) runs the autotuner unless BHX_TUNE=0 or BHX_ENGINE is set: every variant that fits is
// timed with HIP events over a warm-up and three applications to the caller's own x and y (y = H x is
// idempotent), and the fastest is kept in the handle.  Under stream capture the default variant is used.
int bh_apply(bh_ctx* ctx, bh_ham* h, const SpmvArgs<double>* a64, const SpmvArgs<float>* a32, int fmt, int B,
             hipStream_t s) {
    const int prec = a32 ? BH_PREC_MIXED : (fmt == BH_FMT_CSR64 ? BH_PREC_FP64 : 2);
    const int key = bh_tune_key(1, prec, B);
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    auto it = h->tuned.find(key);
    if (it != h->tuned.end()) return spmv_launch(h, it->second.variant, a64, a32, fmt, B, s);
    const int eng = bh_engine_choice(B, h);
    BhTuned t;
    t.variant = eng == 2 ? 1 : (eng == 1 ? 0 : 100 + bh_pick_lanes(h, B));
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    BH_HIP(ctx, hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) return spmv_launch(h, t.variant, a64, a32, fmt, B, s);
    if (bh_tune_enabled() && bh_engine_env() < 0) {
        std::vector<int> cand;
        if (static_cast<int64_t>(h->max_row) * B <= BH_STAGE_PRODUCTS_HOST) cand.push_back(0);
        cand.push_back(1);
        if (B >= 2) cand.push_back(2);
        if (B >= 4 && a32) cand.push_back(4);
        const int G = bh_pick_lanes(h, B);
        if (G > 1) cand.push_back(100 + G);
        hipEvent_t e0 = nullptr, e1 = nullptr;
        BH_HIP(ctx, hipEventCreate(&e0));
        BH_HIP(ctx, hipEventCreate(&e1));
        double best = 0.0;
        for (int var : cand) {
            BH_TRY(spmv_launch(h, var, a64, a32, fmt, B, s));
            BH_HIP(ctx, hipEventRecord(e0, s));
            for (int r = 0; r < 3; ++r) BH_TRY(spmv_launch(h, var, a64, a32, fmt, B, s));
            BH_HIP(ctx, hipEventRecord(e1, s));
            BH_HIP(ctx, hipEventSynchronize(e1));
            float ms = 0.f;
            BH_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
            const double us = 1e3 * ms / 3;
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
    return spmv_launch(h, t.variant, a64, a32, fmt, B, s);
}

template <typename XT, int FMT>
static void launch_spmv_t(const SpmvArgs<XT>& a, int B, int G, hipStream_t s) {
    const int GB = G * B;
    const int64_t slots = BH_WG / GB;
    const unsigned grid = static_cast<unsigned>((a.D + slots - 1) / slots);
    const size_t shm = FMT == BH_FMT_PACKED ? (static_cast<size_t>(1) << a.cb) * sizeof(double) : 0;
#define L(BB, GG)                                                                               \
    if (B == BB && G == GG) {                                                                  \
        hipLaunchKernelGGL((k_spmv<XT, BB, GG, FMT>), dim3(grid), dim3(BH_WG), shm, s, a);     \
        return;                                                                                \
    }
    L(1, 1) L(1, 2) L(1, 4) L(1, 8) L(1, 16) L(1, 32) L(1, 64)
    L(2, 1) L(2, 2) L(2, 4) L(2, 8) L(2, 16) L(2, 32)
    L(4, 1) L(4, 2) L(4, 4) L(4, 8) L(4, 16)
    L(8, 1) L(8, 2) L(8, 4) L(8, 8)
    L(16, 1) L(16, 2) L(16, 4)
    L(32, 1) L(32, 2)
    L(64, 1)
#undef L
}

void bh_launch_spmv(const SpmvArgs<float>& a, int B, int G, hipStream_t s) { launch_spmv_t<float, BH_FMT_PACKED>(a, B, G, s); }
void bh_launch_spmv(const SpmvArgs<double>& a, int B, int G, int fmt, hipStream_t s) {
    if (fmt == BH_FMT_PACKED) launch_spmv_t<double, BH_FMT_PACKED>(a, B, G, s);
    else launch_spmv_t<double, BH_FMT_CSR64>(a, B, G, s);
}

int bh_pick_lanes(const bh_ham* h, int B) {
    // lanes per row: about half the mean row length, as a power of two, G*B <= 64
    const double avg = h->D ? static_cast<double>(h->nnz) / static_cast<double>(h->D) : 1.0;
    int G = 1;
    while (G < 64 && G * 2 <= avg) G <<= 1;
    while (G * B > 64 && G > 1) G >>= 1;
    return G;
}

double bh_spmv_bytes(const bh_ham* h, int prec, int B) {
    const double D = static_cast<double>(h->D), nnz = static_cast<double>(h->nnz);
    if (prec == BH_PREC_FP64) return nnz * 12.0 + 4.0 * (D + 1) + 8.0 * D + B * 2.0 * 8.0 * D;
    return nnz * 4.0 + 4.0 * (D + 1) + 4.0 * D + B * 2.0 * 4.0 * D;
}

int bh_is_pow2_le64(int B) { return B >= 1 && B <= 64 && (B & (B - 1)) == 0; }

// ---- the SpMV entry point that allocates nothing -------------------------------------------------------------
// bh_spmv stages the per-lane coefficients J_b, U_b in two small device arrays it allocates per call.  Callers that
// cannot allocate (code under HIP graph capture, tight serving loops) use bh_spmv_noalloc: the coefficients go to
// this file-scope array instead, whose device address is looked up once per device with hipGetSymbolAddress (the
// array's name in host code is a host-side stand-in, never a device address).  Every call on a device shares
// that device's copy, so calls that overlap on one device race.
namespace {
constexpr int SPMV_SCRATCH_LANES = 64;  // the largest batch width of the SpMV kernels
static __device__ __align__(256) double scratch[2 * SPMV_SCRATCH_LANES];
constexpr int SPMV_SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SPMV_SCRATCH_DEVICES] = {};

// device address of `scratch` on the context's device (the device is current), looked up on first use
int scratch_ptr(bh_ctx* ctx, double** out) {
    if (ctx->device < 0 || ctx->device >= SPMV_SCRATCH_DEVICES)
        return bh_set_error(ctx, BH_EINVAL, "spmv_noalloc: device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[ctx->device]) {
        void* p = nullptr;
        BH_HIP(ctx, hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratch_ptr[ctx->device] = static_cast<double*>(p);
    }
    *out = g_scratch_ptr[ctx->device];
    return BH_OK;
}
}  // namespace

extern "C" {

int bh_spmv_noalloc(bh_ctx* ctx, bh_ham* h, int prec, int B, const double* J, const double* U, const void* x, void* y,
                    void* stream) {
    if (!ctx || !h || !J || !U || !x || !y) return BH_EINVAL;
    if (!bh_is_pow2_le64(B) || B > SPMV_SCRATCH_LANES)
        return bh_set_error(ctx, BH_EINVAL, "spmv_noalloc: B must be a power of two <= 64 (the scratch holds 64 lanes)");
    if (prec != BH_PREC_FP64 && prec != BH_PREC_MIXED) return bh_set_error(ctx, BH_EINVAL, "spmv_noalloc: prec");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    if (prec == BH_PREC_FP64 && !h->col64)
        return bh_set_error(ctx, BH_EINVAL, "spmv_noalloc: FP64 CSR not materialised (bh_ham_prepare first)");
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    double* sc = nullptr;
    BH_TRY(scratch_ptr(ctx, &sc));
    BH_HIP(ctx, hipMemcpyAsync(sc, J, B * sizeof(double), hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(sc + SPMV_SCRATCH_LANES, U, B * sizeof(double), hipMemcpyHostToDevice, s));
    // the variant the autotuner chose for this sector (no tuning here: it times with events), else the default
    // bh_apply uses under capture; the staged engine only when its row-block table for B already exists
    const int key = bh_tune_key(1, prec, B);
    auto it = h->tuned.find(key);
    int var;
    if (it != h->tuned.end()) {
        var = it->second.variant;
    } else {
        const int eng = bh_engine_choice(B, h);
        var = eng == 2 ? 1 : (eng == 1 ? 0 : 100 + bh_pick_lanes(h, B));
    }
    if (var == 0 && h->rowblocks.find(B) == h->rowblocks.end()) var = 1;
    if (prec == BH_PREC_FP64) {
        SpmvArgs<double> a{h->D, h->row_ptr, h->words, h->col64, h->val64, nullptr, h->diag64, h->table, h->code_bits,
                           sc, sc + SPMV_SCRATCH_LANES, static_cast<const double*>(x), static_cast<double*>(y), 1};
        BH_TRY(spmv_launch(h, var, &a, nullptr, BH_FMT_CSR64, B, s));
    } else {
        SpmvArgs<float> a{h->D, h->row_ptr, h->words, nullptr, nullptr, h->diag, nullptr, h->table, h->code_bits,
                          sc, sc + SPMV_SCRATCH_LANES, static_cast<const float*>(x), static_cast<float*>(y), 1};
        BH_TRY(spmv_launch(h, var, nullptr, &a, BH_FMT_PACKED, B, s));
    }
    BH_HIP(ctx, hipGetLastError());
    // outside a capture the call returns with its work done (the host's J and U may be reused at once)
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    BH_HIP(ctx, hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) BH_HIP(ctx, hipStreamSynchronize(s));
    return BH_OK;
}

int bh_spmv(bh_ctx* ctx, bh_ham* h, int prec, int B, const double* J, const double* U, const void* x, void* y,
            void* stream) {
    if (!ctx || !h || !J || !U || !x || !y) return BH_EINVAL;
    if (!bh_is_pow2_le64(B)) return bh_set_error(ctx, BH_EINVAL, "spmv: B must be a power of two <= 64");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream ? static_cast<hipStream_t>(stream) : ctx->stream;
    BH_TRY(bh_ham_prepare_locked(h, prec));
    DevArray<double> dJ, dU;
    BH_TRY(dJ.alloc(ctx, B));
    BH_TRY(dU.alloc(ctx, B));
    BH_HIP(ctx, hipMemcpyAsync(dJ.p, J, B * sizeof(double), hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(dU.p, U, B * sizeof(double), hipMemcpyHostToDevice, s));
    {
        KTimer kt(ctx, "spmv", s);
        if (prec == BH_PREC_FP64) {
            SpmvArgs<double> a{h->D, h->row_ptr, h->words, h->col64, h->val64, nullptr, h->diag64, h->table, h->code_bits,
                               dJ.p, dU.p, static_cast<const double*>(x), static_cast<double*>(y), 1};
            BH_TRY(bh_apply(ctx, h, &a, nullptr, BH_FMT_CSR64, B, s));
        } else {
            SpmvArgs<float> a{h->D, h->row_ptr, h->words, nullptr, nullptr, h->diag, nullptr, h->table, h->code_bits,
                              dJ.p, dU.p, static_cast<const float*>(x), static_cast<float*>(y), 1};
            BH_TRY(bh_apply(ctx, h, nullptr, &a, BH_FMT_PACKED, B, s));
        }
    }
    BH_HIP(ctx, hipGetLastError());
    BH_HIP(ctx, hipStreamSynchronize(s));  // dJ/dU are freed on return
    return BH_OK;
}

int bh_spmv_host(bh_ctx* ctx, bh_ham* h, int prec, int B, const double* J, const double* U, const double* x,
                 double* y) {
    if (!ctx || !h || !x || !y) return BH_EINVAL;
    if (!bh_is_pow2_le64(B)) return bh_set_error(ctx, BH_EINVAL, "spmv: B must be a power of two <= 64");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const size_t n = static_cast<size_t>(h->D) * B;
    if (prec == BH_PREC_FP64) {
        DevArray<double> dx, dy;
        BH_TRY(dx.alloc(ctx, n));
        BH_TRY(dy.alloc(ctx, n));
        BH_HIP(ctx, hipMemcpyAsync(dx.p, x, n * 8, hipMemcpyHostToDevice, s));
        BH_TRY(bh_spmv(ctx, h, prec, B, J, U, dx.p, dy.p, nullptr));
        BH_HIP(ctx, hipMemcpyAsync(y, dy.p, n * 8, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
    } else {
        std::vector<float> xf(n), yf(n);
        for (size_t i = 0; i < n; ++i) xf[i] = static_cast<float>(x[i]);
        DevArray<float> dx, dy;
        BH_TRY(dx.alloc(ctx, n));
        BH_TRY(dy.alloc(ctx, n));
        BH_HIP(ctx, hipMemcpyAsync(dx.p, xf.data(), n * 4, hipMemcpyHostToDevice, s));
        BH_TRY(bh_spmv(ctx, h, prec, B, J, U, dx.p, dy.p, nullptr));
        BH_HIP(ctx, hipMemcpyAsync(yf.data(), dy.p, n * 4, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (size_t i = 0; i < n; ++i) y[i] = yf[i];
    }
    return BH_OK;
}

int bh_spmv_bench(bh_ctx* ctx, bh_ham* h, int prec, int B, int reps, double* avg_ms, double* bytes) {
    if (!ctx || !h || reps < 1) return BH_EINVAL;
    if (!bh_is_pow2_le64(B)) return bh_set_error(ctx, BH_EINVAL, "spmv: B must be a power of two <= 64");
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    BH_TRY(bh_ham_prepare_locked(h, prec));
    const size_t n = static_cast<size_t>(h->D) * B;
    const size_t es = prec == BH_PREC_FP64 ? 8 : 4;
    DevArray<uint8_t> dx, dy;
    DevArray<double> dJ, dU;
    BH_TRY(dx.alloc(ctx, n * es));
    BH_TRY(dy.alloc(ctx, n * es));
    BH_TRY(dJ.alloc(ctx, B));
    BH_TRY(dU.alloc(ctx, B));
    std::vector<double> ones(B, 1.0);
    BH_HIP(ctx, hipMemcpyAsync(dJ.p, ones.data(), B * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(dU.p, ones.data(), B * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemsetAsync(dx.p, 0, n * es, s));
    auto launch = [&]() -> int {
        if (prec == BH_PREC_FP64) {
            SpmvArgs<double> a{h->D, h->row_ptr, h->words, h->col64, h->val64, nullptr, h->diag64, h->table, h->code_bits,
                               dJ.p, dU.p, reinterpret_cast<const double*>(dx.p), reinterpret_cast<double*>(dy.p), 1};
            return bh_apply(ctx, h, &a, nullptr, BH_FMT_CSR64, B, s);
        }
        SpmvArgs<float> a{h->D, h->row_ptr, h->words, nullptr, nullptr, h->diag, nullptr, h->table, h->code_bits,
                          dJ.p, dU.p, reinterpret_cast<const float*>(dx.p), reinterpret_cast<float*>(dy.p), 1};
        return bh_apply(ctx, h, nullptr, &a, BH_FMT_PACKED, B, s);
    };
    BH_TRY(launch());  // warm-up (also builds the row blocks)
    hipEvent_t e0, e1;
    BH_HIP(ctx, hipEventCreate(&e0));
    BH_HIP(ctx, hipEventCreate(&e1));
    BH_HIP(ctx, hipEventRecord(e0, s));
    for (int i = 0; i < reps; ++i) BH_TRY(launch());
    BH_HIP(ctx, hipEventRecord(e1, s));
    BH_HIP(ctx, hipEventSynchronize(e1));
    float ms = 0.f;
    BH_HIP(ctx, hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    BH_HIP(ctx, hipGetLastError());
    if (avg_ms) *avg_ms = ms / reps;
    if (bytes) *bytes = bh_spmv_bytes(h, prec, B);
    return BH_OK;
}

}  // extern "C"
