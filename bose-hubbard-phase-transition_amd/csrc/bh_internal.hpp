// bh_internal.hpp -- shared definitions of libbhx (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/bh_capi.h"

// ---- context -------------------------------------------------------------------------------------

struct bh_ctx {
    int device = 0;
    size_t budget = 0;
    size_t used = 0;
    hipStream_t stream = nullptr;
    mutable std::string err;
    std::recursive_mutex mu;
    int timing = 0;
    std::map<std::string, std::pair<int64_t, double>> kstats;
    // kernel timings recorded but not yet read back (resolved in bulk, so timing adds no host waits)
    struct PendingTimer {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<PendingTimer> ktimers;
    // grow-only device workspaces reused across calls (Krylov basis, locked vectors, scratch)
    static constexpr int NWS = 8;
    void* ws[NWS] = {};
    size_t ws_bytes[NWS] = {};
    // start vector cached in workspace slot 3 for (r0_D, r0_seed), with the inverse norm of its copy in the
    // storage precision r0_mixed (1: FP32)
    int64_t r0_D = -1;
    uint64_t r0_seed = 0;
    int r0_mixed = -1;
    uint64_t r0_layout = 0;  // layout_id of the basis for the lexicographically indexed start vectors, else 0
    double r0_scale = 0.0;
    // Ritz vectors kept by the last bh_lanczos(want_vectors=1): [B][k][D] doubles on the host
    std::vector<double> last_vecs;
    int last_vecs_B = 0, last_vecs_k = 0;
    int64_t last_vecs_D = 0;
    // device allocations made through the context: calls and host time (hipFree waits for the device)
    int64_t n_alloc = 0, n_free = 0;
    double alloc_ms = 0.0, free_ms = 0.0;
};

int bh_set_error(const bh_ctx* ctx, int code, const std::string& msg);

#define BH_HIP(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return bh_set_error((ctx), BH_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

#define BH_TRY(expr)              \
    do {                          \
        int rc_ = (expr);         \
        if (rc_ != BH_OK) return rc_; \
    } while (0)

// budgeted device allocation (bytes counted against ctx->budget)
int bh_dev_alloc(bh_ctx* ctx, void** p, size_t bytes);
void bh_dev_free(bh_ctx* ctx, void* p, size_t bytes);
// workspace slot of at least `bytes` (kept by the context until bh_ctx_destroy)
int bh_ws_get(bh_ctx* ctx, int slot, size_t bytes, void** p);

// RAII device buffer
template <typename T>
struct DevArray {
    bh_ctx* ctx = nullptr;
    T* p = nullptr;
    size_t n = 0;
    DevArray() = default;
    DevArray(const DevArray&) = delete;
    DevArray& operator=(const DevArray&) = delete;
    ~DevArray() { reset(); }
    int alloc(bh_ctx* c, size_t count) {
        reset();
        ctx = c;
        n = count;
        void* q = nullptr;
        int rc = bh_dev_alloc(c, &q, sizeof(T) * (count ? count : 1));
        p = static_cast<T*>(q);
        if (rc != BH_OK) n = 0;
        return rc;
    }
    void reset() {
        if (p) bh_dev_free(ctx, p, sizeof(T) * (n ? n : 1));
        p = nullptr;
        n = 0;
    }
    T* release() {
        T* q = p;
        p = nullptr;
        n = 0;
        return q;
    }
    size_t bytes() const { return sizeof(T) * n; }
};

// kernel timing (HIP events on the launch stream when ctx->timing is on)
void bh_ktimers_flush(bh_ctx* ctx);  // waits for the recorded events and adds them to kstats

struct KTimer {
    bh_ctx* ctx;
    const char* name;
    hipStream_t s;
    hipEvent_t a = nullptr, b = nullptr;
    KTimer(bh_ctx* c, const char* n, hipStream_t st);
    ~KTimer();
};

// ---- basis / Hamiltonian handles ------------------------------------------------------------------

struct bh_basis {
    bh_ctx* ctx = nullptr;
    int m = 0, n = 0, order = 0;
    int64_t D = 0;
    int nw = 1, bits = 4;         // packed-state layout used by the kernels (NW words of BITS-bit sites)
    int32_t* lex_of_pos = nullptr;  // [D]
    int32_t* pos_of_lex = nullptr;  // [D]
    uint32_t* rank_tab = nullptr;   // [(m-1) * (n+1)] g[l][R] = C(R + m-l-2, m-l-1)
    size_t bytes = 0;
    uint64_t layout_id = 0;         // new at every build and permutation (keys host caches of the layout)
};

// the caller-chosen orders index the solver's start vector by lexicographic rank, so that a solve runs the
// lexicographic order's Krylov sequence whatever the permutation
inline bool bh_basis_lex_start(const bh_basis* b) {
    return b && (b->order == BH_ORDER_BAND || b->order == BH_ORDER_CUSTOM);
}
inline bool bh_basis_lex_start(int order) { return order == BH_ORDER_BAND || order == BH_ORDER_CUSTOM; }
uint64_t bh_next_layout_id();

// Autotuner result for one (kernel, precision, batch width) of a sector: the chosen variant and the
// measured time of every candidate (variant, microseconds per launch); no candidates when not timed.
struct BhTuned {
    int variant = 1;
    std::vector<std::pair<int, double>> cand;
};

struct bh_ham {
    bh_ctx* ctx = nullptr;
    const bh_basis* basis = nullptr;  // the basis at build time (not used after it: the basis may be destroyed)
    // caller-chosen orders (bh_basis_lex_start): a copy of the basis's lex_of_pos at build time, which indexes the
    // solver's start vectors, and the basis's layout_id then; nullptr and 0 otherwise
    int32_t* lex_of_pos = nullptr;  // [D]
    uint64_t layout_id = 0;
    int64_t D = 0, nnz = 0;
    int code_bits = 0;
    unsigned flags = 0;
    int max_row = 0;              // longest row
    int32_t* row_ptr = nullptr;   // [D+1]
    uint32_t* words = nullptr;    // [nnz] column << code_bits | code
    float* diag = nullptr;        // [D] V (exact small integers)
    double* table = nullptr;      // [1 << code_bits] sqrt(code), exact FP64
    float* table32 = nullptr;     // [1 << code_bits] sqrt(code) rounded to FP32
    // FP64 CSR (BH_PREC_FP64), materialised on demand
    int32_t* col64 = nullptr;     // [nnz]
    double* val64 = nullptr;      // [nnz] -sqrt(code)
    double* diag64 = nullptr;     // [D]
    size_t bytes = 0, bytes64 = 0;
    // spectral bounds: V in [vmin, vmax]; the largest row sum of |T| (Gershgorin radius of T)
    double vmin = 0.0, vmax = 0.0, tmax = 0.0;
    std::vector<double> vlevels;  // distinct values of V ascending (host, filled on first use)
    // row blocks of the LDS-staged engine, per batch width B: (first row, first entry) of each block, then
    // (D, nnz)
    std::map<int, std::pair<int32_t*, int64_t>> rowblocks;
    size_t bytes_rb = 0;
    // operator-kernel variants chosen by the autotuner, keyed by bh_tune_key(kind, prec, B)
    std::map<int, BhTuned> tuned;
};

// key of a tuning result: kind 0 = Lanczos K1, 1 = SpMV
inline int bh_tune_key(int kind, int prec, int B) { return (kind * 4 + prec) * 128 + B; }
// false when BHX_TUNE=0 (read once): the default variants are used without timing
bool bh_tune_enabled();

// row blocks of at most (products / B) entries for batch width B (built once, cached in the handle);
// returns BH_ERANGE when a row does not fit one block
int bh_ham_rowblocks(bh_ham* h, int B, const int32_t** rb, int64_t* nblocks);
// engine selection of the operator kernels: 0 = lanes per row (vector), 1 = LDS-staged row blocks,
// 2 = one thread per (row, b); automatic unless BHX_ENGINE is set (read once).
int bh_engine_choice(int B, const bh_ham* h);
int bh_engine_env();  // BHX_ENGINE (-1 when unset)

// ---- primes and packed occupation states ------------------------------------------------------------

constexpr int BH_MAX_SITES = 32;
constexpr int BH_STAGE_PRODUCTS_HOST = 2048;  // must equal BH_STAGE_PRODUCTS of the row-block engine
extern const int bh_primes[BH_MAX_SITES];

// A Fock state of up to NW*64/BITS sites packed in NW 64-bit words (BITS bits per site).
template <int NW, int BITS>
struct PState {
    static constexpr int PER = 64 / BITS;
    static constexpr int MAXM = NW * PER;
    static constexpr uint64_t MASK = (1ull << BITS) - 1ull;
    uint64_t w[NW];

    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int a = 0; a < NW; ++a) w[a] = 0;
    }
    // site index may be a runtime value: the word is chosen by a select chain, not an indexed load
    __device__ __forceinline__ int get(int i) const {
        const int wi = i / PER;
        const int sh = (i % PER) * BITS;
        uint64_t x = w[0];
#pragma unroll
        for (int a = 1; a < NW; ++a) x = (wi == a) ? w[a] : x;
        return static_cast<int>((x >> sh) & MASK);
    }
    // i must be a compile-time constant after unrolling
    __device__ __forceinline__ void put(int i, int v) { w[i / PER] |= static_cast<uint64_t>(v) << ((i % PER) * BITS); }
};

// lexicographic unranking, descending order (src/hamiltonian.cpp:60-85):
// rank(s) = sum_{l=0}^{m-2} g[l][Rem_l],  Rem_l = n - sum_{l'<=l} s_l',  g[l][R] = C(R + m-l-2, m-l-1)
template <int NW, int BITS>
__device__ __forceinline__ void bh_unrank(int64_t r, int m, int n, const uint32_t* __restrict__ g,
                                          PState<NW, BITS>& s) {
    s.clear();
    int R = n;
    int64_t rr = r;
#pragma unroll
    for (int i = 0; i < PState<NW, BITS>::MAXM - 1; ++i) {
        if (i < m - 1) {
            const uint32_t* gi = g + static_cast<size_t>(i) * (n + 1);
            int K = 0;
            while (K < R && static_cast<int64_t>(gi[K + 1]) <= rr) ++K;
            rr -= gi[K];
            s.put(i, R - K);
            R = K;
        }
    }
    // last site takes the remainder (static position m-1 is runtime: place via unrolled match)
#pragma unroll
    for (int i = 0; i < PState<NW, BITS>::MAXM; ++i)
        if (i == m - 1) s.put(i, R);
}

// ---- host helpers ----------------------------------------------------------------------------------
int64_t bh_binom_host(int n, int k);  // exact, -1 on overflow of int64
int bh_key_bits(int m, int n);        // bits of the largest key p_{m-1}^n (exact), or > 128
int bh_state_layout(int m, int n, int* nw, int* bits);
