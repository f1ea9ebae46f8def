// bh_basis.hip -- fixed-N Fock basis on the device: lexicographic unranking, exact prime-product
// keys, stable LSD radix sort, permutation / inverse permutation in the reference's order.
//
// Reference: BH::init_lexicographic / next_lexicographic (src/hamiltonian.cpp:60-85) enumerate the
// states in descending lexicographic order; BH::calculate_tags (:91-106) assigns
// tag = sum n_i ln p_i; BH::sort_basis (:109-123) argsorts the tags and applies the argsort with a
// cycle-swap loop that realises the inverse permutation.  Here the ordering key is the exact integer
// prod p_i^n_i (the tag is its logarithm, so the orders agree; distinct states have distinct keys by
// unique factorisation), sorted by a radix sort.
#include <cmath>
#include <cstring>
#include <memory>

#include "bh_scan.hpp"

// ---- exclusive scan ------------------------------------------------------------------------------

__global__ void __launch_bounds__(BH_SCAN_BLOCK) k_scan_reduce(const int32_t* __restrict__ in, int64_t n,
                                                               int64_t* __restrict__ bsum) {
    __shared__ int64_t lds[BH_SCAN_BLOCK / 64 + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * BH_SCAN_TILE + static_cast<int64_t>(threadIdx.x) * BH_SCAN_ITEMS;
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < BH_SCAN_ITEMS; ++i)
        if (base + i < n) s += in[base + i];
    int64_t tot = 0;
    (void)bh_block_exclusive_scan<BH_SCAN_BLOCK>(s, lds, &tot);
    if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// single block: exclusive scan of nb block sums in place; total -> *total
__global__ void __launch_bounds__(1024) k_scan_blocks(int64_t* __restrict__ bsum, int64_t nb, int64_t* __restrict__ total) {
    __shared__ int64_t lds[1024 / 64 + 1];
    const int64_t per = (nb + 1023) / 1024;
    const int64_t b0 = static_cast<int64_t>(threadIdx.x) * per;
    int64_t s = 0;
    for (int64_t i = b0; i < b0 + per && i < nb; ++i) s += bsum[i];
    int64_t tot = 0;
    int64_t run = bh_block_exclusive_scan<1024>(s, lds, &tot);
    for (int64_t i = b0; i < b0 + per && i < nb; ++i) {
        int64_t t = bsum[i];
        bsum[i] = run;
        run += t;
    }
    if (threadIdx.x == 0) *total = tot;
}

__global__ void __launch_bounds__(BH_SCAN_BLOCK) k_scan_apply(const int32_t* __restrict__ in, int64_t n,
                                                              const int64_t* __restrict__ bsum, const int64_t* __restrict__ total,
                                                              int32_t* __restrict__ out) {
    __shared__ int64_t lds[BH_SCAN_BLOCK / 64 + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * BH_SCAN_TILE + static_cast<int64_t>(threadIdx.x) * BH_SCAN_ITEMS;
    int32_t v[BH_SCAN_ITEMS];
    int64_t s = 0;
#pragma unroll
    for (int i = 0; i < BH_SCAN_ITEMS; ++i) {
        v[i] = (base + i < n) ? in[base + i] : 0;
        s += v[i];
    }
    int64_t run = bh_block_exclusive_scan<BH_SCAN_BLOCK>(s, lds, nullptr) + bsum[blockIdx.x];
#pragma unroll
    for (int i = 0; i < BH_SCAN_ITEMS; ++i) {
        if (base + i < n) out[base + i] = static_cast<int32_t>(run);
        run += v[i];
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) out[n] = static_cast<int32_t>(*total);
}

int bh_exclusive_scan_i32(bh_ctx* ctx, const int32_t* in, int64_t n, int32_t* out, int64_t* total_out,
                          hipStream_t s) {
    const int64_t nb = n > 0 ? (n + BH_SCAN_TILE - 1) / BH_SCAN_TILE : 1;
    DevArray<int64_t> bsum, tot;
    BH_TRY(bsum.alloc(ctx, static_cast<size_t>(nb)));
    BH_TRY(tot.alloc(ctx, 1));
    hipLaunchKernelGGL(k_scan_reduce, dim3(static_cast<unsigned>(nb)), dim3(BH_SCAN_BLOCK), 0, s, in, n, bsum.p);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, s, bsum.p, nb, tot.p);
    int64_t total = 0;
    BH_HIP(ctx, hipMemcpyAsync(&total, tot.p, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipStreamSynchronize(s));
    if (total > INT32_MAX) return bh_set_error(ctx, BH_ERANGE, "prefix sum exceeds the int32 index range");
    hipLaunchKernelGGL(k_scan_apply, dim3(static_cast<unsigned>(nb)), dim3(BH_SCAN_BLOCK), 0, s, in, n, bsum.p, tot.p,
                       out);
    BH_HIP(ctx, hipGetLastError());
    if (total_out) *total_out = total;
    return BH_OK;
}

// ---- radix sort ------------------------------------------------------------------------------------

constexpr int RS_BLOCK = 256;
constexpr int RS_ROUNDS = 8;
constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;

__device__ __forceinline__ int rs_digit(uint64_t lo, uint64_t hi, int shift) {
    return static_cast<int>(((shift < 64) ? (lo >> shift) : (hi >> (shift - 64))) & 0xFFull);
}

// per-tile digit histogram, digit-major layout hist[d * nb + tile]
__global__ void __launch_bounds__(RS_BLOCK) k_rs_hist(const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi,
                                                      int64_t n, int shift, int32_t* __restrict__ hist, int nb) {
    __shared__ int32_t cnt[256];
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = static_cast<int64_t>(blockIdx.x) * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t idx = base + r * RS_BLOCK + threadIdx.x;
        if (idx < n) atomicAdd(&cnt[rs_digit(klo[idx], khi ? khi[idx] : 0ull, shift)], 1);
    }
    __syncthreads();
    hist[static_cast<int64_t>(threadIdx.x) * nb + blockIdx.x] = cnt[threadIdx.x];
}

// Stable scatter.  Elements are taken in rounds of RS_BLOCK consecutive positions; inside a round the
// rank of an element among equal digits comes from a ballot match within its wave and the counts of
// the lower waves, so the output keeps input order for equal digits.
__global__ void __launch_bounds__(RS_BLOCK) k_rs_scatter(const uint64_t* __restrict__ klo, const uint64_t* __restrict__ khi,
                                                         const int32_t* __restrict__ vin, uint64_t* __restrict__ olo,
                                                         uint64_t* __restrict__ ohi, int32_t* __restrict__ vout, int64_t n,
                                                         int shift, const int32_t* __restrict__ offs, int nb) {
    __shared__ int32_t run[256];
    __shared__ int32_t wcnt[RS_BLOCK / 64][256];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wid = tid >> 6;
    run[tid] = offs[static_cast<int64_t>(tid) * nb + blockIdx.x];
#pragma unroll
    for (int w = 0; w < RS_BLOCK / 64; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    const uint64_t lt_mask = (1ull << lane) - 1ull;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * RS_TILE;
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const int64_t idx = base + r * RS_BLOCK + tid;
        const bool active = idx < n;
        uint64_t lo = 0, hi = 0;
        int32_t v = 0;
        int d = 0;
        if (active) {
            lo = klo[idx];
            hi = khi ? khi[idx] : 0ull;
            v = vin[idx];
            d = rs_digit(lo, hi, shift);
        }
        uint64_t peers = __ballot(active);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (d >> b) & 1;
            const uint64_t bb = __ballot(active && bit);
            peers &= bit ? bb : ~bb;
        }
        const int rank = __popcll(peers & lt_mask);
        if (active && rank == 0) wcnt[wid][d] = __popcll(peers);
        __syncthreads();
        if (active) {
            int pos = run[d] + rank;
            for (int w = 0; w < wid; ++w) pos += wcnt[w][d];
            olo[pos] = lo;
            if (ohi) ohi[pos] = hi;
            vout[pos] = v;
        }
        __syncthreads();
        int add = 0;
#pragma unroll
        for (int w = 0; w < RS_BLOCK / 64; ++w) {
            add += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        run[tid] += add;
        __syncthreads();
    }
}

int bh_radix_sort_pairs(bh_ctx* ctx, uint64_t* klo_a, uint64_t* khi_a, int32_t* v_a, uint64_t* klo_b,
                        uint64_t* khi_b, int32_t* v_b, int64_t n, int key_bits, int* result_in_b, hipStream_t s) {
    *result_in_b = 0;
    if (n <= 1 || key_bits <= 0) return BH_OK;
    const int nb = static_cast<int>((n + RS_TILE - 1) / RS_TILE);
    const int64_t nh = static_cast<int64_t>(256) * nb;
    DevArray<int32_t> hist, offs;
    BH_TRY(hist.alloc(ctx, static_cast<size_t>(nh)));
    BH_TRY(offs.alloc(ctx, static_cast<size_t>(nh + 1)));
    const bool wide = key_bits > 64;
    uint64_t *ilo = klo_a, *ihi = wide ? khi_a : nullptr, *olo = klo_b, *ohi = wide ? khi_b : nullptr;
    int32_t *iv = v_a, *ov = v_b;
    int passes = 0;
    for (int shift = 0; shift < key_bits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(nb), dim3(RS_BLOCK), 0, s, ilo, ihi, n, shift, hist.p, nb);
        BH_TRY(bh_exclusive_scan_i32(ctx, hist.p, nh, offs.p, nullptr, s));
        hipLaunchKernelGGL(k_rs_scatter, dim3(nb), dim3(RS_BLOCK), 0, s, ilo, ihi, iv, olo, ohi, ov, n, shift, offs.p, nb);
        BH_HIP(ctx, hipGetLastError());
        std::swap(ilo, olo);
        std::swap(ihi, ohi);
        std::swap(iv, ov);
        ++passes;
    }
    *result_in_b = passes & 1;
    return BH_OK;
}

// ---- basis kernels -----------------------------------------------------------------------------------

// keys of all states by lexicographic rank; WIDE: 128-bit products
template <int NW, int BITS, bool WIDE>
__global__ void __launch_bounds__(256) k_basis_keys(int64_t D, int m, int n, const uint32_t* __restrict__ g,
                                                     const uint64_t* __restrict__ pw_lo, const uint64_t* __restrict__ pw_hi,
                                                     uint64_t* __restrict__ klo, uint64_t* __restrict__ khi,
                                                     int32_t* __restrict__ vals) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= D) return;
    PState<NW, BITS> s;
    bh_unrank(p, m, n, g, s);
    if constexpr (WIDE) {
        unsigned __int128 key = 1;
#pragma unroll
        for (int i = 0; i < PState<NW, BITS>::MAXM; ++i) {
            if (i < m) {
                const int e = s.get(i);
                const size_t t = static_cast<size_t>(i) * (n + 1) + e;
                const unsigned __int128 f = (static_cast<unsigned __int128>(pw_hi[t]) << 64) | pw_lo[t];
                key *= f;
            }
        }
        klo[p] = static_cast<uint64_t>(key);
        khi[p] = static_cast<uint64_t>(key >> 64);
    } else {
        uint64_t key = 1;
#pragma unroll
        for (int i = 0; i < PState<NW, BITS>::MAXM; ++i)
            if (i < m) key *= pw_lo[static_cast<size_t>(i) * (n + 1) + s.get(i)];
        klo[p] = key;
    }
    vals[p] = static_cast<int32_t>(p);
}

// sortperm[t] = lex rank of the t-th smallest key.
// reference order: the state of lex rank p sits at position sortperm[p];
// ascending order: position t holds the state of lex rank sortperm[t].
__global__ void k_basis_perm(int64_t D, const int32_t* __restrict__ sortperm, int order, int32_t* __restrict__ lex_of_pos,
                             int32_t* __restrict__ pos_of_lex) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= D) return;
    const int32_t v = sortperm[i];
    if (order == BH_ORDER_LEX) {
        lex_of_pos[i] = static_cast<int32_t>(i);
        pos_of_lex[i] = static_cast<int32_t>(i);
    } else if (order == BH_ORDER_REFERENCE) {
        pos_of_lex[i] = v;
        lex_of_pos[v] = static_cast<int32_t>(i);
    } else {
        lex_of_pos[i] = v;
        pos_of_lex[v] = static_cast<int32_t>(i);
    }
}

template <int NW, int BITS>
__global__ void __launch_bounds__(256) k_basis_states(int64_t D, int m, int n, const uint32_t* __restrict__ g,
                                                       const int32_t* __restrict__ lex_of_pos, uint8_t* __restrict__ out) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= D) return;
    PState<NW, BITS> s;
    bh_unrank(lex_of_pos[k], m, n, g, s);
    for (int i = 0; i < m; ++i) out[k * m + i] = static_cast<uint8_t>(s.get(i));
}

// ---- dispatch over the instantiated state layouts ---------------------------------------------------

#define BH_LAYOUT_DISPATCH(nw, bits, MACRO)               \
    do {                                                  \
        if ((nw) == 1 && (bits) == 4) MACRO(1, 4);        \
        else if ((nw) == 2 && (bits) == 4) MACRO(2, 4);   \
        else if ((nw) == 2 && (bits) == 8) MACRO(2, 8);   \
        else MACRO(4, 8);                                 \
    } while (0)

extern "C" {

int bh_basis_build(bh_ctx* ctx, int m, int n, int order, bh_basis** out) {
    if (!ctx || !out) return BH_EINVAL;
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    if (m < 1 || m > BH_MAX_SITES || n < 0 || n > 255)
        return bh_set_error(ctx, BH_EINVAL, "basis: need 1 <= m <= 32 and 0 <= n <= 255");
    if (order != BH_ORDER_REFERENCE && order != BH_ORDER_ASCENDING && order != BH_ORDER_LEX)
        return bh_set_error(ctx, BH_EINVAL, "basis: unknown order");
    const int64_t D = bh_binom_host(m + n - 1, n);
    if (D <= 0 || D > INT32_MAX) return bh_set_error(ctx, BH_ERANGE, "basis: dimension exceeds the int32 range");
    const int key_bits = bh_key_bits(m, n);
    if (key_bits > 128) return bh_set_error(ctx, BH_ERANGE, "basis: ordering key prod p_i^n_i exceeds 128 bits");
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;

    std::unique_ptr<bh_basis> bp(new bh_basis());
    bh_basis* b = bp.get();
    b->ctx = ctx;
    b->m = m;
    b->n = n;
    b->order = order;
    b->D = D;
    b->layout_id = bh_next_layout_id();
    bh_state_layout(m, n, &b->nw, &b->bits);

    // rank table g[l][R] = C(R + m-l-2, m-l-1), l in [0, m-2], R in [0, n]; values < D
    const int rows = m > 1 ? m - 1 : 1;
    std::vector<uint32_t> g(static_cast<size_t>(rows) * (n + 1), 0u);
    for (int l = 0; l + 1 < m; ++l)
        for (int R = 0; R <= n; ++R) g[static_cast<size_t>(l) * (n + 1) + R] = static_cast<uint32_t>(bh_binom_host(R + m - l - 2, m - l - 1));
    // prime powers p_i^e, e in [0, n], as 128-bit (lo, hi)
    std::vector<uint64_t> plo(static_cast<size_t>(m) * (n + 1)), phi(static_cast<size_t>(m) * (n + 1));
    for (int i = 0; i < m; ++i) {
        unsigned __int128 v = 1;
        for (int e = 0; e <= n; ++e) {
            plo[static_cast<size_t>(i) * (n + 1) + e] = static_cast<uint64_t>(v);
            phi[static_cast<size_t>(i) * (n + 1) + e] = static_cast<uint64_t>(v >> 64);
            if (e < n) {
                // p_i^e <= p_{m-1}^n < 2^128 for every e <= n, so this never wraps when used
                v *= static_cast<unsigned __int128>(bh_primes[i]);
            }
        }
    }
    const bool wide = key_bits > 64;
    int rc = BH_OK;
    {
        DevArray<uint32_t> dg;
        DevArray<uint64_t> dplo, dphi, klo_a, khi_a, klo_b, khi_b;
        DevArray<int32_t> v_a, v_b, lop, pol;
        do {
            if ((rc = dg.alloc(ctx, g.size())) != BH_OK) break;
            if ((rc = dplo.alloc(ctx, plo.size())) != BH_OK) break;
            if ((rc = dphi.alloc(ctx, phi.size())) != BH_OK) break;
            if ((rc = klo_a.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
            if ((rc = klo_b.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
            if (wide) {
                if ((rc = khi_a.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
                if ((rc = khi_b.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
            }
            if ((rc = v_a.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
            if ((rc = v_b.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
            if ((rc = lop.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
            if ((rc = pol.alloc(ctx, static_cast<size_t>(D))) != BH_OK) break;
            BH_HIP(ctx, hipMemcpyAsync(dg.p, g.data(), g.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipMemcpyAsync(dplo.p, plo.data(), plo.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
            BH_HIP(ctx, hipMemcpyAsync(dphi.p, phi.data(), phi.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s));
            const unsigned grid = static_cast<unsigned>((D + 255) / 256);
            {
                KTimer kt(ctx, "basis_keys", s);
#define LAUNCH_KEYS(NW_, BITS_)                                                                                  \
    do {                                                                                                         \
        if (wide)                                                                                                \
            hipLaunchKernelGGL((k_basis_keys<NW_, BITS_, true>), dim3(grid), dim3(256), 0, s, D, m, n, dg.p,     \
                               dplo.p, dphi.p, klo_a.p, khi_a.p, v_a.p);                                         \
        else                                                                                                     \
            hipLaunchKernelGGL((k_basis_keys<NW_, BITS_, false>), dim3(grid), dim3(256), 0, s, D, m, n, dg.p,    \
                               dplo.p, dphi.p, klo_a.p, nullptr, v_a.p);                                         \
    } while (0)
                BH_LAYOUT_DISPATCH(b->nw, b->bits, LAUNCH_KEYS);
#undef LAUNCH_KEYS
            }
            BH_HIP(ctx, hipGetLastError());
            int in_b = 0;
            {
                KTimer kt(ctx, "basis_sort", s);
                if ((rc = bh_radix_sort_pairs(ctx, klo_a.p, khi_a.p, v_a.p, klo_b.p, khi_b.p, v_b.p, D, key_bits, &in_b,
                                              s)) != BH_OK)
                    break;
            }
            const int32_t* sortperm = in_b ? v_b.p : v_a.p;
            hipLaunchKernelGGL(k_basis_perm, dim3(grid), dim3(256), 0, s, D, sortperm, order, lop.p, pol.p);
            BH_HIP(ctx, hipGetLastError());
            BH_HIP(ctx, hipStreamSynchronize(s));
            b->rank_tab = dg.release();
            b->lex_of_pos = lop.release();
            b->pos_of_lex = pol.release();
            b->bytes = g.size() * sizeof(uint32_t) + 2 * static_cast<size_t>(D) * sizeof(int32_t);
        } while (0);
    }
    if (rc != BH_OK) return rc;
    *out = bp.release();
    return BH_OK;
}

int bh_basis_destroy(bh_basis* b) {
    if (!b) return BH_EINVAL;
    bh_ctx* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    const int rows = b->m > 1 ? b->m - 1 : 1;
    bh_dev_free(ctx, b->rank_tab, static_cast<size_t>(rows) * (b->n + 1) * sizeof(uint32_t));
    bh_dev_free(ctx, b->lex_of_pos, static_cast<size_t>(b->D) * sizeof(int32_t));
    bh_dev_free(ctx, b->pos_of_lex, static_cast<size_t>(b->D) * sizeof(int32_t));
    delete b;
    return BH_OK;
}

int64_t bh_basis_dim(const bh_basis* b) { return b ? b->D : -1; }

int bh_basis_copy_states(const bh_basis* b, uint8_t* states) {
    if (!b || !states) return BH_EINVAL;
    bh_ctx* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    DevArray<uint8_t> out;
    BH_TRY(out.alloc(ctx, static_cast<size_t>(b->D) * b->m));
    const unsigned grid = static_cast<unsigned>((b->D + 255) / 256);
#define LAUNCH_STATES(NW_, BITS_)                                                                                  \
    hipLaunchKernelGGL((k_basis_states<NW_, BITS_>), dim3(grid), dim3(256), 0, s, b->D, b->m, b->n, b->rank_tab,  \
                       b->lex_of_pos, out.p)
    BH_LAYOUT_DISPATCH(b->nw, b->bits, LAUNCH_STATES);
#undef LAUNCH_STATES
    BH_HIP(ctx, hipGetLastError());
    BH_HIP(ctx, hipMemcpyAsync(states, out.p, static_cast<size_t>(b->D) * b->m, hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipStreamSynchronize(s));
    return BH_OK;
}

int bh_basis_copy_tags(const bh_basis* b, double* tags) {
    if (!b || !tags) return BH_EINVAL;
    std::vector<uint8_t> st(static_cast<size_t>(b->D) * b->m);
    BH_TRY(bh_basis_copy_states(b, st.data()));
    // same accumulation order as BH::calculate_tag (src/hamiltonian.cpp:91-97)
    double lp[BH_MAX_SITES];
    for (int i = 0; i < b->m; ++i) lp[i] = std::log(static_cast<double>(bh_primes[i]));
    for (int64_t k = 0; k < b->D; ++k) {
        double t = 0.0;
        for (int i = 0; i < b->m; ++i) t += static_cast<double>(st[static_cast<size_t>(k) * b->m + i]) * lp[i];
        tags[k] = t;
    }
    return BH_OK;
}

int bh_basis_copy_perm(const bh_basis* b, int32_t* lex_of_pos, int32_t* pos_of_lex) {
    if (!b) return BH_EINVAL;
    bh_ctx* ctx = b->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    if (lex_of_pos)
        BH_HIP(ctx, hipMemcpyAsync(lex_of_pos, b->lex_of_pos, static_cast<size_t>(b->D) * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
    if (pos_of_lex)
        BH_HIP(ctx, hipMemcpyAsync(pos_of_lex, b->pos_of_lex, static_cast<size_t>(b->D) * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

}  // extern "C"
