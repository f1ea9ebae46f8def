#include <atomic>
// bh_runtime.hip -- context, errors, budgeted device memory, kernel timing, lattices.
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "bh_internal.hpp"

const int bh_primes[BH_MAX_SITES] = {2,  3,  5,  7,  11, 13, 17, 19, 23, 29,  31,  37,  41,  43,  47,  53,
                                     59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131};

int bh_set_error(const bh_ctx* ctx, int code, const std::string& msg) {
    if (ctx) ctx->err = msg;
    return code;
}

int bh_dev_alloc(bh_ctx* ctx, void** p, size_t bytes) {
    *p = nullptr;
    if (ctx->used + bytes > ctx->budget)
        return bh_set_error(ctx, BH_EBUDGET,
                            "HBM budget exceeded: need " + std::to_string(bytes) + " B, used " +
                                std::to_string(ctx->used) + " of " + std::to_string(ctx->budget));
    const auto t0 = std::chrono::steady_clock::now();
    hipError_t e = hipMalloc(p, bytes);
    ctx->n_alloc += 1;
    ctx->alloc_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (e != hipSuccess) {
        *p = nullptr;
        return bh_set_error(ctx, BH_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    ctx->used += bytes;
    return BH_OK;
}

void bh_dev_free(bh_ctx* ctx, void* p, size_t bytes) {
    if (!p) return;
    const auto t0 = std::chrono::steady_clock::now();
    (void)hipFree(p);
    ctx->n_free += 1;
    ctx->free_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->used = ctx->used >= bytes ? ctx->used - bytes : 0;
}

uint64_t bh_next_layout_id() {
    static std::atomic<uint64_t> next{1};
    return next.fetch_add(1);
}

int bh_ws_get(bh_ctx* ctx, int slot, size_t bytes, void** p) {
    if (slot < 0 || slot >= bh_ctx::NWS) return bh_set_error(ctx, BH_EINVAL, "workspace slot");
    if (bytes == 0) bytes = 1;
    if (ctx->ws_bytes[slot] < bytes) {
        bh_dev_free(ctx, ctx->ws[slot], ctx->ws_bytes[slot]);
        ctx->ws[slot] = nullptr;
        ctx->ws_bytes[slot] = 0;
        void* q = nullptr;
        int rc = bh_dev_alloc(ctx, &q, bytes);
        if (rc != BH_OK) return rc;
        ctx->ws[slot] = q;
        ctx->ws_bytes[slot] = bytes;
    }
    *p = ctx->ws[slot];
    return BH_OK;
}

KTimer::KTimer(bh_ctx* c, const char* n, hipStream_t st) : ctx(c), name(n), s(st) {
    if (ctx && ctx->timing) {
        if (hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess) {
            (void)hipEventRecord(a, s);
        } else {
            a = b = nullptr;
        }
    }
}

KTimer::~KTimer() {
    if (a && b && hipEventRecord(b, s) == hipSuccess) {
        ctx->ktimers.push_back({name, a, b});
        if (ctx->ktimers.size() >= 4096) bh_ktimers_flush(ctx);
        return;
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
}

void bh_ktimers_flush(bh_ctx* ctx) {
    for (const auto& t : ctx->ktimers) {
        if (hipEventSynchronize(t.b) == hipSuccess) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
                auto& st = ctx->kstats[t.name];
                st.first += 1;
                st.second += ms;
            }
        }
        (void)hipEventDestroy(t.a);
        (void)hipEventDestroy(t.b);
    }
    ctx->ktimers.clear();
}

bool bh_tune_enabled() {
    static const bool on = [] {
        const char* e = std::getenv("BHX_TUNE");
        return !(e && std::strcmp(e, "0") == 0);
    }();
    return on;
}

int64_t bh_binom_host(int n, int k) {
    if (k < 0 || n < 0 || k > n) return 0;
    if (k > n - k) k = n - k;
    unsigned __int128 r = 1;
    for (int i = 1; i <= k; ++i) {
        r = r * static_cast<unsigned __int128>(n - k + i) / static_cast<unsigned __int128>(i);
        if (r > static_cast<unsigned __int128>(INT64_MAX)) return -1;
    }
    return static_cast<int64_t>(r);
}

int bh_key_bits(int m, int n) {
    // exact bit length of p_{m-1}^n; 129 when it does not fit in 128 bits
    const unsigned __int128 maxv = ~static_cast<unsigned __int128>(0);
    const unsigned __int128 p = static_cast<unsigned __int128>(bh_primes[m - 1]);
    unsigned __int128 k = 1;
    for (int a = 0; a < n; ++a) {
        if (k > maxv / p) return 129;
        k *= p;
    }
    int bits = 0;
    while (k) {
        ++bits;
        k >>= 1;
    }
    return bits;
}

int bh_state_layout(int m, int n, int* nw, int* bits) {
    // instantiated layouts: (1 word, 4 bits) (2, 4) (2, 8) (4, 8)
    if (n <= 15) {
        *bits = 4;
        *nw = m <= 16 ? 1 : 2;
    } else {
        *bits = 8;
        *nw = m <= 16 ? 2 : 4;
    }
    return BH_OK;
}

extern "C" {

int bh_ctx_create(int device, size_t hbm_budget, bh_ctx** out) {
    if (!out) return BH_EINVAL;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BH_EHIP;
    if (device < 0 || device >= ndev) return BH_EINVAL;
    if (hipSetDevice(device) != hipSuccess) return BH_EHIP;
    bh_ctx* c = new bh_ctx();
    c->device = device;
    size_t freeb = 0, totb = 0;
    if (hipMemGetInfo(&freeb, &totb) != hipSuccess) {
        delete c;
        return BH_EHIP;
    }
    c->budget = hbm_budget ? hbm_budget : static_cast<size_t>(0.9 * static_cast<double>(freeb));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return BH_EHIP;
    }
    *out = c;
    return BH_OK;
}

int bh_ctx_destroy(bh_ctx* ctx) {
    if (!ctx) return BH_EINVAL;
    {
        std::lock_guard<std::recursive_mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        (void)hipStreamSynchronize(ctx->stream);
        bh_ktimers_flush(ctx);
        for (int i = 0; i < bh_ctx::NWS; ++i) bh_dev_free(ctx, ctx->ws[i], ctx->ws_bytes[i]);
        (void)hipStreamDestroy(ctx->stream);
    }
    delete ctx;
    return BH_OK;
}

const char* bh_last_error(const bh_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int bh_ctx_memory(const bh_ctx* ctx, size_t* used, size_t* budget) {
    if (!ctx) return BH_EINVAL;
    if (used) *used = ctx->used;
    if (budget) *budget = ctx->budget;
    return BH_OK;
}

int bh_ctx_set_timing(bh_ctx* ctx, int on) {
    if (!ctx) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    ctx->timing = on ? 1 : 0;
    return BH_OK;
}

int bh_ctx_kernel_stats(bh_ctx* ctx, const char* name, int64_t* count, double* total_ms, int reset) {
    if (!ctx || !name) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    bh_ktimers_flush(ctx);
    auto it = ctx->kstats.find(name);
    if (count) *count = it == ctx->kstats.end() ? 0 : it->second.first;
    if (total_ms) *total_ms = it == ctx->kstats.end() ? 0.0 : it->second.second;
    if (reset) ctx->kstats.clear();
    return BH_OK;
}

int bh_device_name(int device, char* buf, int len) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device) != hipSuccess) return BH_EHIP;
    std::snprintf(buf, static_cast<size_t>(len), "%s (%s, %d CUs)", p.name, p.gcnArchName, p.multiProcessorCount);
    return BH_OK;
}

// Neighbour lists.  Chain: src/neighbours.cpp:21-34 (left, right, then the periodic pair).
// Square: src/neighbours.cpp:40-71 order (left, right, top, bottom), written to the real lists.
int bh_lattice_neighbours(int kind, int Lx, int Ly, int periodic, int* off, int* sites) {
    if (!off || !sites) return -BH_EINVAL;
    if (kind == 0) {
        const int m = Lx;
        if (m < 1 || m > BH_MAX_SITES) return -BH_EINVAL;
        std::vector<std::vector<int>> nb(m);
        for (int i = 0; i < m; ++i) {
            if (i > 0) nb[i].push_back(i - 1);
            if (i < m - 1) nb[i].push_back(i + 1);
        }
        if (periodic) {
            nb[0].push_back(m - 1);
            nb[m - 1].push_back(0);
        }
        int t = 0;
        for (int i = 0; i < m; ++i) {
            off[i] = t;
            for (int s : nb[i]) sites[t++] = s;
        }
        off[m] = t;
        return t;
    }
    if (kind == 1) {
        if (Lx < 1 || Ly < 1 || Lx * Ly > BH_MAX_SITES) return -BH_EINVAL;
        int t = 0;
        for (int i = 0; i < Ly; ++i)
            for (int j = 0; j < Lx; ++j) {
                const int idx = i * Lx + j;
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
        off[Lx * Ly] = t;
        return t;
    }
    return -BH_EINVAL;
}

}  // extern "C"
