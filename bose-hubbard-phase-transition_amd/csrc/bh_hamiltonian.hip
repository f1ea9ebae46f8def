// bh_hamiltonian.hip -- two-pass (count / exclusive scan / fill) CSR build of the hopping term T and
// the diagonal interaction V over a neighbour list, in the basis order of a bh_basis.
//
// Reference: BH::fill_hopping (src/hamiltonian.cpp:170-191) and BH::fill_interaction (:194-211),
// selected by BH::fixed_bosons_hamiltonian (:238-256).  The intended operator is built: for every
// site i and every entry j of i's neighbour list with n_j >= 1, the state a_i^+ a_j |k> receives
// -sqrt((n_i+1) n_j); repeated neighbour entries add up, as repeated triplets do in
// setFromTriplets (:190).  The target's column is found exactly: lexicographic rank of the target
// (prefix sums of binomials) -> position through the basis permutation.
//
// Storage: one 32-bit word per entry, column << code_bits | code, with code = (c w)^2 (n_i+1) n_j
// (w: multiplicity of the neighbour entry, c: 2 under BH_CONV_REF_HOP, else 1), so that the value
// -c w sqrt((n_i+1) n_j) = -sqrt(code) is read exactly from a small FP64 table.  Columns ascending per
// row, as in Eigen's compressed storage.
#include <algorithm>
#include <cmath>
#include <map>
#include <memory>

#include <cstdlib>
#include <cstring>

#include "bh_scan.hpp"
#include "bh_spmv.hpp"

constexpr int HB_BLOCK = 64;  // one wave per block: per-thread LDS row buffers stay small

struct EdgeArgs {
    int E;
    const int8_t* ei;    // target site i (gains a boson)
    const int8_t* ej;    // source site j (loses a boson)
    const uint32_t* cm;  // code multiplier (c w)^2
};

// adds delta (+1 or -1) to site i; the field never under/overflows (callers check occupations)
template <int NW, int BITS>
__device__ __forceinline__ void pstate_add(PState<NW, BITS>& s, int i, int delta) {
    const int wi = i / PState<NW, BITS>::PER;
    const int sh = (i % PState<NW, BITS>::PER) * BITS;
    const uint64_t d = static_cast<uint64_t>(static_cast<int64_t>(delta)) << sh;  // two's complement wrap
#pragma unroll
    for (int a = 0; a < NW; ++a)
        if (a == wi) s.w[a] += d;
}

// lexicographic rank of a state; g is the (m-1) x (n+1) table in LDS
template <int NW, int BITS>
__device__ __forceinline__ int64_t pstate_rank(const PState<NW, BITS>& t, int m, int n, const uint32_t* g) {
    int R = n;
    int64_t r = 0;
#pragma unroll
    for (int l = 0; l < PState<NW, BITS>::MAXM - 1; ++l) {
        if (l < m - 1) {
            R -= t.get(l);
            r += g[l * (n + 1) + R];
        }
    }
    return r;
}

template <int NW, int BITS>
__global__ void __launch_bounds__(HB_BLOCK) k_ham_count(int64_t D, int m, int n, const uint32_t* __restrict__ g,
                                                         const int32_t* __restrict__ lex_of_pos, EdgeArgs ed,
                                                         int32_t* __restrict__ rownnz) {
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= D) return;
    PState<NW, BITS> s;
    bh_unrank(lex_of_pos[q], m, n, g, s);
    int c = 0;
    for (int e = 0; e < ed.E; ++e) c += s.get(ed.ej[e]) >= 1;
    rownnz[q] = c;
}

template <int NW, int BITS>
__global__ void __launch_bounds__(HB_BLOCK) k_ham_fill(int64_t D, int m, int n, const uint32_t* __restrict__ g_glob,
                                                        int gsize, const int32_t* __restrict__ lex_of_pos,
                                                        const int32_t* __restrict__ pos_of_lex, EdgeArgs ed, int cb,
                                                        unsigned flags, const int32_t* __restrict__ row_ptr,
                                                        uint32_t* __restrict__ words, float* __restrict__ diag) {
    extern __shared__ uint32_t lds[];
    uint32_t* g = lds;                                   // gsize words
    const int stride = ed.E + 1;                         // odd stride: no LDS bank aliasing between lanes
    uint32_t* ent = lds + gsize + threadIdx.x * stride;  // this thread's row buffer
    for (int a = threadIdx.x; a < gsize; a += blockDim.x) g[a] = g_glob[a];
    __syncthreads();
    const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= D) return;
    PState<NW, BITS> s;
    bh_unrank(lex_of_pos[q], m, n, g, s);
    int cnt = 0;
    for (int e = 0; e < ed.E; ++e) {
        const int i = ed.ei[e], j = ed.ej[e];
        const int nj = s.get(j);
        if (nj < 1) continue;
        const int ni = s.get(i);
        PState<NW, BITS> t = s;
        pstate_add(t, i, 1);
        pstate_add(t, j, -1);
        const int64_t tl = pstate_rank(t, m, n, g);
        const uint32_t col = static_cast<uint32_t>(pos_of_lex[tl]);
        const uint32_t code = ed.cm[e] * static_cast<uint32_t>((ni + 1) * nj);
        // insertion into the sorted row buffer (distinct targets: distinct (i,j) pairs)
        const uint32_t wv = (col << cb) | code;
        int a = cnt - 1;
        while (a >= 0 && ent[a] > wv) {
            ent[a + 1] = ent[a];
            --a;
        }
        ent[a + 1] = wv;
        ++cnt;
    }
    const int32_t base = row_ptr[q];
    for (int a = 0; a < cnt; ++a) words[base + a] = ent[a];
    float v = 0.f;
    for (int i = 0; i < m; ++i) {
        const int ni = s.get(i);
        v += (flags & BH_CONV_REF_INTERACTION) ? static_cast<float>(ni * (ni + 1)) : static_cast<float>((ni * (ni - 1)) / 2);
    }
    diag[q] = v;
}

// Single-particle density matrix rho_ij = <phi| b_i^+ b_j |phi> (real phi), partial sums per workgroup.
// A workgroup is one wave; it walks tiles of 64 basis positions (tile = blockIdx.x, + gridDim.x, ...). A
// thread unranks its state once per tile and, for every ordered pair (i, j), forms phi_q n_i phi_q (i = j)
// or phi_q sqrt((n_i + 1) n_j) phi_t with t the position of b_i^+ b_j s_q (n_j >= 1); the wave sums each pair
// over its 64 positions (fixed shuffle order) and lane 0 adds it to the workgroup's LDS accumulator, tile
// after tile, so the result is reproducible.  part[(i m + j) * gridDim.x + blockIdx.x] receives the sums.
template <int NW, int BITS>
__global__ void __launch_bounds__(HB_BLOCK) k_spdm(int64_t D, int m, int n, const uint32_t* __restrict__ g_glob, int gsize,
                                                    const int32_t* __restrict__ lex_of_pos,
                                                    const int32_t* __restrict__ pos_of_lex, const double* __restrict__ phi,
                                                    double* __restrict__ part) {
    extern __shared__ uint32_t lds[];
    uint32_t* g = lds;                                     // gsize words
    double* acc = reinterpret_cast<double*>(lds + ((gsize + 1) & ~1));  // m * m sums (8-byte aligned)
    for (int a = threadIdx.x; a < gsize; a += blockDim.x) g[a] = g_glob[a];
    for (int a = threadIdx.x; a < m * m; a += blockDim.x) acc[a] = 0.0;
    __syncthreads();
    const int64_t ntile = (D + HB_BLOCK - 1) / HB_BLOCK;
    for (int64_t tile = blockIdx.x; tile < ntile; tile += gridDim.x) {
        const int64_t q = tile * HB_BLOCK + threadIdx.x;
        const bool valid = q < D;
        PState<NW, BITS> s;
        s.clear();
        if (valid) bh_unrank(lex_of_pos[q], m, n, g, s);
        const double pq = valid ? phi[q] : 0.0;
        for (int i = 0; i < m; ++i) {
            const int ni = s.get(i);
            for (int j = 0; j < m; ++j) {
                double c = 0.0;
                if (valid) {
                    if (i == j) {
                        c = pq * pq * ni;
                    } else {
                        const int nj = s.get(j);
                        if (nj >= 1) {
                            PState<NW, BITS> t = s;
                            pstate_add(t, i, 1);
                            pstate_add(t, j, -1);
                            const int64_t tl = pstate_rank(t, m, n, g);
                            c = pq * phi[pos_of_lex[tl]] * std::sqrt(static_cast<double>((ni + 1) * nj));
                        }
                    }
                }
#pragma unroll
                for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o, 64);
                if (threadIdx.x == 0) acc[i * m + j] += c;
            }
        }
    }
    __syncthreads();
    for (int a = threadIdx.x; a < m * m; a += blockDim.x) part[static_cast<int64_t>(a) * gridDim.x + blockIdx.x] = acc[a];
}

// out[p] = sum over the nb workgroup partials of value p, in order (one workgroup per value)
__global__ void __launch_bounds__(256) k_spdm_sum(const double* __restrict__ part, int nb, double* __restrict__ out) {
    __shared__ double red[4];
    double s = 0.0;
    for (int w = threadIdx.x; w < nb; w += 256) s += part[static_cast<int64_t>(blockIdx.x) * nb + w];
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ void k_ham_expand64(int64_t nnz, const uint32_t* __restrict__ words, int cb, const double* __restrict__ table,
                               int32_t* __restrict__ col, double* __restrict__ val) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    const uint32_t w = words[k];
    col[k] = static_cast<int32_t>(w >> cb);
    val[k] = -table[w & ((1u << cb) - 1u)];
}

__global__ void k_f32_to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b) {
    const int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k < n) b[k] = static_cast<double>(a[k]);
}

// out[0]: longest row; out[1]: largest row sum of |T| (FP32 bits, non-negative); out[2], out[3]: min, max of V
__global__ void k_row_stats(int64_t D, const int32_t* __restrict__ row_ptr, const uint32_t* __restrict__ words, int cb,
                            const double* __restrict__ table, const float* __restrict__ diag, int* __restrict__ out) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= D) return;
    const uint32_t mask = (1u << cb) - 1u;
    double s = 0.0;
    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) s += table[words[k] & mask];
    atomicMax(out, row_ptr[r + 1] - row_ptr[r]);
    atomicMax(out + 1, __float_as_int(static_cast<float>(s) * 1.0001f));
    atomicMin(out + 2, __float_as_int(diag[r]));
    atomicMax(out + 3, __float_as_int(diag[r]));
}

#define BH_LAYOUT_DISPATCH(nw, bits, MACRO)             \
    do {                                                \
        if ((nw) == 1 && (bits) == 4) MACRO(1, 4);      \
        else if ((nw) == 2 && (bits) == 4) MACRO(2, 4); \
        else if ((nw) == 2 && (bits) == 8) MACRO(2, 8); \
        else MACRO(4, 8);                               \
    } while (0)

int bh_ham_rowblocks(bh_ham* h, int B, const int32_t** rb, int64_t* nblocks) {
    auto it = h->rowblocks.find(B);
    if (it != h->rowblocks.end()) {
        *rb = it->second.first;
        *nblocks = it->second.second;
        return BH_OK;
    }
    bh_ctx* ctx = h->ctx;
    const int64_t cap = BH_STAGE_PRODUCTS_HOST / B;
    if (static_cast<int64_t>(h->max_row) > cap) return bh_set_error(ctx, BH_ERANGE, "row longer than a staged block");
    std::vector<int32_t> rp(static_cast<size_t>(h->D + 1));
    BH_HIP(ctx, hipMemcpyAsync(rp.data(), h->row_ptr, rp.size() * 4, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    // (first row, first entry) of every block, then (D, nnz): one load gives a block's whole extent
    std::vector<int32_t> blocks;
    auto add = [&](int64_t r) {
        blocks.push_back(static_cast<int32_t>(r));
        blocks.push_back(rp[r]);
    };
    add(0);
    int64_t start = 0;
    for (int64_t r = 0; r < h->D; ++r) {
        if (rp[r + 1] - rp[start] > cap) {
            add(r);
            start = r;
        }
    }
    add(h->D);
    DevArray<int32_t> d;
    BH_TRY(d.alloc(ctx, blocks.size()));
    BH_HIP(ctx, hipMemcpyAsync(d.p, blocks.data(), blocks.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    BH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    const int64_t nb = static_cast<int64_t>(blocks.size()) / 2 - 1;
    h->bytes_rb += blocks.size() * 4;
    h->rowblocks[B] = {d.release(), nb};
    *rb = h->rowblocks[B].first;
    *nblocks = nb;
    return BH_OK;
}

extern "C" int bh_ham_tuned(const bh_ham* h, int kind, int prec, int B, int* variant, int* ncand, int* cand_var,
                            double* cand_us, int cap) {
    if (!h || kind < 0 || kind > 1) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
    auto it = h->tuned.find(bh_tune_key(kind, prec, B));
    if (it == h->tuned.end()) return bh_set_error(h->ctx, BH_EINVAL, "sector not used at this (kind, precision, width)");
    if (variant) *variant = it->second.variant;
    const int n = static_cast<int>(it->second.cand.size());
    if (ncand) *ncand = n;
    for (int i = 0; i < n && i < cap; ++i) {
        if (cand_var) cand_var[i] = it->second.cand[i].first;
        if (cand_us) cand_us[i] = it->second.cand[i].second;
    }
    return BH_OK;
}

extern "C" int bh_ham_set_tuned(bh_ham* h, int kind, int prec, int B, int variant) {
    if (!h || kind < 0 || kind > 1) return BH_EINVAL;
    if (!bh_is_pow2_le64(B) || (prec != BH_PREC_FP64 && prec != BH_PREC_MIXED))
        return bh_set_error(h->ctx, BH_EINVAL, "set_tuned: B a power of two <= 64, precision FP64 or MIXED");
    bool ok = false;
    if (variant == 0) ok = static_cast<int64_t>(h->max_row) * B <= BH_STAGE_PRODUCTS_HOST;
    else if (variant == 1 || variant == 2) ok = variant <= B;
    else if (variant == 4) ok = B >= 4 && prec == BH_PREC_MIXED;
    else if (variant == 8) ok = kind == 0 && B >= 8 && prec == BH_PREC_MIXED;
    else if (variant == 17 || variant == 33 || variant == 18 || variant == 34) ok = kind == 0 && (variant & 15) <= B && B <= 16;
    else if (variant == 20) ok = kind == 0 && B >= 4 && B <= 32 && prec == BH_PREC_MIXED;
    else if (variant > 100 && kind == 1) {
        const int G = variant - 100;
        ok = (G & (G - 1)) == 0 && G * B <= 64;
    }
    if (!ok) return bh_set_error(h->ctx, BH_EINVAL, "set_tuned: variant does not fit this sector and width");
    std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
    BhTuned t;
    t.variant = variant;
    h->tuned[bh_tune_key(kind, prec, B)] = t;
    return BH_OK;
}

int bh_engine_env() {
    static const int env = [] {
        const char* e = std::getenv("BHX_ENGINE");
        return e ? std::atoi(e) : -1;
    }();
    return env;
}

int bh_engine_choice(int B, const bh_ham* h) {
    const int env = bh_engine_env();
    const bool staged_ok = static_cast<int64_t>(h->max_row) * B <= BH_STAGE_PRODUCTS_HOST;
    if (env == 0) return 0;
    if (env == 1) return staged_ok ? 1 : 2;
    if (env == 2) return 2;
    return (B == 1 && staged_ok) ? 1 : 2;
}

static int free_ham(bh_ham* h) {
    bh_ctx* ctx = h->ctx;
    for (auto& kv : h->rowblocks) bh_dev_free(ctx, kv.second.first, static_cast<size_t>(kv.second.second + 1) * 8);
    const size_t tab = (static_cast<size_t>(1) << h->code_bits);
    bh_dev_free(ctx, h->row_ptr, static_cast<size_t>(h->D + 1) * sizeof(int32_t));
    bh_dev_free(ctx, h->words, static_cast<size_t>(h->nnz ? h->nnz : 1) * sizeof(uint32_t));
    bh_dev_free(ctx, h->diag, static_cast<size_t>(h->D) * sizeof(float));
    bh_dev_free(ctx, h->table, tab * sizeof(double));
    bh_dev_free(ctx, h->table32, tab * sizeof(float));
    bh_dev_free(ctx, h->col64, static_cast<size_t>(h->nnz ? h->nnz : 1) * sizeof(int32_t));
    bh_dev_free(ctx, h->val64, static_cast<size_t>(h->nnz ? h->nnz : 1) * sizeof(double));
    bh_dev_free(ctx, h->diag64, static_cast<size_t>(h->D) * sizeof(double));
    bh_dev_free(ctx, h->lex_of_pos, static_cast<size_t>(h->D) * sizeof(int32_t));
    delete h;
    return BH_OK;
}

int bh_ham_prepare_locked(bh_ham* h, int prec) {
    bh_ctx* ctx = h->ctx;
    if (prec != BH_PREC_FP64 || h->col64) return BH_OK;
    const size_t nz = static_cast<size_t>(h->nnz ? h->nnz : 1);
    DevArray<int32_t> c;
    DevArray<double> v, d;
    BH_TRY(c.alloc(ctx, nz));
    BH_TRY(v.alloc(ctx, nz));
    BH_TRY(d.alloc(ctx, static_cast<size_t>(h->D)));
    hipStream_t s = ctx->stream;
    if (h->nnz)
        hipLaunchKernelGGL(k_ham_expand64, dim3(static_cast<unsigned>((h->nnz + 255) / 256)), dim3(256), 0, s, h->nnz,
                           h->words, h->code_bits, h->table, c.p, v.p);
    hipLaunchKernelGGL(k_f32_to_f64, dim3(static_cast<unsigned>((h->D + 255) / 256)), dim3(256), 0, s, h->D, h->diag, d.p);
    BH_HIP(ctx, hipGetLastError());
    BH_HIP(ctx, hipStreamSynchronize(s));
    h->col64 = c.release();
    h->val64 = v.release();
    h->diag64 = d.release();
    h->bytes64 = nz * (sizeof(int32_t) + sizeof(double)) + static_cast<size_t>(h->D) * sizeof(double);
    return BH_OK;
}

extern "C" {

int bh_hamiltonian_build(bh_ctx* ctx, const bh_basis* b, const int* nbr_off, const int* nbr_sites, unsigned flags,
                         bh_ham** out) {
    if (!ctx || !b || !nbr_off || !nbr_sites || !out) return BH_EINVAL;
    *out = nullptr;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int m = b->m, n = b->n;
    const int64_t D = b->D;
    if (flags & ~(BH_CONV_REF_HOP | BH_CONV_REF_INTERACTION)) return bh_set_error(ctx, BH_EINVAL, "hamiltonian: unknown flags");
    // neighbour list: range, no self loops, symmetric as a multiset
    if (nbr_off[0] != 0) return bh_set_error(ctx, BH_EINVAL, "hamiltonian: nbr_off[0] must be 0");
    std::map<std::pair<int, int>, int> mult;  // (i <- j) -> multiplicity
    for (int i = 0; i < m; ++i) {
        if (nbr_off[i + 1] < nbr_off[i]) return bh_set_error(ctx, BH_EINVAL, "hamiltonian: nbr_off not monotone");
        for (int a = nbr_off[i]; a < nbr_off[i + 1]; ++a) {
            const int j = nbr_sites[a];
            if (j < 0 || j >= m) return bh_set_error(ctx, BH_EINVAL, "hamiltonian: neighbour site out of range");
            if (j == i) return bh_set_error(ctx, BH_EINVAL, "hamiltonian: self-loop in the neighbour list");
            mult[{i, j}] += 1;
        }
    }
    for (const auto& kv : mult) {
        auto it = mult.find({kv.first.second, kv.first.first});
        if (it == mult.end() || it->second != kv.second)
            return bh_set_error(ctx, BH_EINVAL, "hamiltonian: neighbour list is not symmetric");
    }
    const int E = static_cast<int>(mult.size());
    if (E > 128) return bh_set_error(ctx, BH_ERANGE, "hamiltonian: more than 128 directed bonds");
    const uint32_t hop = (flags & BH_CONV_REF_HOP) ? 2u : 1u;
    std::vector<int8_t> ei, ej;
    std::vector<uint32_t> cm;
    uint32_t cm_max = 1;
    for (const auto& kv : mult) {
        ei.push_back(static_cast<int8_t>(kv.first.first));
        ej.push_back(static_cast<int8_t>(kv.first.second));
        const uint32_t w = static_cast<uint32_t>(kv.second) * hop;
        cm.push_back(w * w);
        cm_max = std::max(cm_max, w * w);
    }
    // largest (n_i + 1) n_j with n_i + n_j <= n, n_j >= 1
    uint64_t qmax = 0;
    for (int nj = 1; nj <= n; ++nj) qmax = std::max<uint64_t>(qmax, static_cast<uint64_t>(n - nj + 1) * nj);
    const uint64_t code_max = std::max<uint64_t>(1, qmax * cm_max);
    int cb = 0;
    while ((code_max >> cb) != 0) ++cb;
    if (cb > 12 || (D > 1 && static_cast<uint64_t>(D - 1) >> (32 - cb)) != 0)
        return bh_set_error(ctx, BH_ERANGE, "hamiltonian: column and value code do not fit one 32-bit word");
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;

    std::unique_ptr<bh_ham> hp(new bh_ham());
    bh_ham* h = hp.get();
    h->ctx = ctx;
    h->basis = b;
    h->D = D;
    h->code_bits = cb;
    h->flags = flags;

    const size_t tab = static_cast<size_t>(1) << cb;
    std::vector<double> table(tab);
    std::vector<float> table32(tab);
    for (size_t c = 0; c < tab; ++c) {
        table[c] = std::sqrt(static_cast<double>(c));
        table32[c] = static_cast<float>(table[c]);
    }
    const int gsize = (m > 1 ? m - 1 : 1) * (n + 1);

    DevArray<int8_t> dei, dej;
    DevArray<uint32_t> dcm, words, dtab_dummy;
    DevArray<int32_t> rownnz, row_ptr;
    DevArray<float> diag, t32;
    DevArray<double> t64;
    DevArray<int> dmax;
    BH_TRY(dei.alloc(ctx, std::max(E, 1)));
    BH_TRY(dej.alloc(ctx, std::max(E, 1)));
    BH_TRY(dcm.alloc(ctx, std::max(E, 1)));
    BH_TRY(rownnz.alloc(ctx, static_cast<size_t>(D)));
    BH_TRY(row_ptr.alloc(ctx, static_cast<size_t>(D + 1)));
    BH_TRY(diag.alloc(ctx, static_cast<size_t>(D)));
    BH_TRY(t64.alloc(ctx, tab));
    BH_TRY(t32.alloc(ctx, tab));
    BH_TRY(dmax.alloc(ctx, 4));
    if (E) {
        BH_HIP(ctx, hipMemcpyAsync(dei.p, ei.data(), E, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dej.p, ej.data(), E, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dcm.p, cm.data(), E * sizeof(uint32_t), hipMemcpyHostToDevice, s));
    }
    BH_HIP(ctx, hipMemcpyAsync(t64.p, table.data(), tab * sizeof(double), hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(t32.p, table32.data(), tab * sizeof(float), hipMemcpyHostToDevice, s));
    EdgeArgs ed{E, dei.p, dej.p, dcm.p};
    const unsigned grid = static_cast<unsigned>((D + HB_BLOCK - 1) / HB_BLOCK);
    {
        KTimer kt(ctx, "ham_count", s);
#define LAUNCH_COUNT(NW_, BITS_)                                                                                   \
    hipLaunchKernelGGL((k_ham_count<NW_, BITS_>), dim3(grid), dim3(HB_BLOCK), 0, s, D, m, n, b->rank_tab,         \
                       b->lex_of_pos, ed, rownnz.p)
        BH_LAYOUT_DISPATCH(b->nw, b->bits, LAUNCH_COUNT);
#undef LAUNCH_COUNT
    }
    BH_HIP(ctx, hipGetLastError());
    int64_t nnz = 0;
    BH_TRY(bh_exclusive_scan_i32(ctx, rownnz.p, D, row_ptr.p, &nnz, s));
    rownnz.reset();
    BH_TRY(words.alloc(ctx, static_cast<size_t>(nnz ? nnz : 1)));
    const size_t lds_bytes = (static_cast<size_t>(gsize) + static_cast<size_t>(HB_BLOCK) * (E + 1)) * sizeof(uint32_t);
    if (lds_bytes > 64 * 1024) return bh_set_error(ctx, BH_ERANGE, "hamiltonian: rank table too large for LDS");
    {
        KTimer kt(ctx, "ham_fill", s);
#define LAUNCH_FILL(NW_, BITS_)                                                                                    \
    hipLaunchKernelGGL((k_ham_fill<NW_, BITS_>), dim3(grid), dim3(HB_BLOCK), lds_bytes, s, D, m, n, b->rank_tab,   \
                       gsize, b->lex_of_pos, b->pos_of_lex, ed, cb, flags, row_ptr.p, words.p, diag.p)
        BH_LAYOUT_DISPATCH(b->nw, b->bits, LAUNCH_FILL);
#undef LAUNCH_FILL
    }
    BH_HIP(ctx, hipGetLastError());
    int stats[4] = {0, 0, 0x7f7fffff, 0};  // V >= 0, so FP32 bit patterns order like the values
    BH_HIP(ctx, hipMemcpyAsync(dmax.p, stats, sizeof(stats), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_row_stats, dim3(static_cast<unsigned>((D + 255) / 256)), dim3(256), 0, s, D, row_ptr.p, words.p,
                       cb, t64.p, diag.p, dmax.p);
    BH_HIP(ctx, hipMemcpyAsync(stats, dmax.p, sizeof(stats), hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipStreamSynchronize(s));
    auto as_f = [](int v) {
        float f;
        std::memcpy(&f, &v, 4);
        return static_cast<double>(f);
    };
    DevArray<int32_t> lop;
    if (bh_basis_lex_start(b)) {
        BH_TRY(lop.alloc(ctx, static_cast<size_t>(D)));
        BH_HIP(ctx, hipMemcpyAsync(lop.p, b->lex_of_pos, static_cast<size_t>(D) * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
    }
    h->nnz = nnz;
    h->max_row = stats[0];
    h->tmax = as_f(stats[1]);
    h->vmin = as_f(stats[2]);
    h->vmax = as_f(stats[3]);
    h->row_ptr = row_ptr.release();
    h->words = words.release();
    h->diag = diag.release();
    h->table = t64.release();
    h->table32 = t32.release();
    h->bytes = static_cast<size_t>(D + 1) * 4 + static_cast<size_t>(nnz ? nnz : 1) * 4 + static_cast<size_t>(D) * 4 +
               tab * 12;
    if (bh_basis_lex_start(b)) {
        h->lex_of_pos = lop.release();
        h->layout_id = b->layout_id;
        h->bytes += static_cast<size_t>(D) * 4;
    }
    *out = hp.release();
    return BH_OK;
}

int bh_ham_destroy(bh_ham* h) {
    if (!h) return BH_EINVAL;
    bh_ctx* ctx = h->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    return free_ham(h);
}

int64_t bh_ham_dim(const bh_ham* h) { return h ? h->D : -1; }
int64_t bh_ham_nnz(const bh_ham* h) { return h ? h->nnz : -1; }
int64_t bh_ham_device_bytes(const bh_ham* h) { return h ? static_cast<int64_t>(h->bytes + h->bytes64 + h->bytes_rb) : -1; }

int bh_ham_prepare(bh_ham* h, int prec) {
    if (!h) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(h->ctx->mu);
    BH_HIP(h->ctx, hipSetDevice(h->ctx->device));
    return bh_ham_prepare_locked(h, prec);
}

int bh_ham_copy_packed(const bh_ham* h, uint32_t* words, int* code_bits) {
    if (!h) return BH_EINVAL;
    bh_ctx* ctx = h->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    if (code_bits) *code_bits = h->code_bits;
    if (words && h->nnz)
        BH_HIP(ctx, hipMemcpyAsync(words, h->words, static_cast<size_t>(h->nnz) * 4, hipMemcpyDeviceToHost, ctx->stream));
    BH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    return BH_OK;
}

int bh_ham_copy_csr(const bh_ham* h, int64_t* row_ptr, int32_t* col, double* val, double* diag) {
    if (!h) return BH_EINVAL;
    bh_ctx* ctx = h->ctx;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    if (row_ptr) {
        std::vector<int32_t> rp(static_cast<size_t>(h->D + 1));
        BH_HIP(ctx, hipMemcpyAsync(rp.data(), h->row_ptr, rp.size() * 4, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (size_t i = 0; i < rp.size(); ++i) row_ptr[i] = rp[i];
    }
    if ((col || val) && h->nnz) {
        std::vector<uint32_t> w(static_cast<size_t>(h->nnz));
        BH_HIP(ctx, hipMemcpyAsync(w.data(), h->words, w.size() * 4, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        const uint32_t mask = (1u << h->code_bits) - 1u;
        for (size_t k = 0; k < w.size(); ++k) {
            if (col) col[k] = static_cast<int32_t>(w[k] >> h->code_bits);
            if (val) val[k] = -std::sqrt(static_cast<double>(w[k] & mask));
        }
    }
    if (diag) {
        std::vector<float> d(static_cast<size_t>(h->D));
        BH_HIP(ctx, hipMemcpyAsync(d.data(), h->diag, d.size() * 4, hipMemcpyDeviceToHost, s));
        BH_HIP(ctx, hipStreamSynchronize(s));
        for (size_t i = 0; i < d.size(); ++i) diag[i] = d[i];
    }
    return BH_OK;
}

// Single-particle density matrix of one state (Analysis::SPDM / braket, src/analysis.cpp:497-512, 562-594).
int bh_spdm(bh_ctx* ctx, const bh_basis* b, const double* phi, double* rho) {
    if (!ctx || !b || !phi || !rho) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const int m = b->m, n = b->n;
    const int64_t D = b->D;
    if (m < 1) return bh_set_error(ctx, BH_EINVAL, "spdm: empty lattice");
    const int gsize = (m > 1 ? m - 1 : 1) * (n + 1);
    const size_t lds_bytes = static_cast<size_t>((gsize + 1) & ~1) * 4 + static_cast<size_t>(m) * m * 8;
    if (lds_bytes > 64 * 1024) return bh_set_error(ctx, BH_ERANGE, "spdm: rank table too large for LDS");
    const int64_t ntile = (D + HB_BLOCK - 1) / HB_BLOCK;
    const int nb = static_cast<int>(std::min<int64_t>(std::max<int64_t>(ntile, 1), 2048));
    DevArray<double> dphi, part, out;
    BH_TRY(dphi.alloc(ctx, static_cast<size_t>(D)));
    BH_TRY(part.alloc(ctx, static_cast<size_t>(m) * m * nb));
    BH_TRY(out.alloc(ctx, static_cast<size_t>(m) * m));
    BH_HIP(ctx, hipMemcpyAsync(dphi.p, phi, static_cast<size_t>(D) * 8, hipMemcpyHostToDevice, s));
#define LAUNCH_SPDM(NW_, BITS_)                                                                                   \
    hipLaunchKernelGGL((k_spdm<NW_, BITS_>), dim3(nb), dim3(HB_BLOCK), lds_bytes, s, D, m, n, b->rank_tab, gsize, \
                       b->lex_of_pos, b->pos_of_lex, dphi.p, part.p)
    BH_LAYOUT_DISPATCH(b->nw, b->bits, LAUNCH_SPDM);
#undef LAUNCH_SPDM
    hipLaunchKernelGGL(k_spdm_sum, dim3(m * m), dim3(256), 0, s, part.p, nb, out.p);
    BH_HIP(ctx, hipGetLastError());
    BH_HIP(ctx, hipMemcpyAsync(rho, out.p, static_cast<size_t>(m) * m * 8, hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipStreamSynchronize(s));
    return BH_OK;
}

}  // extern "C"
