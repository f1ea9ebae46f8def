// bh_kernels.hpp -- device building blocks of the operator application and of the Lanczos step.
//
// Layout: B vectors are stored row-interleaved, x[r * B + b], so that the gather x[c * B + b] of one
// column for all B vectors is one contiguous B-element segment.  In a workgroup, lane l works on
// vector b = l % B, on the g = (l / B) % G-th strided part of its row, and on row slot l / (G * B).
#pragma once

#include "bh_internal.hpp"

constexpr int BH_FMT_PACKED = 0;  // 32-bit word: column << code_bits | code, value -sqrt(code)
constexpr int BH_FMT_CSR64 = 1;   // int32 column + FP64 value

constexpr int BH_WG = 256;  // workgroup size of the operator kernels (4 waves)

template <typename XT>
__device__ __forceinline__ double to_d(XT v) {
    return static_cast<double>(v);
}

// Partial sum of one lane over its strided part of row r: sum_k value_k * x[col_k * B + b].
// For the packed format the table holds +sqrt(code); the caller negates.
template <typename XT, int B, int G, int FMT>
__device__ __forceinline__ double bh_row_partial(int32_t rs, int32_t re, int g, int b, const uint32_t* __restrict__ words,
                                                 const int32_t* __restrict__ col64, const double* __restrict__ val64,
                                                 const double* __restrict__ tab, int cb, const XT* __restrict__ x) {
    // two entries per iteration, both loaded unconditionally (the second repeats the row's last entry past the
    // end and adds a zero), in the order of k
    double acc = 0.0;
    for (int32_t k = rs + g; k < re; k += 2 * G) {
        const int32_t k1 = min(k + G, re - 1);
        double p0, p1;
        if constexpr (FMT == BH_FMT_PACKED) {
            const uint32_t mask = (1u << cb) - 1u;
            const uint32_t w0 = words[k], w1 = words[k1];
            p0 = tab[w0 & mask] * to_d(x[static_cast<int64_t>(w0 >> cb) * B + b]);
            p1 = tab[w1 & mask] * to_d(x[static_cast<int64_t>(w1 >> cb) * B + b]);
        } else {
            p0 = val64[k] * to_d(x[static_cast<int64_t>(col64[k]) * B + b]);
            p1 = val64[k1] * to_d(x[static_cast<int64_t>(col64[k1]) * B + b]);
        }
        acc += p0;
        acc += k + G < re ? p1 : 0.0;
    }
    return acc;
}

// sum over the G lanes that share (row slot, b); every lane of the wave must call it
template <int B, int G>
__device__ __forceinline__ double bh_group_sum(double v) {
#pragma unroll
    for (int o = B; o < B * G; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums NV values per vector b over the workgroup.  Each lane contributes v[0..NV-1] for its
// b = lane % B.  On return, out[q * B + b] (LDS, NV*B doubles) holds the workgroup totals; all lanes
// must call it.  scratch: (BH_WG/64) * B * NV doubles of LDS.
template <int B, int NV>
__device__ __forceinline__ void bh_block_sum_b(double (&v)[NV], double* scratch, double* out) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NV; ++q) {
        double t = v[q];
#pragma unroll
        for (int o = B; o < 64; o <<= 1) t += __shfl_xor(t, o, 64);
        if (lane < B) scratch[(wid * NV + q) * B + lane] = t;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NV * B; i += BH_WG) {
        const int q = i / B, b = i % B;
        double t = 0.0;
        for (int w = 0; w < BH_WG / 64; ++w) t += scratch[(w * NV + q) * B + b];
        out[i] = t;
    }
    __syncthreads();
}

// true in exactly one workgroup: the last one to arrive.  Partials written before the call by any
// workgroup are visible to the last one after it returns.  The counter is reset by the caller.
__device__ __forceinline__ bool bh_last_block(unsigned* counter) {
    __shared__ int amlast;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) {
        const unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        amlast = (prev == gridDim.x - 1) ? 1 : 0;
    }
    __syncthreads();
    if (amlast) __threadfence();
    return amlast != 0;
}

__device__ __forceinline__ double bh_coherent_load(const double* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void bh_reset_counter(unsigned* counter) {
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Row tile of a workgroup with XCD-aware placement.  Workgroups are dispatched round-robin over the
// 8 XCDs of the chip (workgroup w runs on XCD w % 8); tile ids are renumbered so that each XCD walks
// one contiguous eighth of the rows, keeping the x gathers of neighbouring rows in one L2.
__device__ __forceinline__ int64_t bh_xcd_tile(int64_t wg, int64_t nwg) {
    const int64_t per = nwg / 8;
    if (per == 0 || wg >= per * 8) return wg;  // tail workgroups keep their natural index
    return (wg % 8) * per + wg / 8;
}

// ---- row-block engine ------------------------------------------------------------------------------
// A row block is a run of consecutive rows holding at most BH_STAGE_PRODUCTS / B entries.  A workgroup
// takes one block at a time: phase 1 streams the block's entries (coalesced, independent loads and
// gathers, several per lane in flight) and writes the products value * x[col * B + b] to LDS; phase 2
// gives each (row, b) pair to one thread, which sums its row's products from LDS (four interleaved
// partial sums in a fixed order) and hands the sum to the epilogue (as an array of one, the epilogue
// interface of the row-lane engine).  rb[] holds (first row, first entry) of every block, then (D, nnz).
constexpr int BH_STAGE_PRODUCTS = 2048;  // FP64 products per block (16 KB of LDS)

template <typename XT, int B, int FMT, int WG, typename Epi>
__device__ __forceinline__ void bh_rowblock_run(int64_t t0, int64_t t1, const int32_t* __restrict__ rb,
                                                const int32_t* __restrict__ rp, const uint32_t* __restrict__ words,
                                                const int32_t* __restrict__ col64, const double* __restrict__ val64,
                                                const double* __restrict__ tab, int cb, const XT* __restrict__ x,
                                                double* prod, Epi& epi) {
    constexpr int CAP = BH_STAGE_PRODUCTS;  // products per block (entries * B)
    constexpr int PER = CAP / WG;           // products per thread in phase 1
    const uint32_t mask = (1u << cb) - 1u;
    for (int64_t t = t0; t < t1; ++t) {
        const int2 b0 = *reinterpret_cast<const int2*>(rb + 2 * t), b1 = *reinterpret_cast<const int2*>(rb + 2 * t + 2);
        const int32_t r0 = b0.x, ks = b0.y, r1 = b1.x, ke = b1.y;
        const int np = (ke - ks) * B;
        if (np > 0) {  // uniform
            // Products past the block's end repeat its last one and go to LDS slots that phase 2 never reads
            // (i < PER * WG = CAP always): every load and store is unconditional, so that all of a thread's word
            // loads, then all of its gathers, are in flight together.
            double v[PER];
#pragma unroll
            for (int u = 0; u < PER; ++u) {
                const int ic = min(static_cast<int>(threadIdx.x) + u * WG, np - 1);
                const int32_t k = ks + ic / B;
                const int bb = ic % B;
                if constexpr (FMT == BH_FMT_PACKED) {
                    const uint32_t w = words[k];
                    v[u] = tab[w & mask] * to_d(x[static_cast<int64_t>(w >> cb) * B + bb]);
                } else {
                    v[u] = val64[k] * to_d(x[static_cast<int64_t>(col64[k]) * B + bb]);
                }
            }
#pragma unroll
            for (int u = 0; u < PER; ++u) prod[threadIdx.x + u * WG] = v[u];
        }
        __syncthreads();
        const int nq = (r1 - r0) * B;
        for (int i = threadIdx.x; i < nq; i += WG) {
            const int32_t r = r0 + i / B;
            const int bb = i % B;
            const int32_t a = rp[r] - ks, e = rp[r + 1] - ks;
            // four interleaved partial sums (independent LDS reads), combined in a fixed order
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int32_t q = a;
            for (; q + 3 < e; q += 4) {
                s0 += prod[q * B + bb];
                s1 += prod[(q + 1) * B + bb];
                s2 += prod[(q + 2) * B + bb];
                s3 += prod[(q + 3) * B + bb];
            }
            for (; q < e; ++q) s0 += prod[q * B + bb];
            double sum[1] = {(s0 + s1) + (s2 + s3)};
            epi(static_cast<int64_t>(r), bb, sum);
        }
        __syncthreads();
    }
}

// ---- row-lane engine ---------------------------------------------------------------------------------
// B / V consecutive threads share a row; each holds V consecutive vectors b0 .. b0 + V - 1 of it, so an
// entry's word is one broadcast load and its gather one V-element vector load per thread (the row's B
// elements are contiguous).  A wavefront covers 64 V / B rows.  The entry loop is unrolled by U with all
// loads of an unrolled step issued before the multiply-adds; only the words stay live across the
// gathers (the values are looked up afterwards).  Rows [r0, r1) of a workgroup are walked in groups of
// WG V / B.  The sum of every (row, b) runs over the row's entries in storage order whatever V and U are.
template <typename T, int V>
struct alignas(sizeof(T) * V) BhVec {
    T v[V];
};

template <int V, typename T>
__device__ __forceinline__ BhVec<T, V> bh_vload(const T* p) {
    return *reinterpret_cast<const BhVec<T, V>*>(p);
}

template <typename XT, int B, int V, int FMT, int WG, int U, typename Epi>
__device__ __forceinline__ void bh_rowlane_run(int64_t r0, int64_t r1, const int32_t* __restrict__ rp,
                                               const uint32_t* __restrict__ words, const int32_t* __restrict__ col64,
                                               const double* __restrict__ val64, const double* __restrict__ tab, int cb,
                                               const XT* __restrict__ x, Epi& epi) {
    static_assert(B % V == 0, "V divides B");
    constexpr int TPR = B / V;  // threads per row
    const uint32_t mask = (1u << cb) - 1u;
    const int b0 = (threadIdx.x % TPR) * V;
    for (int64_t base = r0; base < r1; base += WG / TPR) {
        const int64_t r = base + threadIdx.x / TPR;
        if (r >= r1) continue;
        const int32_t ks = rp[r], ke = rp[r + 1];
        double acc[V];
#pragma unroll
        for (int v = 0; v < V; ++v) acc[v] = 0.0;
        for (int32_t k = ks; k < ke; k += U) {
            uint32_t w[U];
            double val[U];
            BhVec<XT, V> xv[U];
            // Entries past the row's end read the row's last entry again (a cache hit) with a zero value, so
            // every word load and gather of the unrolled step is unconditional and all of them are in flight
            // together (a load under a branch is waited for before the next one is issued).  Such an entry
            // adds a zero product (the x it gathers is one the row already uses).
            if constexpr (FMT == BH_FMT_PACKED) {
#pragma unroll
                for (int u = 0; u < U; ++u) w[u] = words[min(k + u, ke - 1)];
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = bh_vload<V>(x + static_cast<int64_t>(w[u] >> cb) * B + b0);
#pragma unroll
                for (int u = 0; u < U; ++u) val[u] = k + u < ke ? tab[w[u] & mask] : 0.0;
            } else {
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int32_t kk = min(k + u, ke - 1);
                    w[u] = static_cast<uint32_t>(col64[kk]);
                    const double vv = val64[kk];
                    val[u] = k + u < ke ? vv : 0.0;
                }
#pragma unroll
                for (int u = 0; u < U; ++u) xv[u] = bh_vload<V>(x + static_cast<int64_t>(w[u]) * B + b0);
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int v = 0; v < V; ++v) acc[v] += val[u] * to_d(xv[u].v[v]);
        }
        epi(r, b0, acc);
    }
}
