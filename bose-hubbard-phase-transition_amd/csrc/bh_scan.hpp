// bh_scan.hpp -- block-level scan helpers and the device-wide primitives shared by the basis and
// Hamiltonian builders (exclusive scan, stable LSD radix sort of (key, value) pairs).
#pragma once

#include "bh_internal.hpp"

constexpr int BH_SCAN_BLOCK = 256;
constexpr int BH_SCAN_ITEMS = 8;
constexpr int BH_SCAN_TILE = BH_SCAN_BLOCK * BH_SCAN_ITEMS;

// Exclusive scan of a wave64 block of BLOCK threads.  lds must hold BLOCK/64 + 1 values.
template <int BLOCK>
__device__ __forceinline__ int64_t bh_block_exclusive_scan(int64_t v, int64_t* lds, int64_t* total) {
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    int64_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int64_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) lds[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t run = 0;
        for (int w = 0; w < BLOCK / 64; ++w) {
            int64_t t = lds[w];
            lds[w] = run;
            run += t;
        }
        lds[BLOCK / 64] = run;
    }
    __syncthreads();
    const int64_t res = lds[wid] + x - v;
    if (total) *total = lds[BLOCK / 64];
    __syncthreads();
    return res;
}

// out[0..n] = exclusive prefix sums of in[0..n-1] (out[n] = total).  Fails with BH_ERANGE when the
// total does not fit in int32.  total_out (host) may be NULL.
int bh_exclusive_scan_i32(bh_ctx* ctx, const int32_t* in, int64_t n, int32_t* out, int64_t* total_out,
                          hipStream_t s);

// Stable ascending LSD radix sort of n pairs by the low `key_bits` bits of the key (lo word, and the
// hi word when key_bits > 64).  Input/output buffers ping-pong; the sorted result is returned in
// *_a when the number of 8-bit passes is even, in *_b otherwise (*result_in_b tells which).
int bh_radix_sort_pairs(bh_ctx* ctx, uint64_t* klo_a, uint64_t* khi_a, int32_t* v_a, uint64_t* klo_b,
                        uint64_t* khi_b, int32_t* v_b, int64_t n, int key_bits, int* result_in_b, hipStream_t s);
