// bh_spmv.hpp -- launch interface of the operator kernels (shared with the Lanczos driver).
#pragma once

#include "bh_internal.hpp"

constexpr int BH_TAB_MAX = 4096;  // entries of the sqrt(code) table held in LDS (code_bits <= 12)

template <typename XT>
struct SpmvArgs {
    int64_t D;
    const int32_t* row_ptr;
    const uint32_t* words;  // packed format
    const int32_t* col64;   // FP64 CSR
    const double* val64;
    const float* diag32;    // one of diag32 / diag64
    const double* diag64;
    const double* table;    // sqrt(code)
    int cb;
    const double* J;        // [B]
    const double* U;        // [B]
    const XT* x;            // [D][B]
    XT* y;                  // [D][B]
    int xcd;                // XCD-aware tile order
};

void bh_launch_spmv(const SpmvArgs<float>& a, int B, int G, hipStream_t s);
void bh_launch_spmv(const SpmvArgs<double>& a, int B, int G, int fmt, hipStream_t s);
int bh_pick_lanes(const bh_ham* h, int B);
double bh_spmv_bytes(const bh_ham* h, int prec, int B);
int bh_is_pow2_le64(int B);
int bh_ham_prepare_locked(bh_ham* h, int prec);
void bh_launch_spmv_rb(const SpmvArgs<float>& a, int B, const int32_t* rb, int64_t nb, hipStream_t s);
void bh_launch_spmv_rb(const SpmvArgs<double>& a, int B, int fmt, const int32_t* rb, int64_t nb, hipStream_t s);
int bh_apply(bh_ctx* ctx, bh_ham* h, const SpmvArgs<double>* a64, const SpmvArgs<float>* a32, int fmt, int B,
             hipStream_t s);
