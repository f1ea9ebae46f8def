// bh_dense.hip -- dense FP64 eigensolve of a symmetric sparse matrix on the device.
//
// Replaces Op::exact_eigen (src/operator.cpp:92-101), which copies the sparse Hamiltonian into a dense
// Eigen matrix and runs Eigen's SelfAdjointEigenSolver (tridiagonalisation + implicit QR) on the host.
// Here the CSR is scattered into a dense column-major matrix in HBM by one kernel and the symmetric
// eigenproblem is solved by rocSOLVER's divide-and-conquer dsyevd on the same stream; eigenvalues come
// back ascending, eigenvectors as the columns of the (overwritten) dense matrix, as Eigen returns them.
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <cstring>
#include <string>

#include "bh_internal.hpp"

namespace {

// A[col * n + row] += val for the CSR entries of the rows of this thread (A zeroed before).  Each entry
// (row, col) of a CSR is written once, so the scatter needs no atomics.
__global__ void __launch_bounds__(256) k_dense_scatter(int64_t n, const int64_t* __restrict__ rp, const int32_t* __restrict__ col,
                                                       const double* __restrict__ val, double* __restrict__ A) {
    const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int64_t q = rp[r]; q < rp[r + 1]; ++q) A[static_cast<int64_t>(col[q]) * n + r] += val[q];
}

}  // namespace

extern "C" int bh_dense_eigen(bh_ctx* ctx, int64_t n, const int64_t* row_ptr, const int32_t* col, const double* val,
                              double* evals, double* evecs) {
    if (!ctx || n < 1 || !row_ptr || !evals || (row_ptr[n] > 0 && (!col || !val))) return BH_EINVAL;
    if (n > 2147483647 / 2) return bh_set_error(ctx, BH_ERANGE, "dense_eigen: n too large for a dense solve");
    for (int64_t r = 0; r < n; ++r) {
        if (row_ptr[r + 1] < row_ptr[r]) return bh_set_error(ctx, BH_EINVAL, "dense_eigen: row_ptr not ascending");
        for (int64_t q = row_ptr[r]; q < row_ptr[r + 1]; ++q)
            if (col[q] < 0 || col[q] >= n) return bh_set_error(ctx, BH_EINVAL, "dense_eigen: column out of range");
    }
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    const int64_t nnz = row_ptr[n];
    DevArray<double> A, w, e, dval;
    DevArray<int64_t> drp;
    DevArray<int32_t> dcol;
    DevArray<int> info;
    BH_TRY(A.alloc(ctx, static_cast<size_t>(n) * n));
    BH_TRY(w.alloc(ctx, n));
    BH_TRY(e.alloc(ctx, n));
    BH_TRY(drp.alloc(ctx, n + 1));
    BH_TRY(dcol.alloc(ctx, nnz > 0 ? nnz : 1));
    BH_TRY(dval.alloc(ctx, nnz > 0 ? nnz : 1));
    BH_TRY(info.alloc(ctx, 1));
    hipStream_t s = ctx->stream;
    BH_HIP(ctx, hipMemcpyAsync(drp.p, row_ptr, (n + 1) * 8, hipMemcpyHostToDevice, s));
    if (nnz > 0) {
        BH_HIP(ctx, hipMemcpyAsync(dcol.p, col, nnz * 4, hipMemcpyHostToDevice, s));
        BH_HIP(ctx, hipMemcpyAsync(dval.p, val, nnz * 8, hipMemcpyHostToDevice, s));
    }
    BH_HIP(ctx, hipMemsetAsync(A.p, 0, static_cast<size_t>(n) * n * 8, s));
    hipLaunchKernelGGL(k_dense_scatter, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, s, n, drp.p, dcol.p, dval.p,
                       A.p);
    BH_HIP(ctx, hipGetLastError());
    rocblas_handle hnd = nullptr;
    if (rocblas_create_handle(&hnd) != rocblas_status_success) return bh_set_error(ctx, BH_EHIP, "rocblas_create_handle");
    struct HandleGuard {
        rocblas_handle h;
        ~HandleGuard() { rocblas_destroy_handle(h); }
    } hg{hnd};
    if (rocblas_set_stream(hnd, s) != rocblas_status_success) return bh_set_error(ctx, BH_EHIP, "rocblas_set_stream");
    const rocblas_status st = rocsolver_dsyevd(hnd, evecs ? rocblas_evect_original : rocblas_evect_none, rocblas_fill_lower,
                                               static_cast<rocblas_int>(n), A.p, static_cast<rocblas_int>(n), w.p, e.p, info.p);
    if (st != rocblas_status_success)
        return bh_set_error(ctx, BH_EHIP, "rocsolver_dsyevd: status " + std::to_string(static_cast<int>(st)));
    int hinfo = 0;
    BH_HIP(ctx, hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipMemcpyAsync(evals, w.p, n * 8, hipMemcpyDeviceToHost, s));
    if (evecs) BH_HIP(ctx, hipMemcpyAsync(evecs, A.p, static_cast<size_t>(n) * n * 8, hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipStreamSynchronize(s));
    if (hinfo != 0) return bh_set_error(ctx, BH_ENOCONV, "rocsolver_dsyevd: no convergence (info " + std::to_string(hinfo) + ")");
    return BH_OK;
}
