// bh_meanfield.hip -- batched self-consistent mean-field solves of the Bose-Hubbard model (gfx950).
//
// Replaces Analysis::SCMF / Analysis::SF_density (src/analysis.cpp:95-178) and BH::h_MF
// (src/hamiltonian.cpp:290-309).  Each (mu, J) point is an independent fixed-point iteration on psi, each
// step a sequence of ground states of the tridiagonal single-site Hamiltonian h_MF on 2p+1 occupations
// (p = 1, 2, ... until the ground energy moves by at most eps).  The work is scalar FP64 with a
// data-dependent trip count per point, so one thread carries one point: a wavefront runs 64 neighbouring
// grid points, whose iteration counts are close except near the Mott-lobe boundary, and the hardware
// refills a workgroup slot as soon as its points finish.
//
// Arithmetic: every expression keeps one fixed operand order and the file's functions are compiled without
// floating-point contraction, so each result is a fixed sequence of correctly rounded IEEE operations
// (+, -, *, / and the sqrt(i) table computed on the host).  A ground state needs only a few O(2p+1) pivot
// recurrences (no eigenvector, no per-thread arrays), so every point's state lives in registers.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bh_internal.hpp"

namespace {

constexpr int MF_PMAX = 32;             // largest p a point may reach: h_MF on 65 occupations
constexpr int MF_NMAX = 2 * MF_PMAX + 1;
constexpr int MF_WG = 256;
constexpr int MF_GRID = 2048;           // workgroups of one launch: 8 per CU; later points are claimed

__device__ __forceinline__ double mf_diag(int i, double mu, double w) {
#pragma clang fp contract(off)
    return -mu * static_cast<double>(i) + 0.5 * static_cast<double>(i) * static_cast<double>(i - 1) + w;
}

// LDL^T pivots u_i of T - x I and their x-derivatives: true when every pivot is positive (x below the lowest
// eigenvalue), with S = sum_i u_i'/u_i = d/dx ln det(T - x I).  Squared couplings e_{i-1}^2 = c^2 i (c2 = c^2).
//   u_0 = d_0 - x, u_i = (d_i - x) - e_{i-1}^2 / u_{i-1};  u_0' = -1, u_i' = -1 + e_{i-1}^2 u_{i-1}' / u_{i-1}^2
__device__ bool mf_pivots(int N, double mu, double w, double c2, double x, double& S) {
#pragma clang fp contract(off)
    double u = mf_diag(0, mu, w) - x, du = -1.0;
    if (!(u > 0.0)) return false;
    double r = 1.0 / u;
    double s = du * r;
    for (int i = 1; i < N; ++i) {
        const double ee = c2 * static_cast<double>(i);
        du = -1.0 + ee * du * r * r;
        u = (mf_diag(i, mu, w) - x) - ee * r;
        if (!(u > 0.0)) return false;
        r = 1.0 / u;
        s = s + du * r;
    }
    S = s;
    return true;
}

// Logarithmic derivatives of det(T - x I) in x and in the coupling scale c (diagonal held fixed) at two points
// x1, x2 at once (two independent recurrences interleaved); ok1 / ok2 tell whether every pivot was positive.
// d(e_{i-1}^2)/dc = 2 c i (tc = 2c);  u_i'(c) = -(2 c i) / u_{i-1} + e_{i-1}^2 u_{i-1}'(c) / u_{i-1}^2.
__device__ void mf_logdet_deriv2(int N, double mu, double w, double c2, double tc, double x1, double x2, bool& ok1,
                                 bool& ok2, double& Sx1, double& Sc1, double& Sx2, double& Sc2) {
#pragma clang fp contract(off)
    const double d0 = mf_diag(0, mu, w);
    double u1 = d0 - x1, u2 = d0 - x2, dux1 = -1.0, dux2 = -1.0, duc1 = 0.0, duc2 = 0.0;
    ok1 = u1 > 0.0;
    ok2 = u2 > 0.0;
    double r1 = 1.0 / u1, r2 = 1.0 / u2;
    double sx1 = dux1 * r1, sx2 = dux2 * r2, sc1 = 0.0, sc2 = 0.0;
    for (int i = 1; i < N && (ok1 || ok2); ++i) {
        const double ee = c2 * static_cast<double>(i);
        const double dee = tc * static_cast<double>(i);
        const double di = mf_diag(i, mu, w);
        dux1 = -1.0 + ee * dux1 * r1 * r1;
        dux2 = -1.0 + ee * dux2 * r2 * r2;
        duc1 = -(dee * r1) + ee * duc1 * r1 * r1;
        duc2 = -(dee * r2) + ee * duc2 * r2 * r2;
        u1 = (di - x1) - ee * r1;
        u2 = (di - x2) - ee * r2;
        ok1 = ok1 && u1 > 0.0;
        ok2 = ok2 && u2 > 0.0;
        r1 = 1.0 / u1;
        r2 = 1.0 / u2;
        sx1 = sx1 + dux1 * r1;
        sx2 = sx2 + dux2 * r2;
        sc1 = sc1 + duc1 * r1;
        sc2 = sc2 + duc2 * r2;
    }
    Sx1 = sx1;
    Sc1 = sc1;
    Sx2 = sx2;
    Sc2 = sc2;
}

// Ground state of h_MF(psi, p):  d_i = -mu i + i(i-1)/2 + q J psi^2,  h(i, i+1) = c sqrt(i+1), c = -q J psi.
// Returns its energy and sf = <phi|a|phi>, all in registers (no eigenvector is formed).
// Energy: Newton's iteration on ln det(T - x I) from below (x <- x - 1/S).  Every root is real, so each exact
// step from below the lowest root stays below it and the iterates rise monotonically to it (quadratically
// once close).  Starts at `hint` when its pivots are all positive (the previous p's eigenvalue less 1e-3
// (|lambda| + 1): more occupations only lower it), else at the Gershgorin lower bound; hi = min_i d_i, or the
// rejected hint, lies at or above the root.  An iterate without all pivots positive, or a step reaching hi,
// can only come from rounding at the root: that iterate (resp. hi) is returned, as is the last iterate when
// a step no longer moves x.
// <a>: Hellmann-Feynman, d lambda/dc = 2 sum_i sqrt(i+1) phi_i phi_{i+1} = 2 <a> and d lambda/dc =
// -(d ln det/dc) / (d ln det/dx) as x -> lambda; the ratio is taken at lambda - delta and lambda - 2 delta,
// delta = 1e-9 (|lambda| + 1), and extrapolated linearly to lambda (O(delta^2) error).
// c = 0 (diagonal h): the first smallest diagonal entry and <a> = 0.
__device__ double mf_ground(double psi, int p, double mu, double J, int q, const double* __restrict__ sq, double hint,
                            double& sf) {
#pragma clang fp contract(off)
    const int N = 2 * p + 1;
    const double c = static_cast<double>(-q) * J * psi;
    const double w = static_cast<double>(q) * J * psi * psi;
    double lo = 0.0, hi = 0.0;
    for (int i = 0; i < N; ++i) {
        const double d = mf_diag(i, mu, w);
        const double a0 = i > 0 ? fabs(c * sq[i]) : 0.0;
        const double a1 = i < N - 1 ? fabs(c * sq[i + 1]) : 0.0;
        const double g = d - a0 - a1;
        if (i == 0 || g < lo) lo = g;
        if (i == 0 || d < hi) hi = d;
    }
    sf = 0.0;
    if (c == 0.0) return hi;
    const double c2 = c * c, tc = 2.0 * c;
    double S = 0.0;
    // the start: the hint when its pivots are all positive, else the Gershgorin bound (moved down if rounding
    // put it at the eigenvalue)
    double xk = lo;
    bool have_start = false;
    if (hint > lo && hint < hi) {
        if (mf_pivots(N, mu, w, c2, hint, S)) {
            xk = hint;
            have_start = true;
        } else {
            hi = hint;
        }
    }
    if (!have_start) {
        for (int t = 0; t < 64 && !mf_pivots(N, mu, w, c2, lo, S); ++t) lo = lo - (fabs(lo) + 1.0);
        xk = lo;
    }
    double lam = xk;
    for (int it = 0; it < 200; ++it) {
        if (!mf_pivots(N, mu, w, c2, xk, S)) break;  // a Newton iterate at the root within rounding
        const double xn = xk - 1.0 / S;
        if (!(xn < hi)) {  // the root lies at hi within rounding
            lam = hi;
            break;
        }
        if (!(xn > xk)) break;  // no progress
        xk = xn;
        lam = xk;
    }
    double delta = 1e-9 * (fabs(lam) + 1.0);
    for (int t = 0; t < 8; ++t) {
        double sx1, sc1, sx2, sc2;
        bool ok1, ok2;
        mf_logdet_deriv2(N, mu, w, c2, tc, lam - delta, lam - 2.0 * delta, ok1, ok2, sx1, sc1, sx2, sc2);
        if (ok1 && ok2) {
            const double f1 = -0.5 * sc1 / sx1;
            const double f2 = -0.5 * sc2 / sx2;
            sf = 2.0 * f1 - f2;
            break;
        }
        delta = 2.0 * delta;
    }
    return lam;
}

// the start for p + 1 after the eigenvalue e of p
__device__ __forceinline__ double mf_hint(double e) {
#pragma clang fp contract(off)
    return e - 1e-3 * (fabs(e) + 1.0);
}

// Analysis::SCMF with its two stopping rules (ground energy over p, psi over the self-consistency loop),
// U = 1, as a per-lane state machine: every pass of the loop is one ground state per lane, after which the
// lane either moves to the next p, starts the next self-consistency iteration, or finishes its point (status
// 0 converged, 1 max_outer iterations reached with psi the last iterate, 2 p would exceed pmax) and takes the
// next unclaimed point.  Points near the Mott-lobe boundary need up to ~100x the iterations of their
// neighbours; per-lane refill keeps the other lanes of the wavefront busy meanwhile.  Each lane starts at
// its global thread index; later points are claimed in wavefront batches from *next (zero at launch), offset
// by the number of threads.  A wavefront leaves the loop when none of its lanes holds a point.
__global__ void __launch_bounds__(MF_WG) k_scmf(int64_t npts, const double* __restrict__ mu_in,
                                                const double* __restrict__ J_in, const double* __restrict__ psi0,
                                                int q, double eps, int max_outer, int pmax,
                                                const double* __restrict__ sq_in, double* __restrict__ psi_out,
                                                int32_t* __restrict__ status_out, int32_t* __restrict__ iters_out,
                                                int32_t* __restrict__ p_out, unsigned long long* __restrict__ next) {
#pragma clang fp contract(off)
    __shared__ double sq[MF_NMAX + 1];
    for (int i = threadIdx.x; i <= MF_NMAX; i += MF_WG) sq[i] = sq_in[i];
    __syncthreads();
    const int64_t nthreads = static_cast<int64_t>(gridDim.x) * MF_WG;
    const int lane = static_cast<int>(__lane_id());
    int64_t k = static_cast<int64_t>(blockIdx.x) * MF_WG + threadIdx.x;
    double mu = 0.0, J = 0.0, psi = 0.0, e0 = 0.0, sf = 0.0;
    int p = 1, outer = 1;
    bool have = k < npts, exhausted = !have;
    if (have) {
        mu = mu_in[k];
        J = J_in[k];
        psi = psi0[k];
    }
    for (;;) {
        if (have) {
            const double e = mf_ground(psi, p, mu, J, q, sq, p == 1 ? -INFINITY : mf_hint(e0), sf);
            int status = -1;
            if (p == 1) {
                e0 = e;
                p = 2;
            } else {
                const double tmp = fabs(e0 - e);
                e0 = e;
                if (tmp > eps) {
                    if (p + 1 > pmax)
                        status = 2;
                    else
                        ++p;
                } else {
                    const double dpsi = fabs(sf - psi);
                    psi = sf;
                    if (dpsi <= eps) {
                        status = 0;
                    } else if (outer >= max_outer) {
                        status = 1;
                    } else {
                        ++outer;
                        p = 1;
                    }
                }
            }
            if (status >= 0) {
                psi_out[k] = psi;
                status_out[k] = status;
                iters_out[k] = outer;
                p_out[k] = p;
                have = false;
            }
        }
        // refill: the lanes without a point claim consecutive points with one atomic per wavefront
        const unsigned long long need = __ballot(!have && !exhausted);
        if (need) {
            const int cnt = __popcll(need);
            const int leader = __ffsll(static_cast<long long>(need)) - 1;
            unsigned long long base = 0;
            if (lane == leader) base = atomicAdd(next, static_cast<unsigned long long>(cnt));
            base = __shfl(base, leader);
            if (!have && !exhausted) {
                const int rank = __popcll(need & ((1ull << lane) - 1ull));
                k = nthreads + static_cast<int64_t>(base) + rank;
                if (k < npts) {
                    have = true;
                    mu = mu_in[k];
                    J = J_in[k];
                    psi = psi0[k];
                    p = 1;
                    outer = 1;
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__ballot(have)) break;
    }
}

}  // namespace

extern "C" {

int bh_mean_field(bh_ctx* ctx, int64_t npts, const double* mu, const double* J, const double* psi0, int q,
                  double eps, int max_outer, int pmax, double* psi, int32_t* status, int32_t* iters,
                  int32_t* p_last) {
    if (!ctx) return BH_EINVAL;
    if (npts < 0 || (npts > 0 && (!mu || !J || !psi0 || !psi)))
        return bh_set_error(ctx, BH_EINVAL, "mean_field: need npts >= 0 and the mu, J, psi0, psi arrays");
    if (pmax < 2 || pmax > MF_PMAX)
        return bh_set_error(ctx, BH_EINVAL, "mean_field: need 2 <= pmax <= " + std::to_string(MF_PMAX));
    if (max_outer < 1) return bh_set_error(ctx, BH_EINVAL, "mean_field: need max_outer >= 1");
    if (!(eps >= 0.0) || std::isinf(eps)) return bh_set_error(ctx, BH_EINVAL, "mean_field: need finite eps >= 0");
    if (npts > (static_cast<int64_t>(1) << 31) * MF_WG / 2)
        return bh_set_error(ctx, BH_ERANGE, "mean_field: too many points for one launch");
    if (npts == 0) return BH_OK;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    BH_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = ctx->stream;
    const size_t n = static_cast<size_t>(npts);
    std::vector<double> sqh(MF_NMAX + 1);
    for (int i = 0; i <= MF_NMAX; ++i) sqh[i] = std::sqrt(static_cast<double>(i));
    DevArray<double> dmu, dJ, dpsi0, dpsi, dsq;
    DevArray<int32_t> dst, dit, dp;
    DevArray<unsigned long long> dnext;
    BH_TRY(dmu.alloc(ctx, n));
    BH_TRY(dJ.alloc(ctx, n));
    BH_TRY(dpsi0.alloc(ctx, n));
    BH_TRY(dpsi.alloc(ctx, n));
    BH_TRY(dsq.alloc(ctx, MF_NMAX + 1));
    BH_TRY(dst.alloc(ctx, n));
    BH_TRY(dit.alloc(ctx, n));
    BH_TRY(dp.alloc(ctx, n));
    BH_TRY(dnext.alloc(ctx, 1));
    const unsigned long long zero = 0;
    BH_HIP(ctx, hipMemcpyAsync(dmu.p, mu, n * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(dJ.p, J, n * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(dpsi0.p, psi0, n * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(dsq.p, sqh.data(), sqh.size() * 8, hipMemcpyHostToDevice, s));
    BH_HIP(ctx, hipMemcpyAsync(dnext.p, &zero, sizeof(zero), hipMemcpyHostToDevice, s));
    {
        KTimer kt(ctx, "scmf", s);
        const unsigned nb = static_cast<unsigned>(std::min<int64_t>((npts + MF_WG - 1) / MF_WG, MF_GRID));
        hipLaunchKernelGGL(k_scmf, dim3(nb), dim3(MF_WG), 0, s, npts, dmu.p, dJ.p, dpsi0.p, q, eps, max_outer, pmax,
                           dsq.p, dpsi.p, dst.p, dit.p, dp.p, dnext.p);
    }
    BH_HIP(ctx, hipGetLastError());
    BH_HIP(ctx, hipMemcpyAsync(psi, dpsi.p, n * 8, hipMemcpyDeviceToHost, s));
    if (status) BH_HIP(ctx, hipMemcpyAsync(status, dst.p, n * 4, hipMemcpyDeviceToHost, s));
    if (iters) BH_HIP(ctx, hipMemcpyAsync(iters, dit.p, n * 4, hipMemcpyDeviceToHost, s));
    if (p_last) BH_HIP(ctx, hipMemcpyAsync(p_last, dp.p, n * 4, hipMemcpyDeviceToHost, s));
    BH_HIP(ctx, hipStreamSynchronize(s));
    return BH_OK;
}

}  // extern "C"
