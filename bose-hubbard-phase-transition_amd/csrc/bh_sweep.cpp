// bh_sweep.cpp -- batched phase-diagram sweep over the sectors n-1, n, n+1 of one lattice.
//
// Replaces the grid loop of Analysis::calculate_and_save (src/analysis.cpp:266-387): there every grid
// point merges three sparse matrices (:311) and runs its own IRLM (:314) on an OpenMP thread (:302).
// Here each sector's CSR is built once, the J points of a sector go through bh_lanczos in batches of
// a width chosen per sector (one CSR pass per Lanczos step for the whole batch), and mu enters afterwards
// as the N-dependent shift -mu N.  The gap ratio follows Analysis::gap_ratios (src/analysis.cpp:433-454).
#include <cmath>
#include <cstdlib>
#include <vector>

#include "../../include/bh_capi.h"

namespace {

double gap_ratio(const double* ev, int k) {
    std::vector<double> v;
    for (int i = 0; i < k; ++i)
        if (std::isfinite(ev[i])) v.push_back(ev[i]);
    if (v.size() < 3) return NAN;
    double s = 0.0;
    for (size_t i = 1; i + 1 < v.size(); ++i) {
        const double a = v[i] - v[i - 1], b = v[i + 1] - v[i];
        const double mn = a < b ? a : b, mx = a < b ? b : a;
        s += mx != 0.0 ? mn / mx : 0.0;
    }
    return s / static_cast<double>(v.size() - 2);
}

struct Owned {
    bh_basis* b = nullptr;
    bh_ham* h = nullptr;
    ~Owned() {
        if (h) bh_ham_destroy(h);
        if (b) bh_basis_destroy(b);
    }
};

// Block-width tuner of a sector: the operator is timed at widths 8, 16, 32, 64 (up to the batch that
// holds every J point) and the narrowest width whose cost per vector is within 10% of the best is kept.
// A wider batch amortises the matrix stream over more vectors; a narrower one leaves fewer lanes idle
// while the slowest solve of the batch finishes and groups neighbouring J, which converge alike.
int sweep_width(bh_ctx* ctx, bh_ham* h, int prec, int nJ, int* W) {
    int top = 1;
    while (top < nJ) top <<= 1;
    if (top <= 8) {
        *W = top;
        return BH_OK;
    }
    std::vector<int> ws;
    std::vector<double> cost;
    double best = 0.0;
    for (int w = 8; w <= top && w <= 64; w <<= 1) {
        double ms = 0.0, bytes = 0.0;
        const int rc = bh_spmv_bench(ctx, h, prec, w, 5, &ms, &bytes);
        if (rc != BH_OK) return rc;
        ws.push_back(w);
        cost.push_back(ms / w);
        if (best == 0.0 || ms / w < best) best = ms / w;
    }
    *W = top;
    for (size_t i = 0; i < ws.size(); ++i)
        if (cost[i] <= 1.1 * best) {
            *W = ws[i];
            break;
        }
    return BH_OK;
}

}  // namespace

extern "C" int bh_sweep(bh_ctx* ctx, const bh_sweep_spec* sp, const bh_lanczos_opts* opts, bh_sweep_out* out) {
    if (!ctx || !sp || !opts || !out || !sp->J || (sp->nmu > 0 && !sp->mu)) return BH_EINVAL;
    if (sp->nJ < 1 || sp->nJ > 64 || sp->nmu < 0 || sp->n < 0) return BH_EINVAL;
    if (sp->width < 0 || sp->width > 64 || (sp->width & (sp->width - 1)) != 0) return BH_EINVAL;
    const int m = sp->kind == 0 ? sp->Lx : sp->Lx * sp->Ly;
    if (m < 1 || m > 32) return BH_EINVAL;
    std::vector<int> off(m + 1), sites(4 * m + 4);
    const int ne = bh_lattice_neighbours(sp->kind, sp->Lx, sp->Ly, sp->periodic, off.data(), sites.data());
    if (ne < 0) return BH_EINVAL;
    const int k = opts->k, nJ = sp->nJ;
    std::vector<double> Uv(nJ, sp->U);
    for (int q = 0; q < 3; ++q) {
        const int N = sp->n - 1 + q;
        if (out->width) out->width[q] = 0;
        for (int i = 0; i < nJ; ++i) {
            if (out->status) out->status[q * nJ + i] = BH_OK;
            if (out->used_fp64) out->used_fp64[q * nJ + i] = 0;
            for (int l = 0; l < k; ++l) out->E[(q * nJ + i) * k + l] = NAN;
            out->gap_ratio[q * nJ + i] = NAN;
        }
        if (N < 0) continue;
        Owned o;
        int rc = bh_basis_build(ctx, m, N, sp->order, &o.b);
        if (rc != BH_OK) return rc;
        rc = bh_hamiltonian_build(ctx, o.b, off.data(), sites.data(), sp->flags, &o.h);
        if (rc != BH_OK) return rc;
        int W = sp->width;
        if (W == 0) {
            rc = sweep_width(ctx, o.h, opts->prec, nJ, &W);
            if (rc != BH_OK) return rc;
        }
        if (out->width) out->width[q] = W;
        std::vector<double> ev(static_cast<size_t>(nJ) * k), rs(ev.size()), ee(ev.size());
        std::vector<int> it(nJ), uf(nJ), st(nJ);
        for (int c0 = 0; c0 < nJ; c0 += W) {
            const int nb = nJ - c0 < W ? nJ - c0 : W;
            const size_t e0 = static_cast<size_t>(c0) * k;
            bh_lanczos_result r{ev.data() + e0, rs.data() + e0, ee.data() + e0, it.data() + c0, uf.data() + c0,
                                st.data() + c0};
            rc = bh_lanczos(ctx, o.h, nb, sp->J + c0, Uv.data() + c0, opts, &r);
            if (rc != BH_OK && rc != BH_ENOCONV) return rc;
        }
        for (int i = 0; i < nJ; ++i) {
            for (int l = 0; l < k; ++l) out->E[(q * nJ + i) * k + l] = ev[static_cast<size_t>(i) * k + l];
            out->gap_ratio[q * nJ + i] = gap_ratio(&ev[static_cast<size_t>(i) * k], k);
            if (out->status) out->status[q * nJ + i] = st[i];
            if (out->used_fp64) out->used_fp64[q * nJ + i] = uf[i];
        }
    }
    for (int i = 0; i < nJ; ++i) {
        const double e_m = out->E[(0 * nJ + i) * k], e_0 = out->E[(1 * nJ + i) * k], e_p = out->E[(2 * nJ + i) * k];
        const double mup = e_p - e_0, mum = std::isfinite(e_m) ? e_0 - e_m : NAN;
        out->mu_plus[i] = mup;
        out->mu_minus[i] = mum;
        for (int c = 0; c < sp->nmu; ++c) {
            const double mu = sp->mu[c];
            int best = -1;
            double om = INFINITY;
            for (int q = 0; q < 3; ++q) {
                const double e = out->E[(q * nJ + i) * k];
                if (!std::isfinite(e)) continue;
                const double w = e - mu * (sp->n - 1 + q);
                if (w < om) {
                    om = w;
                    best = q;
                }
            }
            out->density[i * sp->nmu + c] = best < 0 ? NAN : static_cast<double>(sp->n - 1 + best) / m;
            double kap = NAN;
            if (best == 1 && std::isfinite(mum)) {
                const double gap = mup - mum;
                kap = gap > 0 ? 1.0 / (m * gap) : INFINITY;
            }
            out->compressibility[i * sp->nmu + c] = kap;
        }
    }
    return BH_OK;
}
