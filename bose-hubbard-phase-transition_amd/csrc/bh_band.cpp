// bh_band.cpp -- caller-chosen basis orders, and the band order: a permutation of a sector's basis that keeps
// every hop of a lattice's Hamiltonian close to the diagonal.
//
// The reference fixes one order (BH::fixed_set_basis, src/hamiltonian.cpp:60-122); the eigenvalues do not
// depend on it.  The operator pass reads x at the columns of each row, and which of those reads an XCD's L2
// holds depends on how far the columns of nearby rows spread.  In the lexicographic order a hop on the first
// sites moves the rank by a large fraction of D (chain m=12 n=12: half of the entries lie more than 262,144
// rows off the diagonal); after reverse Cuthill-McKee every entry lies within 65,536 rows, and a window of
// 16,384 consecutive rows gathers from 2.5 MiB of a 16-lane FP32 vector instead of 10.6 MiB.
#include <algorithm>
#include <cstring>
#include <vector>

#include "bh_internal.hpp"

namespace {

// Breadth-first levels from `start` over the vertices not yet placed (placed[v] == 0); returns the last
// level's vertices and the number of levels.
int bfs_levels(const std::vector<int32_t>& rp, const std::vector<int32_t>& col, int32_t start,
               const std::vector<uint8_t>& placed, std::vector<int32_t>& dist, std::vector<int32_t>& last) {
    std::vector<int32_t> q;
    q.push_back(start);
    dist[start] = 0;
    size_t h = 0;
    int depth = 0;
    std::vector<int32_t> touched{start};
    while (h < q.size()) {
        const int32_t v = q[h++];
        depth = std::max(depth, dist[v]);
        for (int32_t a = rp[v]; a < rp[v + 1]; ++a) {
            const int32_t u = col[a];
            if (placed[u] || dist[u] >= 0) continue;
            dist[u] = dist[v] + 1;
            q.push_back(u);
            touched.push_back(u);
        }
    }
    last.clear();
    for (int32_t v : q)
        if (dist[v] == depth) last.push_back(v);
    for (int32_t v : touched) dist[v] = -1;
    return depth + 1;
}

// Reverse Cuthill-McKee order of the symmetric graph (rp, col) on D vertices: order[k] = vertex placed at
// position k.  Each connected component starts from a pseudo-peripheral vertex (George and Liu: repeat a
// breadth-first search from a smallest-degree vertex of the previous search's last level while the number
// of levels grows); the Cuthill-McKee sweep appends each vertex's unplaced neighbours by increasing degree
// (ties by index); the whole order is then reversed.  Deterministic.
void rcm_order(int64_t D, const std::vector<int32_t>& rp, const std::vector<int32_t>& col, std::vector<int32_t>& order) {
    std::vector<uint8_t> placed(static_cast<size_t>(D), 0);
    std::vector<int32_t> dist(static_cast<size_t>(D), -1), last;
    auto deg = [&](int32_t v) { return rp[v + 1] - rp[v]; };
    order.clear();
    order.reserve(static_cast<size_t>(D));
    std::vector<std::pair<int32_t, int32_t>> nb;
    for (int64_t seed = 0; seed < D; ++seed) {
        if (placed[seed]) continue;
        int32_t start = static_cast<int32_t>(seed);
        int levels = bfs_levels(rp, col, start, placed, dist, last);
        for (int it = 0; it < 8; ++it) {
            int32_t cand = last[0];
            for (int32_t v : last)
                if (deg(v) < deg(cand) || (deg(v) == deg(cand) && v < cand)) cand = v;
            const int l2 = bfs_levels(rp, col, cand, placed, dist, last);
            if (l2 <= levels) break;
            levels = l2;
            start = cand;
        }
        size_t h = order.size();
        order.push_back(start);
        placed[start] = 1;
        while (h < order.size()) {
            const int32_t v = order[h++];
            nb.clear();
            for (int32_t a = rp[v]; a < rp[v + 1]; ++a) {
                const int32_t u = col[a];
                if (!placed[u]) {
                    placed[u] = 1;
                    nb.push_back({deg(u), u});
                }
            }
            std::sort(nb.begin(), nb.end());
            for (const auto& p : nb) order.push_back(p.second);
        }
    }
    std::reverse(order.begin(), order.end());
}

}  // namespace

extern "C" {

int bh_basis_permute(bh_ctx* ctx, bh_basis* b, const int32_t* lex_of_pos) {
    if (!ctx || !b || !lex_of_pos || b->ctx != ctx) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int64_t D = b->D;
    std::vector<int32_t> pol(static_cast<size_t>(D), -1);
    for (int64_t k = 0; k < D; ++k) {
        const int32_t l = lex_of_pos[k];
        if (l < 0 || l >= D || pol[l] >= 0) return bh_set_error(ctx, BH_EINVAL, "basis_permute: lex_of_pos is not a permutation");
        pol[l] = static_cast<int32_t>(k);
    }
    BH_HIP(ctx, hipSetDevice(ctx->device));
    BH_HIP(ctx, hipMemcpyAsync(b->lex_of_pos, lex_of_pos, static_cast<size_t>(D) * sizeof(int32_t), hipMemcpyHostToDevice,
                               ctx->stream));
    BH_HIP(ctx, hipMemcpyAsync(b->pos_of_lex, pol.data(), static_cast<size_t>(D) * sizeof(int32_t), hipMemcpyHostToDevice,
                               ctx->stream));
    BH_HIP(ctx, hipStreamSynchronize(ctx->stream));
    b->order = BH_ORDER_CUSTOM;
    b->layout_id = bh_next_layout_id();
    return BH_OK;
}

int bh_basis_reorder_band(bh_ctx* ctx, bh_basis* b, const int* nbr_off, const int* nbr_sites) {
    if (!ctx || !b || !nbr_off || !nbr_sites || b->ctx != ctx) return BH_EINVAL;
    std::lock_guard<std::recursive_mutex> lk(ctx->mu);
    const int64_t D = b->D;
    // the hopping graph in the current order: the packed words of the Hamiltonian (column << code_bits | code)
    std::vector<int32_t> rp(static_cast<size_t>(D) + 1), lop(static_cast<size_t>(D));
    std::vector<uint32_t> words;
    int cb = 0;
    {
        bh_ham* h = nullptr;
        BH_TRY(bh_hamiltonian_build(ctx, b, nbr_off, nbr_sites, 0u, &h));
        words.resize(static_cast<size_t>(std::max<int64_t>(h->nnz, 1)));
        cb = h->code_bits;
        int rc = BH_OK;
        hipStream_t s = ctx->stream;
        if (hipMemcpyAsync(rp.data(), h->row_ptr, rp.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            (h->nnz > 0 && hipMemcpyAsync(words.data(), h->words, static_cast<size_t>(h->nnz) * sizeof(uint32_t),
                                          hipMemcpyDeviceToHost, s) != hipSuccess) ||
            hipMemcpyAsync(lop.data(), b->lex_of_pos, lop.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            rc = bh_set_error(ctx, BH_EHIP, "basis_reorder_band: copy of the hopping graph");
        bh_ham_destroy(h);
        BH_TRY(rc);
    }
    std::vector<int32_t> col(words.size());
    for (size_t a = 0; a < words.size(); ++a) col[a] = static_cast<int32_t>(words[a] >> cb);
    words.clear();
    words.shrink_to_fit();
    std::vector<int32_t> order;
    rcm_order(D, rp, col, order);
    std::vector<int32_t> nl(static_cast<size_t>(D));
    for (int64_t k = 0; k < D; ++k) nl[k] = lop[order[k]];
    BH_TRY(bh_basis_permute(ctx, b, nl.data()));
    b->order = BH_ORDER_BAND;
    return BH_OK;
}

}  // extern "C"
