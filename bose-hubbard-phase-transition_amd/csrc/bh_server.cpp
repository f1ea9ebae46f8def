// bh_server.cpp -- continuous-batching request server over sector engines.
//
// No counterpart in the reference.  The reference's parameter grid (src/analysis.cpp:302) schedules
// independent solves dynamically over OpenMP threads, capped by memory (src/resource.cpp:77-82); here
// independent requests of uneven (lattice, m, n) share the device: one Lanczos engine per sector in
// flight (its lanes are the batch), each on its own stream, lanes refilled as solves finish, sectors
// admitted and evicted against an HBM budget.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <unordered_map>

#include "bh_engine.hpp"
#include "bh_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;
constexpr int64_t SRV_SMALL_D = 1 << 15;  // sectors driven to completion within a pass (see worker_loop)
constexpr double SRV_SMALL_MS = 5.0;
using SectorKey = std::tuple<int, int, int, int, int>;  // kind, Lx, Ly, periodic, n

struct Pending {
    uint64_t id;
    bh_request req;
    double t_submit;
    int prec;
};

struct Sector {
    SectorKey key;
    int m = 0;
    bh_basis* basis = nullptr;
    bh_ham* ham = nullptr;
    std::deque<Pending> queue[2];            // per precision
    std::unique_ptr<LzEngineBase> eng[2];
    hipStream_t stream[2] = {nullptr, nullptr};
    std::vector<Pending> lane_req[2];        // request held by each lane
    std::vector<char> lane_busy[2];
    std::vector<double> lane_start[2];       // admission time of each lane's request
    double last_use = 0.0;
    size_t ham_bytes = 0, eng_bytes[2] = {0, 0};
    bool never_fits = false;
    bool inflight[2] = {false, false};       // a chunk of the engine's steps is on its stream
};

}  // namespace

struct bh_server {
    bh_ctx* ctx = nullptr;  // the server's own context
    bh_server_opts opts{};
    Clock::time_point t0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    uint64_t next_id = 1;
    std::deque<Pending> inbox;
    std::unordered_map<uint64_t, bh_result> done;
    std::map<SectorKey, std::unique_ptr<Sector>> sectors;
    int64_t submitted = 0, completed = 0;
    std::vector<double> latencies;
    double first_submit = -1.0, last_done = 0.0;
    int64_t in_flight = 0;
    // read by bh_server_stats_get while the worker runs: kept as counters the worker updates, so that the
    // reader never walks the sector map (which the worker changes without the lock)
    std::atomic<int> engines{0};
    std::atomic<size_t> hbm_used{0};

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
};

namespace {

void complete(bh_server* S, const Pending& p, int status, const double* ev, const double* rs, int k, int iters,
              int used64, double t_start) {
    bh_result r{};
    r.id = p.id;
    r.status = status;
    r.k = k;
    for (int i = 0; i < 16; ++i) {
        r.evals[i] = i < k && ev ? ev[i] : NAN;
        r.resid[i] = i < k && rs ? rs[i] : NAN;
    }
    r.iters = iters;
    r.used_fp64 = used64;
    r.t_submit = p.t_submit;
    r.t_start = t_start;
    r.t_done = S->now();
    std::lock_guard<std::mutex> lk(S->mu);
    S->done[p.id] = r;
    S->completed += 1;
    S->in_flight -= 1;
    S->latencies.push_back(r.t_done - r.t_submit);
    S->last_done = r.t_done;
    S->cv_done.notify_all();
}

size_t server_used(bh_server* S) {
    size_t u = 0;
    for (auto& kv : S->sectors) {
        const Sector& sc = *kv.second;
        u += sc.ham_bytes + sc.eng_bytes[0] + sc.eng_bytes[1];
    }
    return u;
}

int lanes_for(bh_server* S, int64_t D, int prec) {
    int L = S->opts.lanes > 0 ? S->opts.lanes : 16;
    const size_t cap = S->opts.engine_bytes ? S->opts.engine_bytes : (static_cast<size_t>(16) << 30);
    while (L > 1 && bh_engine_bytes(D, prec, L, S->opts.lanczos) > cap) L >>= 1;
    return L;
}

// releases one sector engine: its stream is drained first (a check may have left speculative steps on it)
void release_engine(bh_server* S, Sector& sc, int p) {
    if (sc.stream[p]) (void)hipStreamSynchronize(sc.stream[p]);
    if (sc.eng[p]) S->engines -= 1;
    sc.eng[p].reset();
    sc.eng_bytes[p] = 0;
    sc.inflight[p] = false;
    if (sc.stream[p]) (void)hipStreamDestroy(sc.stream[p]);
    sc.stream[p] = nullptr;
}

// Frees memory until `need` more bytes fit: first idle engines (no busy lane) with nothing queued for them,
// any sector's, least recently used first (an engine is rebuilt on demand; without this step two sectors
// whose FP64 reruns wait for memory while each holds an idle mixed engine would wait for each other
// forever), then idle sectors (no queue, no busy lane) whole.  `keep` / `keep_p` name the engine being made.
void evict_for(bh_server* S, size_t need, Sector* keep, int keep_p = -1) {
    const size_t budget = S->opts.hbm_budget;
    while (server_used(S) + need > budget) {
        Sector* victim = nullptr;
        int vp = -1;
        for (auto& kv : S->sectors) {
            Sector* sc = kv.second.get();
            for (int p = 0; p < 2; ++p) {
                if ((sc == keep && p == keep_p) || !sc->eng[p] || !sc->queue[p].empty()) continue;
                bool busy = false;
                for (char c : sc->lane_busy[p]) busy |= c != 0;
                if (busy) continue;
                if (!victim || sc->last_use < victim->last_use) {
                    victim = sc;
                    vp = p;
                }
            }
        }
        if (!victim) break;
        release_engine(S, *victim, vp);
    }
    while (server_used(S) + need > budget) {
        Sector* victim = nullptr;
        for (auto& kv : S->sectors) {
            Sector* sc = kv.second.get();
            if (sc == keep || !sc->queue[0].empty() || !sc->queue[1].empty()) continue;
            bool busy = false;
            for (int p = 0; p < 2; ++p)
                for (char c : sc->lane_busy[p]) busy |= c != 0;
            if (busy) continue;
            if (!victim || sc->last_use < victim->last_use) victim = sc;
        }
        if (!victim) break;
        for (int p = 0; p < 2; ++p) release_engine(S, *victim, p);
        if (victim->ham) bh_ham_destroy(victim->ham);
        if (victim->basis) bh_basis_destroy(victim->basis);
        S->sectors.erase(victim->key);
    }
    // Still short and no lane anywhere busy (nothing would ever free memory): the other sectors' CSRs are
    // released too, their queues kept (a sector builds its CSR again when it is served), so that the sector
    // being served always progresses.
    if (server_used(S) + need <= budget) return;
    for (auto& kv : S->sectors)
        for (int p = 0; p < 2; ++p)
            for (char c : kv.second->lane_busy[p])
                if (c) return;
    for (auto& kv : S->sectors) {
        Sector* sc = kv.second.get();
        if (sc == keep || !sc->ham || server_used(S) + need <= budget) continue;
        for (int p = 0; p < 2; ++p) release_engine(S, *sc, p);
        bh_ham_destroy(sc->ham);
        sc->ham = nullptr;
        if (sc->basis) bh_basis_destroy(sc->basis);
        sc->basis = nullptr;
        sc->ham_bytes = 0;
    }
}

int ensure_ham(bh_server* S, Sector& sc) {
    if (sc.ham) return BH_OK;
    int kind, Lx, Ly, per, n;
    std::tie(kind, Lx, Ly, per, n) = sc.key;
    const int m = kind == 0 ? Lx : Lx * Ly;
    sc.m = m;
    const int64_t D = bh_binom_host(m + n - 1, n);
    if (D <= 0) return BH_ERANGE;
    // estimate of the CSR footprint before building: at most deg * m entries per row
    const size_t est = static_cast<size_t>(D) * (4 * (2 * m + 3));
    evict_for(S, est, &sc);
    std::vector<int> off(m + 1), sites(4 * m + 4);
    if (bh_lattice_neighbours(kind, Lx, Ly, per, off.data(), sites.data()) < 0) return BH_EINVAL;
    BH_TRY(bh_basis_build(S->ctx, m, n, BH_ORDER_LEX, &sc.basis));
    int rc = bh_hamiltonian_build(S->ctx, sc.basis, off.data(), sites.data(), S->opts.flags, &sc.ham);
    if (rc != BH_OK) return rc;
    sc.ham_bytes = static_cast<size_t>(bh_ham_device_bytes(sc.ham)) + sc.basis->bytes;
    return BH_OK;
}

// Lanes of a sector's engine: enough for the requests queued when it is built (next power of two), within
// the memory cap.  A lane of an engine costs about a full lane's bandwidth per step even when idle (a row's B
// lane values are one contiguous segment), so engines are sized to demand rather than to the maximum; an idle
// engine with more requests queued than lanes is rebuilt larger.
int lanes_wanted(bh_server* S, Sector& sc, int p) {
    const int prec = p == 0 ? BH_PREC_MIXED : BH_PREC_FP64;
    const int cap = p == 0 ? lanes_for(S, sc.ham->D, prec) : std::min(4, lanes_for(S, sc.ham->D, prec));
    int L = 1;
    while (L < cap && L < static_cast<int>(sc.queue[p].size())) L <<= 1;
    return std::min(L, cap);
}

int ensure_engine(bh_server* S, Sector& sc, int p) {
    if (sc.eng[p]) {
        // rebuild an idle engine that is too small for its queue
        if (sc.eng[p]->active() > 0 || static_cast<int>(sc.queue[p].size()) <= sc.eng[p]->lanes()) return BH_OK;
        if (lanes_wanted(S, sc, p) <= sc.eng[p]->lanes()) return BH_OK;
        (void)hipStreamSynchronize(sc.stream[p]);
        sc.eng[p].reset();
        S->engines -= 1;
        sc.eng_bytes[p] = 0;
    }
    BH_TRY(ensure_ham(S, sc));
    const int prec = p == 0 ? BH_PREC_MIXED : BH_PREC_FP64;
    // the FP64 engine serves the reruns of rejected mixed solves
    const bh_lanczos_opts lo = p == 0 ? S->opts.lanczos : bh_fallback_opts(S->opts.lanczos);
    const int L = lanes_wanted(S, sc, p);
    const size_t need = bh_engine_bytes(sc.ham->D, prec, L, lo) +
                        (prec == BH_PREC_FP64 ? static_cast<size_t>(sc.ham->nnz) * 12 : 0);
    evict_for(S, need, &sc, p);
    if (server_used(S) + need > S->opts.hbm_budget) {
        if (need + sc.ham_bytes > S->opts.hbm_budget) sc.never_fits = true;
        return BH_EBUDGET;  // wait for lanes elsewhere to drain
    }
    if (!sc.stream[p]) BH_HIP(S->ctx, hipStreamCreateWithFlags(&sc.stream[p], hipStreamNonBlocking));
    int rc = BH_OK;
    sc.eng[p] = bh_make_engine(S->ctx, sc.ham, prec, L, lo, sc.stream[p], false, &rc);
    if (!sc.eng[p]) return rc;
    S->engines += 1;
    sc.eng_bytes[p] = sc.eng[p]->device_bytes();
    sc.lane_req[p].assign(sc.eng[p]->lanes(), Pending{});
    sc.lane_busy[p].assign(sc.eng[p]->lanes(), 0);
    sc.lane_start[p].assign(sc.eng[p]->lanes(), 0.0);
    return BH_OK;
}

void serve_diagonal(bh_server* S, Sector& sc, const Pending& pd) {
    // J = 0: the levels are the distinct values of U V (bh_lanczos takes its exact diagonal path)
    const double t = S->now();
    const int k = S->opts.lanczos.k;
    double J = 0.0, U = pd.req.U;
    std::vector<double> ev(k), rs(k), ee(k);
    int it = 0, uf = 0, st = 0;
    bh_lanczos_result r{ev.data(), rs.data(), ee.data(), &it, &uf, &st};
    bh_lanczos_opts o = S->opts.lanczos;
    int rc = bh_lanczos(S->ctx, sc.ham, 1, &J, &U, &o, &r);
    for (int i = 0; i < k; ++i) ev[i] -= pd.req.mu * pd.req.n;
    complete(S, pd, rc == BH_OK ? st : rc, ev.data(), rs.data(), k, it, uf, t);
}

void serve_full(bh_server* S, Sector& sc, const Pending& pd, int prior_iters, double t_start) {
    // an FP64 rerun whose levels are not certified: FP64 with full reorthogonalisation, solved at once
    const int k = S->opts.lanczos.k;
    double J = pd.req.J, U = pd.req.U;
    std::vector<double> ev(k), rs(k), ee(k);
    int it = 0, uf = 0, st = 0;
    bh_lanczos_result r{ev.data(), rs.data(), ee.data(), &it, &uf, &st};
    bh_lanczos_opts o = bh_fallback_full_opts(S->opts.lanczos);
    int rc = bh_lanczos(S->ctx, sc.ham, 1, &J, &U, &o, &r);
    for (int i = 0; i < k; ++i) ev[i] -= pd.req.mu * pd.req.n;
    complete(S, pd, rc == BH_OK ? st : rc, ev.data(), rs.data(), k, prior_iters + it, 1, t_start);
}

// Fills the free lanes of a sector's engine for precision p from its queue (the engine is idle).
void admit_queued(bh_server* S, Sector& sc, int p) {
    const int k = S->opts.lanczos.k;
    while (!sc.queue[p].empty()) {
        Pending pd = sc.queue[p].front();
        int rc = ensure_ham(S, sc);
        if (rc == BH_OK && pd.req.J == 0.0) {
            sc.queue[p].pop_front();
            serve_diagonal(S, sc, pd);
            continue;
        }
        if (rc == BH_OK) rc = ensure_engine(S, sc, p);
        if (rc != BH_OK) {
            if (rc == BH_EBUDGET && !sc.never_fits) break;  // retry when memory frees up
            sc.queue[p].pop_front();
            complete(S, pd, rc, nullptr, nullptr, k, 0, p, S->now());
            continue;
        }
        int lane = -1;
        for (int l = 0; l < static_cast<int>(sc.lane_busy[p].size()); ++l)
            if (!sc.lane_busy[p][l]) {
                lane = l;
                break;
            }
        if (lane < 0) break;
        sc.queue[p].pop_front();
        if (sc.eng[p]->admit(lane, pd.req.J, pd.req.U) != BH_OK) {
            complete(S, pd, BH_EHIP, nullptr, nullptr, k, 0, p, S->now());
            continue;
        }
        sc.lane_busy[p][lane] = 1;
        sc.lane_req[p][lane] = pd;
        sc.lane_start[p][lane] = S->now();
        sc.last_use = S->now();
    }
}

// Analyses a sector engine whose chunk has completed: retires converged lanes, completes their requests or
// queues their reruns (FP64, then FP64 with full reorthogonalisation).
// An engine call failed (a device allocation of its analysis or finish over the budget, a HIP error): every
// request on its lanes completes with that status and the engine is released (its state is not trusted), so
// that the worker never spins on an engine that cannot advance.
void fail_engine(bh_server* S, Sector& sc, int p, int rc) {
    const int k = S->opts.lanczos.k;
    for (size_t l = 0; l < sc.lane_busy[p].size(); ++l) {
        if (!sc.lane_busy[p][l]) continue;
        sc.lane_busy[p][l] = 0;
        complete(S, sc.lane_req[p][l], rc, nullptr, nullptr, k, 0, p, sc.lane_start[p][l]);
    }
    release_engine(S, sc, p);
}

void harvest(bh_server* S, Sector& sc, int p) {
    const int k = S->opts.lanczos.k;
    std::vector<int> fin;
    int rc = sc.eng[p]->check(fin);
    if (rc != BH_OK) {
        fail_engine(S, sc, p, rc);
        return;
    }
    if (fin.empty()) return;
    std::vector<LaneResult> res;
    rc = sc.eng[p]->finish(fin, res, false);
    if (rc != BH_OK) {
        fail_engine(S, sc, p, rc);
        return;
    }
    for (size_t i = 0; i < fin.size(); ++i) {
        const int lane = fin[i];
        Pending pd = sc.lane_req[p][lane];
        sc.lane_busy[p][lane] = 0;
        const LaneResult& L = res[i];
        // full reorthogonalisation needs no certificate; the mixed engine and the FP64 engine with selective
        // orthogonalisation do
        const bool fullre_eng =
            sc.eng[p]->fp64() && bh_lz_fullre(sc.ham->D, p == 0 ? S->opts.lanczos : bh_fallback_opts(S->opts.lanczos));
        const bool ok = fullre_eng ||
                        bh_lane_accepted(L.status, L.evals.data(), L.errest.data(), k, S->opts.lanczos.accept_rel);
        if (p == 0 && !ok && S->opts.lanczos.fallback >= 0) {
            pd.prec = BH_PREC_FP64;
            sc.queue[1].push_back(pd);  // rerun in FP64
            continue;
        }
        if (p == 1 && !ok) {
            serve_full(S, sc, pd, L.iters, sc.lane_start[p][lane]);
            continue;
        }
        std::vector<double> ev(k, NAN), rs(k, NAN);
        for (int q = 0; q < k && q < static_cast<int>(L.evals.size()); ++q) {
            ev[q] = L.evals[q] - pd.req.mu * pd.req.n;
            rs[q] = L.resid[q];
        }
        complete(S, pd, L.status, ev.data(), rs.data(), k, L.iters, sc.eng[p]->fp64() ? 1 : 0, sc.lane_start[p][lane]);
    }
    sc.last_use = S->now();
}

// The engines run independently: each pass of the loop visits every sector engine; one whose chunk is still
// on its stream is skipped, one whose stream is idle is analysed (finished lanes retired), refilled from its
// queue and given its next chunk at once.  A small sector thus cycles through its steps while a large one's
// chunk is running, and no engine waits for another's.  Each engine still runs the same sequence of chunks
// and analyses, so a request's result does not depend on the other requests.
void worker_loop(bh_server* S) {
    (void)hipSetDevice(S->ctx->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> lk(S->mu);
            bool have_work = !S->inbox.empty();
            for (auto& kv : S->sectors)
                for (int p = 0; p < 2; ++p)
                    have_work |= !kv.second->queue[p].empty() || (kv.second->eng[p] && kv.second->eng[p]->active() > 0);
            if (!have_work) S->cv_work.wait(lk, [&] { return S->stop || !S->inbox.empty(); });
            if (S->stop) return;
            while (!S->inbox.empty()) {
                Pending pd = S->inbox.front();
                S->inbox.pop_front();
                const bh_request& q = pd.req;
                SectorKey key{q.kind, q.Lx, q.kind == 0 ? 1 : q.Ly, q.periodic, q.n};
                auto& sp = S->sectors[key];
                if (!sp) {
                    sp.reset(new Sector());
                    sp->key = key;
                }
                sp->queue[pd.prec == BH_PREC_FP64 ? 1 : 0].push_back(pd);
            }
        }
        bool progressed = false;
        for (auto it = S->sectors.begin(); it != S->sectors.end(); ++it) {
            Sector& sc = *it->second;
            for (int p = 0; p < 2; ++p) {
                // a small sector's chunks take well under a millisecond: it is driven chunk after chunk for up
                // to SRV_SMALL_MS instead of one chunk per pass, so that it does not wait for the host work of
                // every other sector (CSR builds, tuning, finishes) between its chunks
                const bool small = sc.ham && sc.ham->D <= SRV_SMALL_D;
                const auto t0 = Clock::now();
                for (;;) {
                    if (sc.inflight[p]) {
                        if (small)
                            (void)hipStreamSynchronize(sc.stream[p]);
                        else if (hipStreamQuery(sc.stream[p]) == hipErrorNotReady)
                            break;
                        sc.inflight[p] = false;
                        progressed = true;
                        harvest(S, sc, p);
                    }
                    admit_queued(S, sc, p);
                    if (!(sc.eng[p] && sc.eng[p]->active() > 0)) break;
                    const int rc = sc.eng[p]->run(sc.eng[p]->suggest_chunk());
                    if (rc != BH_OK) {
                        fail_engine(S, sc, p, rc);
                        progressed = true;
                        break;
                    }
                    sc.inflight[p] = true;
                    progressed = true;
                    if (!small || std::chrono::duration<double, std::milli>(Clock::now() - t0).count() > SRV_SMALL_MS) break;
                }
            }
        }
        S->hbm_used = S->ctx->used;
        if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

}  // namespace

extern "C" {

int bh_server_create(bh_ctx* ctx, const bh_server_opts* opts, bh_server** out) {
    if (!ctx || !opts || !out) return BH_EINVAL;
    *out = nullptr;
    if (opts->lanczos.k < 1 || opts->lanczos.k > 16) return bh_set_error(ctx, BH_EINVAL, "server: 1 <= k <= 16");
    if (opts->lanes < 0 || opts->lanes > 64) return bh_set_error(ctx, BH_EINVAL, "server: lanes <= 64");
    auto* S = new bh_server();
    S->opts = *opts;
    size_t freeb = 0, totb = 0;
    if (hipSetDevice(ctx->device) != hipSuccess || hipMemGetInfo(&freeb, &totb) != hipSuccess) {
        delete S;
        return BH_EHIP;
    }
    if (!S->opts.hbm_budget) S->opts.hbm_budget = static_cast<size_t>(0.8 * static_cast<double>(freeb));
    // The budget admits CSRs and engines; the context also holds the transient buffers of an engine's
    // analysis, deflations and finish (one at a time: the worker is one thread), up to ~2 (2k + 8) vectors of
    // the largest engine -- a few GB at the engine cap -- so it gets that much on top of the budget.
    const size_t slack = std::min<size_t>(static_cast<size_t>(16) << 30,
                                          std::max<size_t>(static_cast<size_t>(1) << 30, freeb / 10));
    int rc = bh_ctx_create(ctx->device, S->opts.hbm_budget + slack, &S->ctx);
    if (rc != BH_OK) {
        delete S;
        return rc;
    }
    S->t0 = Clock::now();
    S->worker = std::thread(worker_loop, S);
    *out = S;
    return BH_OK;
}

int bh_server_submit(bh_server* S, const bh_request* req, uint64_t* id) {
    if (!S || !req) return BH_EINVAL;
    const int m = req->kind == 0 ? req->Lx : req->Lx * req->Ly;
    if (m < 1 || m > BH_MAX_SITES || req->n < 0 || req->n > 255 || (req->kind != 0 && req->kind != 1)) return BH_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    Pending pd;
    pd.id = S->next_id++;
    pd.req = *req;
    pd.t_submit = S->now();
    pd.prec = S->opts.prec;
    if (S->first_submit < 0) S->first_submit = pd.t_submit;
    S->inbox.push_back(pd);
    S->submitted += 1;
    S->in_flight += 1;
    if (id) *id = pd.id;
    S->cv_work.notify_one();
    return BH_OK;
}

int bh_server_poll(bh_server* S, uint64_t id, bh_result* out) {
    if (!S) return BH_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    auto it = S->done.find(id);
    if (it == S->done.end()) return BH_EAGAIN;
    if (out) *out = it->second;
    S->done.erase(it);
    return BH_OK;
}

int bh_server_wait(bh_server* S, uint64_t id, double timeout_s, bh_result* out) {
    if (!S) return BH_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    const bool ok = S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return S->done.count(id) > 0; });
    if (!ok) return BH_EAGAIN;
    if (out) *out = S->done[id];
    S->done.erase(id);
    return BH_OK;
}

int bh_server_drain(bh_server* S, double timeout_s) {
    if (!S) return BH_EINVAL;
    std::unique_lock<std::mutex> lk(S->mu);
    const bool ok = S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), [&] { return S->in_flight == 0; });
    return ok ? BH_OK : BH_EAGAIN;
}

int bh_server_stats_get(bh_server* S, bh_server_stats* st) {
    if (!S || !st) return BH_EINVAL;
    std::lock_guard<std::mutex> lk(S->mu);
    std::memset(st, 0, sizeof(*st));
    st->submitted = S->submitted;
    st->completed = S->completed;
    st->in_flight = S->in_flight;
    st->queued = static_cast<int64_t>(S->inbox.size());
    st->hbm_used = S->hbm_used;
    st->engines = S->engines;
    if (!S->latencies.empty()) {
        std::vector<double> l(S->latencies);
        std::sort(l.begin(), l.end());
        auto pct = [&](double q) { return l[std::min(l.size() - 1, static_cast<size_t>(q * (l.size() - 1) + 0.5))]; };
        st->p50_latency = pct(0.50);
        st->p99_latency = pct(0.99);
        double s = 0;
        for (double v : l) s += v;
        st->mean_latency = s / l.size();
        st->busy_seconds = S->last_done - S->first_submit;
    }
    return BH_OK;
}

int bh_server_destroy(bh_server* S) {
    if (!S) return BH_EINVAL;
    {
        std::lock_guard<std::mutex> lk(S->mu);
        S->stop = true;
        S->cv_work.notify_all();
    }
    if (S->worker.joinable()) S->worker.join();
    for (auto& kv : S->sectors) {
        Sector& sc = *kv.second;
        for (int p = 0; p < 2; ++p) {
            if (sc.stream[p]) (void)hipStreamSynchronize(sc.stream[p]);
            sc.eng[p].reset();
            if (sc.stream[p]) (void)hipStreamDestroy(sc.stream[p]);
        }
        if (sc.ham) bh_ham_destroy(sc.ham);
        if (sc.basis) bh_basis_destroy(sc.basis);
    }
    S->sectors.clear();
    bh_ctx_destroy(S->ctx);
    delete S;
    return BH_OK;
}

}  // extern "C"
