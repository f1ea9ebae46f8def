// bh_engine.hpp -- the Lanczos engine shared by the batch solver, the sweep driver and the server.
#pragma once

#include <memory>
#include <vector>

#include "bh_internal.hpp"

struct LaneResult {
    std::vector<double> evals, resid, errest;  // k entries (NaN when missing)
    int iters = 0;
    int status = BH_ENOCONV;
    std::vector<double> vecs;  // k * D (when requested)
};

// host-side time profile of an engine (BHX_LZ_PROFILE=1 prints it per batch solve)
struct LzProfile {
    int checks = 0;
    double wait_ms = 0.0;    // host waiting for the device at the analyses
    double host_ms = 0.0;    // host side of the analyses: decisions, locks, restarts, retirements
    double analysis_ms = 0.0;  // of which the tridiagonal analyses of the lanes
    double deflate_ms = 0.0;   // of which the deflations (Ritz vectors frozen, fresh starts)
    double lock_ms = 0.0;      // of which the locks (locked Ritz vectors formed)
    double finish_ms = 0.0;  // FP64 finish, wall time
    int refine_steps = 0;    // Davidson steps of the FP64 refinement run in the finishes (one per batch of lanes)
    int refined_lanes = 0;   // lane-steps of the refinement
    double run_ms = 0.0;     // enqueueing steps, graph captures included
    double capture_ms = 0.0; // capturing and instantiating step graphs
    double setup_ms = 0.0;   // engine construction and admission
};

// B lanes over one Hamiltonian; each lane runs one solve (J_b, U_b) with its own step counter, so lanes
// are admitted and retired independently.  All work goes to the engine's stream.
class LzEngineBase {
   public:
    LzProfile prof;
    virtual ~LzEngineBase() = default;
    virtual int admit(int slot, double J, double U) = 0;
    virtual int run(int steps) = 0;                     // enqueue steps for the active lanes
    virtual int check(std::vector<int>& finished) = 0;  // wait, analyse, retire converged lanes
    virtual int finish(const std::vector<int>& lanes, std::vector<LaneResult>& out, bool want_vectors) = 0;
    virtual int lanes() const = 0;
    virtual int active() const = 0;
    virtual int steps_of(int slot) const = 0;
    virtual int suggest_chunk() const = 0;  // steps until the next useful analysis
    virtual size_t device_bytes() const = 0;
    virtual bool fp64() const = 0;  // Krylov vectors stored in FP64
    // lane 0's Lanczos coefficients after `steps` steps without analyses (alpha[t], t < steps; beta[t] couples
    // t - 1 and t, t = 1 .. steps) and, when basis is given, its normalised vectors v_t as [steps][D]
    virtual int export_tridiag(int steps, double* alpha, double* beta, double* basis) = 0;
    // a lane that completes a deflation pass waits until no lane iterates, then the waiting lanes start their
    // next passes together (off: each lane starts its next pass at once, as the server wants)
    virtual void set_deflation_barrier(bool on) = 0;
};

// sectors with D at most opts.fullre_d (0: 16384, < 0: none) run FP64 with full reorthogonalisation
bool bh_lz_fullre(int64_t D, const bh_lanczos_opts& o);
// use_ctx_ws: keep the Krylov basis in the context's workspaces (one engine at a time per context);
// otherwise the engine owns its buffers (several engines may coexist, e.g. in the server)
std::unique_ptr<LzEngineBase> bh_make_engine(bh_ctx* ctx, bh_ham* h, int prec, int B, const bh_lanczos_opts& o,
                                             hipStream_t s, bool use_ctx_ws, int* rc);
// acceptance of a mixed-precision solve after the FP64 finish (else it is rerun in FP64): converged,
// E0 within accept_rel (0: 1e-9) by its bound ||r||^2/gap, and the levels above it within a looser bound
bool bh_lane_accepted(int status, const double* evals, const double* errest, int k, double accept_rel);
// options of the FP64 rerun of a rejected mixed solve, and of the full-reorthogonalisation rerun of an
// FP64 rerun that is not certified either
bh_lanczos_opts bh_fallback_opts(const bh_lanczos_opts& o);
bh_lanczos_opts bh_fallback_full_opts(const bh_lanczos_opts& o);
// bytes an engine of B lanes allocates for itself (use_ctx_ws = false)
size_t bh_engine_bytes(int64_t D, int prec, int B, const bh_lanczos_opts& o);
