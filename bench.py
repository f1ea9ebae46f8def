"""Benchmark of the exact-diagonalisation hot path on one MI355X.

Workload (the single large solve of the benchmark configurations): periodic chain, m = 12 sites,
n = 12 bosons (D = C(23, 11) = 1,352,078), H = -J sum (b_i^+ b_j + h.c.) + U/2 sum n(n-1) - mu N at
J/U = 0.1, U = 1, mu = 0.5, lowest k = 4 eigenvalues, seeded Park-Miller start vector.

A step is one batched mixed-precision eigensolve: B = 16 solves of the sector (J/U = 0.1 and 15
neighbouring sweep points 0.1 * (1 + 0.01 b)) sharing one CSR pass per Lanczos step, on the resident
CSR, including the FP64 finish (Ritz vectors, Rayleigh quotients with the exact matrix values,
residuals).  value = eigensolves per second.  The solver layout is the lexicographic basis order
(same spectrum as the reference order; the CSR is rebuilt in it, 3 ms).

Beside it: the plain FP64 GPU Lanczos (FP64 CSR and FP64 vectors, one solve at a time, its fastest
engine), the error of the mixed E0 against it, the dominant kernel's HBM roofline (HIP events on the
library stream over the timed region), and the oracle's FP64 CPU Lanczos on the host cores.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bose-hubbard-phase-transition_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

import bhx  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(m, n, J, U, k, gpu_steps, sample_steps=300):
    """The oracle's FP64 CPU Lanczos (full reorthogonalisation, OpenMP over this host's cores) on the
    same sector: the first `sample_steps` steps are timed and the full solve is extrapolated linearly
    to the step count of the FP64 GPU solve (linear is generous to the CPU: its per-step
    reorthogonalisation grows with the step index)."""
    import oracle as O

    off, sites = O.chain_neighbours(m, True)
    h = O.hamiltonian(m, n, off, sites, order=O.ORDER_ASCENDING)
    t0 = time.perf_counter()
    rc, ev, bd, it = O.lanczos(h, J, U, k, max_iter=sample_steps)
    dt = time.perf_counter() - t0
    per_solve = dt / it * gpu_steps
    return {"value": 1.0 / per_solve, "unit": "eigensolves/s", "cores": omp_threads(), "kind": "port",
            "sample": f"first {it} steps of the oracle's FP64 Lanczos (full reorthogonalisation, OpenMP CSR) on the "
                      f"m=12 n=12 sector took {dt:.2f} s; extrapolated linearly to the {gpu_steps} steps of the FP64 "
                      f"GPU solve"}


def omp_threads():
    """threads the oracle's OpenMP loops use on this host"""
    try:
        return int(os.environ.get("OMP_NUM_THREADS", ""))
    except ValueError:
        return os.cpu_count()


def mf_grid(n):
    """the mean-field grid of the reference CLI (src/analysis.cpp:60-79): mu over [0, 4) in steps 4/n, J over
    [0, 0.25) in steps 0.25/n, accumulated in double; mu outer"""
    def rng(lo, hi, d):
        out, x = [], lo
        while x < hi:
            out.append(x)
            x += d
        return out
    mus, Js = rng(0.0, 4.0, 4.0 / n), rng(0.0, 0.25, 0.25 / n)
    return np.repeat(mus, len(Js)), np.tile(Js, len(mus))


# the report's mean-field phase diagram: 10^6 points at precision 1e-5 in 20 min 12 s on the authors' PC
MF_REFERENCE_S = 20 * 60 + 12


def mean_field_leg(ctx, cpu_sample, n=1000, precision=5):
    """The reference CLI's mean-field phase diagram (-t mean -i 1000 -e 5): 10^6 self-consistent solves in one
    launch; start values uniform in [0, 1) (seeded).  cpu_sample > 0 also times the oracle's scalar restatement
    of the same solves on that many grid points (one host thread)."""
    mu, J = mf_grid(n)
    psi0 = np.random.default_rng(0).uniform(0.0, 1.0, mu.size)
    eps = 10.0 ** -precision
    ctx.mean_field(mu[:4096], J[:4096], psi0[:4096], eps=eps)
    ctx.set_timing(True)
    ctx.kernel_stats("scmf", reset=True)
    t0 = time.perf_counter()
    r = ctx.mean_field(mu, J, psi0, eps=eps)
    wall = time.perf_counter() - t0
    ctx.set_timing(False)
    cnt, kms = ctx.kernel_stats("scmf")
    out = {"points": int(mu.size), "precision": precision, "kernel_ms": kms, "wall_s_with_copies": wall,
           "points_per_s": mu.size / wall, "converged": int(np.sum(r["status"] == 0)),
           "mean_iters": float(r["iters"].mean()), "max_iters": int(r["iters"].max()), "max_p": int(r["p"].max()),
           "reference_published_s": MF_REFERENCE_S, "speedup_vs_published": MF_REFERENCE_S / wall}
    if cpu_sample > 0:
        import oracle as O

        idx = np.random.default_rng(1).choice(mu.size, cpu_sample, replace=False)
        t0 = time.perf_counter()
        for k in idx:
            O.scmf(mu[k], J[k], 1, psi0[k], eps)
        dt = time.perf_counter() - t0
        out["cpu_oracle"] = {"points_per_s": cpu_sample / dt, "cores": 1, "kind": "port",
                             "sample": f"{cpu_sample} random points of the grid in {dt:.2f} s, one thread"}
    return out


PROFILE = os.path.join(ROOT, "profiles", "k1_lanczos_mixed_b16_chain12.txt")


def traffic_from_profile():
    """HBM bytes per K1 launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE passes on this workload (the
    committed summary holds the command, the copy-kernel calibration and the figure); None without it."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def serving_sectors(dmax=5_000_000):
    """The admissible sectors of the serving mix: chain m in [6, 14] periodic, square 3x3, 3x4, 4x4 periodic, with
    n in [m-2, m+2] and dimension at most 5e6 (chain 14 and 4x4 have none)."""
    from math import comb

    out = []
    geoms = [("chain", m, 0, 0) for m in range(6, 15)] + [("square", 0, 3, 3), ("square", 0, 3, 4), ("square", 0, 4, 4)]
    for g, m, Lx, Ly in geoms:
        mm = m if g == "chain" else Lx * Ly
        for n in range(mm - 2, mm + 3):
            if n >= 1 and comb(mm + n - 1, n) <= dmax:
                out.append((g, m, Lx, Ly, n))
    return out


def serving(ctx, nreq, rate, seed=7):
    """A Poisson stream of `nreq` requests at `rate` req/s offered load: each a seed-drawn admissible sector,
    J/U log-uniform in [0.01, 1], mu/U uniform in [0, 3], k = 4, mixed precision with FP64 rerun.  Returns
    p50/p99 latency (submit to result) and completed solves per second over the stream's span."""
    rng = np.random.default_rng(seed)
    secs = serving_sectors()
    with bhx.Server(ctx, lanes=16, k=4, ncv=160) as srv:
        t0 = time.perf_counter()
        tnext = t0
        ids = []
        for _ in range(nreq):
            g, m, Lx, Ly, n = secs[rng.integers(len(secs))]
            J = float(10 ** rng.uniform(-2, 0))
            mu = float(rng.uniform(0, 3))
            tnext += rng.exponential(1.0 / rate)
            dt = tnext - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            ids.append(srv.submit(g, n, J, 1.0, mu, m=m, Lx=Lx, Ly=Ly))
        ok = srv.drain(600.0)
        span = time.perf_counter() - t0
        st = srv.stats()
        res = [srv.poll(i) for i in ids]
    bad = sum(1 for r in res if r is None or r.status != 0)
    reruns = sum(1 for r in res if r is not None and r.used_fp64)
    lat = np.array([r.t_done - r.t_submit for r in res if r is not None])
    return {"requests": nreq, "offered_rps": rate, "completed": int(st.completed), "failed": bad, "drained": bool(ok),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "solves_per_s": float(st.completed / span), "fp64_results": reruns, "sectors": len(secs),
            "mix": "chain m=6..13, square 3x3 and 3x4, n in [m-2, m+2], D <= 5e6; J/U log-uniform [0.01, 1], "
                   "mu/U uniform [0, 3], k=4"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy kernel of BYTES bytes (calibrates HBM counters) and exit")
    ap.add_argument("--serve", type=int, default=120, metavar="N",
                    help="requests of the serving stream at 200 req/s offered load (0: skip)")
    ap.add_argument("--no-sweep", action="store_true", help="skip the phase-diagram sweep leg")
    ap.add_argument("--no-mean-field", action="store_true", help="skip the mean-field phase-diagram leg")
    ap.add_argument("--mean-field-only", action="store_true",
                    help="run only the mean-field leg (1000 x 1000 points, precision 1e-5) and print its JSON: for profilers")
    ap.add_argument("--quick", action="store_true",
                    help="time the mixed batch only (no plain FP64 comparator, SpMV table or CPU leg): for profilers")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported. Solves are independent: a "
              "multi-GPU run shards batches or requests across devices.", file=sys.stderr)
        sys.exit(2)

    if args.copy_calib:
        import torch  # the copy kernel of the calibration only

        a = torch.ones(args.copy_calib, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calib} bytes")
        return

    if args.mean_field_only:
        ctx = bhx.Context(0)
        print(json.dumps(mean_field_leg(ctx, 0 if args.no_cpu else 20000)))
        ctx.close()
        return

    m, n, J0, U, mu, k = 12, 12, 0.1, 1.0, 0.5, 4
    B = args.batch
    Js = J0 * (1.0 + 0.01 * np.arange(B))
    ctx = bhx.Context(0)
    log("device:", bhx.device_name(0))
    off, sites = bhx.lattice("chain", m=m, periodic=True)
    t0 = time.perf_counter()
    basis = ctx.basis(m, n, bhx.ORDER_LEX)
    ham = basis.hamiltonian(off, sites)
    build_s = time.perf_counter() - t0
    log(f"build D={ham.D} nnz={ham.nnz} {build_s * 1e3:.1f} ms")

    # plain FP64 GPU Lanczos: FP64 CSR and vectors, unbatched (accuracy comparator and baseline rate)
    fp64_s, r64 = float("nan"), None
    if not args.quick:
        r64 = ham.lanczos(J0, U, k=k, prec=bhx.PREC_FP64)
        reps64 = 3
        t0 = time.perf_counter()
        for _ in range(reps64):
            r64 = ham.lanczos(J0, U, k=k, prec=bhx.PREC_FP64)
        fp64_s = (time.perf_counter() - t0) / reps64
        log(f"plain fp64: {fp64_s * 1e3:.1f} ms/solve, steps={r64.iters[0]}, E={r64.evals[0]}")

    for _ in range(args.warmup):
        ham.lanczos(Js, U, k=k, prec=bhx.PREC_MIXED)
    # timed region: kernel timing off (the steps replay captured HIP graphs)
    t0 = time.perf_counter()
    res = None
    for _ in range(args.steps):
        res = ham.lanczos(Js, U, k=k, prec=bhx.PREC_MIXED)
    dt = time.perf_counter() - t0
    # one more batch with per-kernel HIP events on the library stream (launched eagerly) for the K1 and K2
    # averages; the kernels are the ones the graphs replay
    ctx.set_timing(True)
    ctx.kernel_stats("lanczos_spmv", reset=True)
    ham.lanczos(Js, U, k=k, prec=bhx.PREC_MIXED)
    ctx.set_timing(False)
    k1_count, k1_total = ctx.kernel_stats("lanczos_spmv")
    upd_count, upd_total = ctx.kernel_stats("lanczos_update")
    ms_step = dt / args.steps * 1e3
    value = B * args.steps / dt
    rel0 = abs(res.evals[0, 0] - r64.evals[0, 0]) / abs(r64.evals[0, 0]) if r64 is not None else float("nan")
    log(f"mixed batched B={B}: {ms_step:.1f} ms/step ({value:.1f} solves/s), steps={res.iters.tolist()}, "
        f"fp64 reruns={int(res.used_fp64.sum())}, |dE0|/|E0| vs fp64 = {rel0:.2e}")

    # dominant kernel: the fused Lanczos operator pass K1 (row-lane engine at B = 16).  Algorithmic bytes per
    # launch: packed matrix 4 B/entry + row pointers 4 B/row + diagonal 4 B/row, and per vector and row
    # v_j read (4 B), v_{j-1} read (4 B), w written in FP64 (8 B).
    D, nnz = ham.D, ham.nnz
    k1_bytes = 4.0 * nnz + 4.0 * (D + 1) + 4.0 * D + B * D * (4 + 4 + 8)
    k1_ms = k1_total / max(k1_count, 1)
    achieved = k1_bytes / (k1_ms * 1e-3) / 1e9
    k1_tuned = ham.tuned(0, bhx.PREC_MIXED, B)
    log(f"K1: {k1_count} launches, avg {k1_ms * 1e3:.1f} us, {k1_bytes / 1e6:.0f} MB -> {achieved:.0f} GB/s; "
        f"K2 avg {upd_total / max(upd_count, 1) * 1e3:.1f} us; autotuner {k1_tuned}")

    spmv = {}
    for prec, name in () if args.quick else ((bhx.PREC_MIXED, "mixed"), (bhx.PREC_FP64, "fp64")):
        for Bs in (1, B):
            ms, by = ham.spmv_bench(prec, Bs, 30)
            spmv[f"{name}_B{Bs}"] = {"us": ms * 1e3, "alg_GBs": by / ms / 1e6, "tuned": ham.tuned(1, prec, Bs)}
    log("spmv:", json.dumps(spmv))
    # SpMV at the 4x4 periodic square lattice, n=10 (D=3,268,760): its vectors and matrix exceed the 256 MB
    # Infinity Cache, so this is the HBM-roofline SpMV figure.  Besides this build's own algorithmic bytes,
    # the fraction is also given against the bytes of an FP32-valued CSR (the unit of the workload
    # definition: nnz (4 + 4) + 4 (D + 1) + 4 D + 2 * 4 D).
    spmv3 = {}
    if not args.quick:
        ham.close()
        basis.close()
        off3, sites3 = bhx.lattice("square", Lx=4, Ly=4, periodic=True)
        b3 = ctx.basis(16, 10, bhx.ORDER_LEX)
        h3 = b3.hamiltonian(off3, sites3)
        csr32 = 8.0 * h3.nnz + 4.0 * (h3.D + 1) + 4.0 * h3.D + 8.0 * h3.D
        csr64 = 12.0 * h3.nnz + 4.0 * (h3.D + 1) + 8.0 * h3.D + 16.0 * h3.D
        for prec, name, ref_bytes in ((bhx.PREC_MIXED, "mixed", csr32), (bhx.PREC_FP64, "fp64", csr64)):
            ms, by = h3.spmv_bench(prec, 1, 30)
            spmv3[f"{name}_B1"] = {"us": ms * 1e3, "alg_GBs": by / ms / 1e6, "frac_of_peak": by / ms / 1e6 / HBM_PEAK_GBS,
                                   "csr_equiv_GBs": ref_bytes / ms / 1e6,
                                   "csr_equiv_frac": ref_bytes / ms / 1e6 / HBM_PEAK_GBS,
                                   "tuned": h3.tuned(1, prec, 1)}
        spmv3["D"], spmv3["nnz"] = h3.D, h3.nnz
        log("spmv 4x4 n=10:", json.dumps(spmv3))
        # the 4x4 workload itself: J/U = 0.05 and 0.2 solved together in the FP32-mixed form (k = 4); its
        # operator pass K1 at width 2 is this sector's SpMV inside the solver
        J3 = np.array([0.05, 0.2])
        h3.lanczos(J3, 1.0, k=4, prec=bhx.PREC_MIXED)
        ctx.set_timing(True)
        ctx.kernel_stats("lanczos_spmv", reset=True)
        t3 = time.perf_counter()
        r3 = h3.lanczos(J3, 1.0, k=4, prec=bhx.PREC_MIXED)
        t3 = time.perf_counter() - t3
        ctx.set_timing(False)
        c3, tot3 = ctx.kernel_stats("lanczos_spmv")
        k1b3 = 4.0 * h3.nnz + 4.0 * (h3.D + 1) + 4.0 * h3.D + 2 * h3.D * (4 + 4 + 8)
        k1us3 = tot3 / max(c3, 1) * 1e3
        spmv3["lanczos_mixed_B2"] = {"s": t3, "iters": r3.iters.tolist(), "E0": r3.evals[:, 0].tolist(),
                                     "fp64_reruns": int(r3.used_fp64.sum()), "k1_us": k1us3,
                                     "k1_alg_GBs": k1b3 / k1us3 / 1e3, "k1_frac_of_peak": k1b3 / k1us3 / 1e3 / HBM_PEAK_GBS,
                                     "tuned": h3.tuned(0, bhx.PREC_MIXED, 2)}
        log("lanczos 4x4 n=10:", json.dumps(spmv3["lanczos_mixed_B2"]))
        h3.close()
        b3.close()

    # batched phase-diagram sweep: chain m=10 periodic, sectors n = 9, 10, 11, 32 log-spaced J/U in
    # [0.01, 0.4], 64 mu/U in [0, 3], k = 4; block width chosen per sector by the tuner
    sweep = None
    if not args.quick and not args.no_sweep:
        Jsw = 0.01 * 40.0 ** (np.arange(32) / 31.0)
        musw = 3.0 * np.arange(64) / 63.0
        ctx.sweep("chain", 10, Jsw, 1.0, musw, m=10, k=4)
        t0 = time.perf_counter()
        rsw = ctx.sweep("chain", 10, Jsw, 1.0, musw, m=10, k=4)
        tsw = time.perf_counter() - t0
        sweep = {"s": tsw, "solves": 3 * Jsw.size, "solves_per_s": 3 * Jsw.size / tsw, "width": rsw["width"].tolist(),
                 "fp64_reruns": int(rsw["used_fp64"].sum()), "all_ok": bool(np.all(rsw["status"] == 0))}
        log("sweep chain m=10 n=9/10/11:", json.dumps(sweep))

    mf = None
    if not args.quick and not args.no_mean_field:
        mf = mean_field_leg(ctx, 0 if args.no_cpu else 20000)
        log("mean field 1000x1000:", json.dumps(mf))

    serve = None
    if args.serve > 0 and not args.quick:
        serve = serving(ctx, args.serve, 200.0)
        log("serving:", json.dumps(serve))

    cpu = None
    if not args.no_cpu and not args.quick:
        cpu = cpu_baseline(m, n, J0, U, k, int(r64.iters[0]))
        log("cpu baseline:", json.dumps(cpu))

    out = {
        "metric": "eigensolves/sec mixed-batched vs plain FP64 @|dE0|/|E0|<=1e-9",
        "value": value,
        "unit": "eigensolves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-stored/f64-reduced",
        "data": "synthetic: exact Hamiltonian of the sector, seeded Park-Miller start vector",
        "config": {"workload": "chain m=12 n=12 periodic, J/U=0.1..0.115 (16 points), U=1, mu=0.5, k=4",
                   "batch": B, "D": D, "nnz": nnz, "layout": "lexicographic basis order"},
        "plain_fp64_eigensolves_per_s": 1.0 / fp64_s,
        "speedup_vs_plain_fp64": value * fp64_s,
        "rel_err_E0_vs_fp64": rel0,
        "spmv": spmv,
        "spmv_square4x4_n10": spmv3 or None,
        "sweep_chain10": sweep,
        "mean_field_1000x1000": mf,
        "serving": serve,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic_from_profile()},
        "k1": {"kernel": "k_lz_a_rl (fused Lanczos operator pass)", "launches": k1_count, "avg_us": k1_ms * 1e3,
               "alg_bytes_per_launch": k1_bytes, "k2_avg_us": upd_total / max(upd_count, 1) * 1e3,
               "tuned": k1_tuned},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    ctx.close()


if __name__ == "__main__":
    main()
