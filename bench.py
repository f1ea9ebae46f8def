"""Benchmark of the exact-diagonalisation hot path on one MI355X.

Workload (the single large solve of the benchmark configurations): periodic chain, m = 12 sites,
n = 12 bosons (D = C(23, 11) = 1,352,078), H = -J sum (b_i^+ b_j + h.c.) + U/2 sum n(n-1) - mu N at
J/U = 0.1, U = 1, mu = 0.5, lowest k = 4 eigenvalues counted with their multiplicity (as the reference's
Spectra solve returns them), seeded Park-Miller start vector.

A step is one batched mixed-precision eigensolve: B = 16 solves of the sector (J/U = 0.1 and 15
neighbouring sweep points 0.1 * (1 + 0.01 b)) sharing one CSR pass per Lanczos step, on the resident
CSR, including the FP64 finish (Ritz vectors, Rayleigh quotients with the exact matrix values,
residuals).  value = eigensolves per second.  The solver layout is the lexicographic basis order
(same spectrum as the reference order; the CSR is rebuilt in it, 3 ms).

Beside it: the plain FP64 GPU Lanczos (FP64 vectors, exact FP64 matrix values, one solve at a time, its fastest
engine), the error of every mixed value against the FP64 solve of its J point, the dominant kernel's HBM
roofline (HIP events on the library stream over the timed region), and the reference's own CPU solver
(Op::IRLM_eigen = Spectra GenEigs, compiled from the reference's sources into oracle/_ref) on the host
cores, one solve per OpenMP thread as the reference's grid loop runs them.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "bose-hubbard-phase-transition_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import numpy as np  # noqa: E402

if __name__ == "__main__" and len(sys.argv) > 2 and sys.argv[1] == "--cpu-child":
    bhx = None  # the CPU child of a baseline leg never touches the GPU library
else:
    import bhx  # noqa: E402

HBM_PEAK_GBS = 8000.0


def bhx_prec(name):
    return getattr(bhx, "PREC_" + name) if bhx is not None else None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    """CPU model and logical CPU count of this host (lscpu / os.cpu_count)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count()


def run_cpu_child(mode, args, cap_s, threads, setup_cap_s=600.0):
    """Runs `bench.py --cpu-child MODE ARGS` (the reference's CPU solver, no GPU) with OMP_NUM_THREADS = threads;
    the child's setup (building the sector on the host) is not timed, then its solves get at most cap_s seconds
    of wall time.  Returns (records of the solves finished within the cap, elapsed seconds)."""
    import selectors
    import subprocess

    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    cmd = [sys.executable, "-u", os.path.abspath(__file__), "--cpu-child", mode] + [str(a) for a in args]
    t_launch = time.perf_counter()
    proc = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True, env=env)
    sel = selectors.DefaultSelector()
    sel.register(proc.stdout, selectors.EVENT_READ)
    recs, start = [], None
    while True:
        deadline = start + cap_s if start is not None else t_launch + setup_cap_s
        left = deadline - time.perf_counter()
        if left <= 0:
            break
        if not sel.select(timeout=min(left, 1.0)):
            if proc.poll() is not None:
                break
            continue
        line = proc.stdout.readline()
        if not line:
            break
        if line.startswith("point "):
            _, p, rc, sec = line.split()
            rec = {"point": int(p), "rc": int(rc), "s": float(sec)}
        else:
            rec = json.loads(line)
        if rec.get("event") == "start":
            start = time.perf_counter()
        else:
            recs.append(rec)
    elapsed = min(time.perf_counter(), start + cap_s) - start if start is not None else 0.0
    if proc.poll() is None:
        proc.kill()
    proc.wait()
    return recs, elapsed


def cpu_child(mode, argv):
    """Child side of run_cpu_child: builds the sector(s) with the oracle (host CSR in the reference's basis
    order), then runs the reference's own solver, printing one JSON record per finished solve."""
    import oracle as O

    if mode == "sweep":
        # the reference's exact sweep on one sector: Op::IRLM_eigen per J point, one point per OpenMP thread
        # (src/analysis.cpp:302-314), H assembled per point (:311)
        kind, m, n, k = argv[0], int(argv[1]), int(argv[2]), int(argv[3])
        J = np.array([float(x) for x in argv[4].split(",")])
        off, sites = O.chain_neighbours(m, True) if kind == "chain" else O.square_neighbours(4, 4, True)
        h = O.hamiltonian(m, n, off, sites)
        print(json.dumps({"event": "start", "D": h.D}), flush=True)
        # the driver reports each finished point on stderr ("point p rc seconds"): sent to the parent as it happens
        os.dup2(sys.stdout.fileno(), 2)
        O.ref_irlm_sweep(h, J, 1.0, k, progress=True)
        return
    if mode == "requests":
        # one request per thread: (geometry, m, Lx, Ly, n, J, mu) each built with the oracle and solved by the
        # reference's Op::IRLM_eigen (Spectra GenEigs nev=4, ncv=9)
        from concurrent.futures import ThreadPoolExecutor, as_completed

        reqs = json.loads(argv[0])
        nthreads = int(os.environ.get("OMP_NUM_THREADS", "1"))
        os.environ["OMP_NUM_THREADS"] = "1"
        R = O.ref()
        print(json.dumps({"event": "start"}), flush=True)

        def solve(q):
            g, m, Lx, Ly, n, J, mu = q
            mm = m if g == "chain" else Lx * Ly
            off, sites = O.chain_neighbours(mm, True) if g == "chain" else O.square_neighbours(Lx, Ly, True)
            t0 = time.perf_counter()
            h = O.hamiltonian(mm, n, off, sites)
            ev = np.zeros(4)
            rc = R.ref_irlm_eigen(h.D, h.row_ptr, h.col, h.val, h.diag, J, 1.0, 4, ev)
            return {"D": h.D, "rc": int(rc), "s": time.perf_counter() - t0}

        with ThreadPoolExecutor(nthreads) as ex:
            for f in as_completed([ex.submit(solve, q) for q in reqs]):
                print(json.dumps(f.result()), flush=True)
        return
    raise SystemExit(f"unknown child mode {mode}")


def cpu_baseline(Js, k, cap_s=45.0):
    """The reference's own CPU path on the headline workload: Op::IRLM_eigen (Spectra GenEigs, nev = k,
    ncv = 2k + 1, compiled from the reference's sources in oracle/_ref) on the m=12 n=12 sector at the batch's
    J points, one point per OpenMP thread as src/analysis.cpp:302 runs its grid, for at most cap_s seconds.
    value = solves converged / elapsed (0 when none converges within the cap)."""
    model, ncpu = cpu_model()
    threads = min(len(Js), omp_threads() or ncpu)
    recs, el = run_cpu_child("sweep", ["chain", 12, 12, k, ",".join(f"{x:.17g}" for x in Js)], cap_s, threads)
    ok = [r for r in recs if r["rc"] == 0]
    return {"value": len(ok) / max(el, 1e-9), "unit": "eigensolves/s", "cores": threads, "kind": "reference",
            "cpu_model": model, "nproc": ncpu,
            "sample": f"the reference's Op::IRLM_eigen (Spectra nev={k}, ncv={2 * k + 1}) on the m=12 n=12 sector at the "
                      f"{len(Js)} J points of the batch, one per OpenMP thread ({threads}), for {el:.0f} s: "
                      f"{len(ok)} converged, {len(recs) - len(ok)} failed, {len(Js) - len(recs)} still running at the "
                      f"cap (a full run of the same sweep failed on every point after ~500 s: Spectra's 1000-restart "
                      f"limit)",
            "finished": recs}


def rerun_causes(ctx, Js, widths, m=10, k=4):
    """Why mixed solves of the sweep are rerun in FP64 (untimed): the sweep's three sectors solved again in
    mixed precision with the rerun switched off (fallback = -1), each rejected solve classified by the first
    failing test of the acceptance rule -- 'uncertified': the mixed solve itself reports it did not certify its
    k values (a copy of a degenerate level missed or left unconverged in a dense cluster); 'E0_bound':
    ||r||^2 / gap of E0 above 1e-9 |E0| (tiny |E0| of the Mott regime); 'excited_bound': a level above E0
    whose bound exceeds 1e-6 of its scale."""
    off, sites = bhx.lattice("chain", m=m)
    out = {"uncertified": 0, "E0_bound": 0, "excited_bound": 0}
    for n, W in zip((m - 1, m, m + 1), widths):
        b = ctx.basis(m, n)
        h = b.hamiltonian(off, sites)
        W = W or 32
        for c0 in range(0, len(Js), W):
            r = h.lanczos(Js[c0:c0 + W], 1.0, k=k, prec=bhx.PREC_MIXED, fallback=-1)
            for ev, ee, st in zip(r.evals, r.errest, r.status):
                if st != 0:
                    out["uncertified"] += 1
                elif not ee[0] <= 1e-9 * abs(ev[0]):
                    out["E0_bound"] += 1
                elif not np.all(ee[1:] <= 1e-6 * np.maximum(np.abs(ev[1:]), np.nanmax(ev) - ev[0])):
                    out["excited_bound"] += 1
        h.close()
        b.close()
    return out


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


def request_stream(nreq, seed):
    """nreq requests of the serving mix: (geometry, m, Lx, Ly, n, J, mu), J/U log-uniform in [0.01, 1], mu/U uniform
    in [0, 3], the sector uniform over the admissible ones."""
    rng = np.random.default_rng(seed)
    secs = serving_sectors()
    out = []
    for _ in range(nreq):
        g, m, Lx, Ly, n = secs[rng.integers(len(secs))]
        out.append((g, m, Lx, Ly, n, float(10 ** rng.uniform(-2, 0)), float(rng.uniform(0, 3))))
    return out, rng


def stream(srv, reqs, rate, rng):
    """Submits reqs with exponential gaps of mean 1/rate (Poisson arrivals), samples the server's queue
    (requests submitted and not completed) as it goes, drains.  Returns results, span, queue samples."""
    t0 = time.perf_counter()
    tnext = t0
    ids, qs = [], []
    done0 = srv.stats().completed
    for q in reqs:
        g, m, Lx, Ly, n, J, mu = q
        tnext += rng.exponential(1.0 / rate)
        dt = tnext - time.perf_counter()
        if dt > 0:
            time.sleep(dt)
        ids.append(srv.submit(g, n, J, 1.0, mu, m=m, Lx=Lx, Ly=Ly))
        if len(ids) % 20 == 0:
            qs.append((time.perf_counter() - t0, int(srv.stats().in_flight)))
        if len(ids) % 500 == 0:
            log(f"  submitted {len(ids)}/{len(reqs)}, {srv.stats().completed - done0} served")
    t_sub = time.perf_counter() - t0
    # drain, reporting each further tenth of the stream served, each request of the last tenth, and otherwise
    # the count every 20 s while requests keep completing (a line only when requests have completed)
    ok, step = False, max(len(reqs) // 10, 1)
    last = srv.stats().completed - done0
    nxt = last + step
    t_last = time.perf_counter()
    while time.perf_counter() - t0 < 1800.0:
        ok = srv.drain(2.0)
        served = srv.stats().completed - done0
        if ok or served >= nxt or (served > last and (len(reqs) - served < step or time.perf_counter() - t_last > 20.0)):
            t_last = time.perf_counter()
            st = srv.stats()
            log(f"  served {served}/{len(reqs)} at {time.perf_counter() - t0:.1f} s (in flight {st.in_flight}, "
                f"engines {st.engines}, HBM {st.hbm_used / 2**30:.1f} GiB)")
            nxt = served + step
            last = served
        if ok:
            break
    span = time.perf_counter() - t0
    res = [srv.poll(i) for i in ids]
    return res, span, t_sub, qs, ok


# sectors up to this dimension are solved in FP64 with full reorthogonalisation by design (bh_lanczos_opts.fullre_d)
FULLRE_D = 16384


def request_dim(q):
    g, m, Lx, Ly, n = q[:5]
    return math.comb((m if g == "chain" else Lx * Ly) + n - 1, n)


def rerun_breakdown(res, big, reqs):
    """FP64 reruns of mixed-lane requests against the mixed-lane requests, by J/U band and by filling (n against
    the number of sites): "band": [reruns, requests]."""
    bands = [("J<0.03", 0.0, 0.03), ("0.03<=J<0.1", 0.03, 0.1), ("0.1<=J<0.3", 0.1, 0.3), ("J>=0.3", 0.3, 2.0)]
    out = {}
    for r, bg, q in zip(res, big, reqs):
        if r is None or not bg:
            continue
        g, m, Lx, Ly, n, J = q[:6]
        sites = m if g == "chain" else Lx * Ly
        keys = [name for name, lo, hi in bands if lo <= J < hi]
        keys.append("n<sites" if n < sites else ("n=sites" if n == sites else "n>sites"))
        for key in keys:
            c = out.setdefault(key, [0, 0])
            c[0] += int(bool(r.used_fp64))
            c[1] += 1
    return out


def stream_summary(res, span, rate, t_sub, qs, ok, reqs):
    lat = np.array([r.t_done - r.t_submit for r in res if r is not None])
    big = [request_dim(q) > FULLRE_D for q in reqs]
    return {"requests": len(res), "offered_rps": rate, "completed": int(sum(r is not None for r in res)),
            "failed": int(sum(1 for r in res if r is None or r.status != 0)), "drained": bool(ok),
            "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "solves_per_s": float(len(lat) / span), "submit_s": t_sub, "span_s": span,
            "fp64_results": int(sum(1 for r in res if r is not None and r.used_fp64)),
            "fp64_by_design": int(sum(1 for r, bg in zip(res, big) if r is not None and not bg)),
            "mixed_lane_requests": int(sum(big)),
            "fp64_reruns_of_mixed": int(sum(1 for r, bg in zip(res, big) if r is not None and bg and r.used_fp64)),
            "reruns_by": rerun_breakdown(res, big, reqs),
            "queue_first_half_mean": float(np.mean([q for t, q in qs if t < t_sub / 2])) if qs else None,
            "queue_second_half_mean": float(np.mean([q for t, q in qs if t >= t_sub / 2])) if qs else None}


def serving(ctx, nreq, rate, sustained_n, seed=7):
    """The serving stream: nreq requests at `rate` req/s offered load (Poisson) from a cold server (sector
    CSRs built and the operator tuned on first use), then, on the same warm server, sustained_n requests at
    0.4 of the rate the first stream completed (the sustained operating point: the queue is sampled to show it
    does not grow).  k = 4, mixed precision with FP64 rerun on rejection, values with multiplicity."""
    reqs, rng = request_stream(nreq, seed)
    with bhx.Server(ctx, lanes=16, k=4, ncv=160) as srv:
        res, span, t_sub, qs, ok = stream(srv, reqs, rate, rng)
        cold = stream_summary(res, span, rate, t_sub, qs, ok, reqs)
        log("serving cold:", json.dumps(cold))
        sus = None
        if sustained_n > 0:
            # the cold rate comes from full batches over a deep backlog; at 0.7 of it the queue still grew
            r2 = 0.4 * cold["solves_per_s"]
            reqs2, rng2 = request_stream(sustained_n, seed + 1)
            res2, span2, t_sub2, qs2, ok2 = stream(srv, reqs2, r2, rng2)
            sus = stream_summary(res2, span2, r2, t_sub2, qs2, ok2, reqs2)
            log("serving sustained:", json.dumps(sus))
    return {"cold_stream": cold, "sustained": sus, "sectors": len(serving_sectors()),
            "mix": "chain m=6..13, square 3x3 and 3x4, n in [m-2, m+2], D <= 5e6; J/U log-uniform [0.01, 1], "
                   "mu/U uniform [0, 3], k=4"}


def serving_cpu(nreq=24, seed=11, cap_s=40.0):
    """CPU-only rate on the same mix: nreq requests drawn as the stream draws them, each built on the host by the
    oracle and solved by the reference's Op::IRLM_eigen (Spectra, nev=4, ncv=9), one request per thread, for at
    most cap_s seconds (host build included: the reference builds its Hamiltonian too)."""
    reqs, _ = request_stream(nreq, seed)
    model, ncpu = cpu_model()
    threads = min(nreq, omp_threads() or ncpu)
    recs, el = run_cpu_child("requests", [json.dumps(reqs)], cap_s, threads, setup_cap_s=60.0)
    ok = [r for r in recs if r["rc"] == 0]
    return {"solves_per_s": len(ok) / max(el, 1e-9), "cores": threads, "kind": "reference", "cpu_model": model,
            "sample": f"{nreq} requests of the mix, {len(ok)} solved ({len(recs) - len(ok)} failed) in {el:.0f} s "
                      f"(cap {cap_s:.0f} s), mean D of the solved {np.mean([r['D'] for r in ok]) if ok else 0:.0f}, "
                      f"of the sample {np.mean([math.comb((q[1] if q[0] == 'chain' else q[2] * q[3]) + q[4] - 1, q[4]) for q in reqs]):.0f}"}


def sweep_cpu(Js, cap_s=40.0):
    """The reference's CPU sweep on the chain m=10, n=10 sector (D = 92,378) at the sweep's J points: one point per
    OpenMP thread (src/analysis.cpp:302), Spectra nev=4, ncv=9, at most cap_s seconds."""
    model, ncpu = cpu_model()
    threads = min(len(Js), omp_threads() or ncpu)
    recs, el = run_cpu_child("sweep", ["chain", 10, 10, 4, ",".join(f"{x:.17g}" for x in Js)], cap_s, threads)
    ok = [r for r in recs if r["rc"] == 0]
    return {"solves_per_s": len(ok) / max(el, 1e-9), "cores": threads, "kind": "reference", "cpu_model": model,
            "sample": f"sector n=10 at the {len(Js)} J points: {len(ok)} solved, {len(recs) - len(ok)} failed in {el:.0f} s"}


def spmv_table(h, Bs, precs=((bhx_prec("MIXED"), "mixed"), (bhx_prec("FP64"), "fp64"))):
    out = {}
    for prec, name in precs:
        for B in Bs:
            ms, by = h.spmv_bench(prec, B, 30)
            out[f"{name}_B{B}"] = {"us": ms * 1e3, "alg_GBs": by / ms / 1e6, "frac_of_peak": by / ms / 1e6 / HBM_PEAK_GBS,
                                   "tuned": h.tuned(1, prec, B)}
    return out


def spmv_only():
    """The SpMV table alone: chain m=12 n=12 (lexicographic layout) at B = 1, 16 and the 4x4 torus n=10 at B = 1,
    both precisions, 30 launches each after the tuner's choice."""
    ctx = bhx.Context(0)
    out = {}
    for key, (geom, m, Lx, Ly, sites, n, Bs) in {"chain12": ("chain", 12, 0, 0, 12, 12, (1, 16)),
                                                 "square4x4_n10": ("square", 0, 4, 4, 16, 10, (1,))}.items():
        off, nb = bhx.lattice(geom, m=m, Lx=Lx, Ly=Ly, periodic=True)
        b = ctx.basis(sites, n, bhx.ORDER_LEX)
        h = b.hamiltonian(off, nb)
        out[key] = {"D": h.D, "nnz": h.nnz, **spmv_table(h, Bs)}
        log(key, json.dumps(out[key]))
        h.close()
        b.close()
    ctx.close()
    return out


def main():
    if len(sys.argv) > 2 and sys.argv[1] == "--cpu-child":
        cpu_child(sys.argv[2], sys.argv[3:])
        return
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", type=int, default=0, metavar="BYTES",
                    help="only run one device-to-device copy kernel of BYTES bytes (calibrates HBM counters) and exit")
    ap.add_argument("--serve", type=int, default=600, metavar="N",
                    help="requests of the serving stream at 200 req/s offered load (0: skip; the full stream is "
                         "--serving-only --serve 2000)")
    ap.add_argument("--serve-sustained", type=int, default=200, metavar="N",
                    help="requests of the sustained-rate stream on the warm server (0: skip)")
    ap.add_argument("--spmv-only", action="store_true",
                    help="run only the SpMV table (chain m=12 n=12 at B = 1, 16; 4x4 n=10 at B = 1) and print its JSON: "
                         "for counter passes")
    ap.add_argument("--serving-only", action="store_true",
                    help="run only the serving legs (and their CPU comparison unless --no-cpu) and print their JSON")
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

    if args.spmv_only:
        print(json.dumps(spmv_only()))
        return

    if args.serving_only:
        ctx = bhx.Context(0)
        log("device:", bhx.device_name(0))
        serve = serving(ctx, args.serve, 200.0, args.serve_sustained)
        if not args.no_cpu:
            serve["cpu_reference"] = serving_cpu()
        print(json.dumps({"serving": serve}))
        ctx.close()
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

    # plain FP64 GPU Lanczos: FP64 vectors and matrix values, unbatched (accuracy comparator and baseline rate), and the
    # FP64 values of every J point of the batch (the comparator of each mixed lane's four values)
    fp64_s, r64, E64 = float("nan"), None, None
    if not args.quick:
        r64 = ham.lanczos(J0, U, k=k, prec=bhx.PREC_FP64)
        reps64 = 3
        t0 = time.perf_counter()
        for _ in range(reps64):
            r64 = ham.lanczos(J0, U, k=k, prec=bhx.PREC_FP64)
        fp64_s = (time.perf_counter() - t0) / reps64
        log(f"plain fp64: {fp64_s * 1e3:.1f} ms/solve, steps={r64.iters[0]}, E={r64.evals[0]}")
        E64 = np.stack([ham.lanczos(Jv, U, k=k, prec=bhx.PREC_FP64).evals[0] for Jv in Js])

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
    # |dE_i| / |E_i| of every value of every lane against its own FP64 solve (values with multiplicity)
    relerr = (np.abs(res.evals - E64) / np.abs(E64)) if E64 is not None else None
    log(f"mixed batched B={B}: {ms_step:.1f} ms/step ({value:.1f} solves/s), steps={res.iters.tolist()}, "
        f"fp64 reruns={int(res.used_fp64.sum())}, |dE0|/|E0| vs fp64 = {rel0:.2e}")

    # dominant kernel: the fused Lanczos operator pass K1 (row-lane engine at B = 16).  Algorithmic bytes per
    # launch: packed matrix 4 B/entry + row pointers 4 B/row + diagonal 4 B/row, and per vector and row
    # v_j read (4 B), v_{j-1} read (4 B), the shifted w written in FP32 (4 B).
    D, nnz = ham.D, ham.nnz
    k1_bytes = 4.0 * nnz + 4.0 * (D + 1) + 4.0 * D + B * D * (4 + 4 + 4)
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
        E3_64 = np.stack([h3.lanczos(Jv, 1.0, k=4, prec=bhx.PREC_FP64).evals[0] for Jv in J3])
        ctx.set_timing(True)
        ctx.kernel_stats("lanczos_spmv", reset=True)
        t3 = time.perf_counter()
        r3 = h3.lanczos(J3, 1.0, k=4, prec=bhx.PREC_MIXED)
        t3 = time.perf_counter() - t3
        ctx.set_timing(False)
        c3, tot3 = ctx.kernel_stats("lanczos_spmv")
        k1b3 = 4.0 * h3.nnz + 4.0 * (h3.D + 1) + 4.0 * h3.D + 2 * h3.D * (4 + 4 + 4)
        k1us3 = tot3 / max(c3, 1) * 1e3
        spmv3["lanczos_mixed_B2"] = {"s": t3, "iters": r3.iters.tolist(), "E0": r3.evals[:, 0].tolist(),
                                     "evals": r3.evals.tolist(), "evals_fp64": E3_64.tolist(),
                                     "rel_err_vs_fp64": (np.abs(r3.evals - E3_64) / np.abs(E3_64)).tolist(),
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
                 "fp64_reruns": int(rsw["used_fp64"].sum()), "all_ok": bool(np.all(rsw["status"] == 0)),
                 "rerun_causes": rerun_causes(ctx, Jsw, rsw["width"].tolist())}
        if not args.no_cpu:
            sweep["cpu_reference"] = sweep_cpu(Jsw)
        log("sweep chain m=10 n=9/10/11:", json.dumps(sweep))

    mf = None
    if not args.quick and not args.no_mean_field:
        mf = mean_field_leg(ctx, 0 if args.no_cpu else 20000)
        log("mean field 1000x1000:", json.dumps(mf))

    serve = None
    if args.serve > 0 and not args.quick:
        serve = serving(ctx, args.serve, 200.0, args.serve_sustained)
        if not args.no_cpu:
            serve["cpu_reference"] = serving_cpu()
        log("serving:", json.dumps(serve))

    cpu = None
    if not args.no_cpu and not args.quick:
        cpu = cpu_baseline(Js, k)
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
        "rel_err_vs_fp64_per_level_max": relerr.max(axis=0).tolist() if relerr is not None else None,
        "rel_err_vs_fp64": relerr.tolist() if relerr is not None else None,
        "evals_lane0": res.evals[0].tolist(),
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
