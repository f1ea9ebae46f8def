"""GPU serving loop: continuous batching over sector engines.  Results against individual FP64 solves
(E0 within 1e-9 relative, levels include -mu n, values with their multiplicity), alone-vs-mix equality,
budget rejection.  Sectors of at most 16384 states are served in FP64 with full reorthogonalisation; the
larger ones (MIXED_SECTORS) run FP32-stored lanes with FP64 reductions and an FP64 rerun on rejection."""
import numpy as np
import pytest

import bhx
import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = bhx.Context(0)
    yield c
    c.close()


MIX = [("chain", 6, 0, 0, 6), ("chain", 8, 0, 0, 7), ("square", 0, 3, 3, 8), ("chain", 7, 0, 0, 9)]


def fp64_levels(ctx, geom, m, Lx, Ly, n, J, U, k=4):
    mm = m if geom == "chain" else Lx * Ly
    off, sites = (O.chain_neighbours(mm, True) if geom == "chain" else O.square_neighbours(Lx, Ly, True))
    b = ctx.basis(mm, n, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    r = h.lanczos(J, U, k=k, prec=bhx.PREC_FP64)
    h.close()
    b.close()
    return r.evals[0]


def test_server_mixed_stream(ctx):
    rng = np.random.default_rng(7)
    reqs = []
    with bhx.Server(ctx, lanes=8) as srv:
        for i in range(40):
            g = MIX[i % len(MIX)]
            J = float(10 ** rng.uniform(-2, 0))
            mu = float(rng.uniform(0, 3))
            rid = srv.submit(g[0], g[4], J, 1.0, mu, m=g[1], Lx=g[2], Ly=g[3])
            reqs.append((rid, g, J, mu))
        assert srv.drain(300.0)
        st = srv.stats()
        assert st.completed == 40 and st.p99_latency >= st.p50_latency > 0
        for rid, g, J, mu in reqs[::5]:
            r = srv.poll(rid)
            assert r is not None and r.status == 0
            want = fp64_levels(ctx, g[0], g[1], g[2], g[3], g[4], J, 1.0) - mu * g[4]
            assert abs(r.evals[0] - want[0]) <= 1e-9 * abs(want[0])
            assert np.allclose(np.array(r.evals[:4]), want, rtol=1e-5)


def test_alone_equals_in_mix(ctx):
    with bhx.Server(ctx, lanes=8) as srv:
        alone = srv.wait(srv.submit("chain", 8, 0.17, 1.0, 0.4, m=8))
        rng = np.random.default_rng(3)
        ids = []
        for i in range(30):
            g = MIX[i % len(MIX)]
            ids.append(srv.submit(g[0], g[4], float(rng.uniform(0.01, 1)), 1.0, 0.0, m=g[1], Lx=g[2], Ly=g[3]))
            if i == 11:
                target = srv.submit("chain", 8, 0.17, 1.0, 0.4, m=8)
        mixed = srv.wait(target)
        assert srv.drain(300.0)
        assert alone.status == 0 and mixed.status == 0
        assert list(alone.evals[:4]) == list(mixed.evals[:4])  # bitwise: a lane does not see its neighbours


def test_diagonal_request_and_budget(ctx):
    with bhx.Server(ctx, lanes=4, hbm_budget=64 << 20) as srv:
        r = srv.wait(srv.submit("chain", 6, 0.0, 1.0, 0.5, m=6))
        # J = 0, unit filling: the Mott state, then U with multiplicity m (m - 1)
        assert r.status == 0 and np.allclose(r.evals[:4], np.array([0.0, 1.0, 1.0, 1.0]) - 0.5 * 6)
        big = srv.submit("chain", 13, 0.1, 1.0, 0.0, m=13)  # D = C(24, 12): the CSR alone exceeds 64 MiB
        small = srv.submit("chain", 6, 0.1, 1.0, 0.0, m=6)
        rs = srv.wait(small)
        rb = srv.wait(big, 120.0)
        assert rs.status == 0
        assert rb.status == bhx.EBUDGET


# sectors above the full-reorthogonalisation size: chain m=10 n=10 (D = 92,378), 3x3 n=11 (D = 75,582),
# chain m=12 n=11 (D = 705,432)
MIXED_SECTORS = [("chain", 10, 0, 0, 10), ("square", 0, 3, 3, 11), ("chain", 12, 0, 0, 11)]


def test_server_mixed_lanes(ctx):
    """Requests on sectors served by FP32-stored lanes: every result's E0 within 1e-9 of an individual FP64 solve,
    the other three values within 1e-5 (1e-9 for a result that was rerun in FP64), most results from mixed lanes."""
    rng = np.random.default_rng(11)
    reqs = []
    with bhx.Server(ctx, lanes=8) as srv:
        for i in range(18):
            g = MIXED_SECTORS[i % len(MIXED_SECTORS)]
            J = float(10 ** rng.uniform(-1.5, -0.3))
            mu = float(rng.uniform(0, 3))
            reqs.append((srv.submit(g[0], g[4], J, 1.0, mu, m=g[1], Lx=g[2], Ly=g[3]), g, J, mu))
        assert srv.drain(600.0)
        res = [(srv.poll(rid), g, J, mu) for rid, g, J, mu in reqs]
    mixed = 0
    for r, g, J, mu in res:
        print("served", g, J, r.status, r.used_fp64, list(r.evals[:4]), flush=True)
        assert r is not None and r.status == 0
        want = fp64_levels(ctx, g[0], g[1], g[2], g[3], g[4], J, 1.0) - mu * g[4]
        got = np.array(r.evals[:4])
        assert abs(got[0] - want[0]) <= 1e-9 * abs(want[0])
        tol = 1e-9 if r.used_fp64 else 1e-5
        assert np.all(np.abs(got - want) <= tol * np.maximum(np.abs(want), 1.0)), (g, J, got, want)
        mixed += r.used_fp64 == 0
    assert mixed >= len(res) // 2


def test_alone_equals_in_mix_mixed_lane(ctx):
    """A request on an FP32-stored lane gives the same values alone and in a mix of other sectors' requests
    (within 1e-10 relative on E0 and 1e-8 on the others: the engine's width follows the queue, and the
    operator variant and the partial sums of the reductions follow the width)."""
    with bhx.Server(ctx, lanes=8) as srv:
        alone = srv.wait(srv.submit("chain", 10, 0.2, 1.0, 0.7, m=10), 300.0)
        rng = np.random.default_rng(5)
        ids, target = [], None
        for i in range(16):
            g = (MIXED_SECTORS + MIX)[i % (len(MIXED_SECTORS) + len(MIX))]
            ids.append(srv.submit(g[0], g[4], float(rng.uniform(0.02, 0.5)), 1.0, 0.0, m=g[1], Lx=g[2], Ly=g[3]))
            if i == 5:
                target = srv.submit("chain", 10, 0.2, 1.0, 0.7, m=10)
        mixed = srv.wait(target, 300.0)
        assert srv.drain(600.0)
    assert alone.status == 0 and mixed.status == 0
    assert alone.used_fp64 == 0 and mixed.used_fp64 == 0
    a, b = np.array(alone.evals[:4]), np.array(mixed.evals[:4])
    assert abs(a[0] - b[0]) <= 1e-10 * abs(a[0])
    assert np.all(np.abs(a - b) <= 1e-8 * np.maximum(np.abs(a), 1.0)), (a, b)


def test_server_budget_contention(ctx):
    """Three sectors under a budget that holds one sector's engines at a time, every mixed result rejected (an
    acceptance bound no solve meets), so that each sector's FP64 reruns wait for memory while the other sectors
    hold idle mixed engines and CSRs: the server must release those (an idle engine with nothing queued for
    it, and when nothing runs, other sectors' CSRs) instead of waiting for memory that never frees.  Every
    request completes, from an FP64 solve, with E0 within 1e-9 of a direct FP64 solve."""
    sectors = [("chain", 9, 9), ("chain", 10, 8), ("chain", 8, 10)]  # D = 24,310 / 24,310 / 19,448
    rng = np.random.default_rng(5)
    reqs = []
    with bhx.Server(ctx, lanes=4, hbm_budget=160 << 20, ncv=64, accept_rel=1e-300) as srv:
        for g, m, n in sectors:
            for _ in range(4):
                J = float(10 ** rng.uniform(-1.2, -0.5))
                reqs.append((srv.submit(g, n, J, 1.0, 0.0, m=m), m, n, J))
        assert srv.drain(300.0), srv.stats().in_flight
        for rid, m, n, J in reqs:
            r = srv.poll(rid)
            assert r is not None and r.status == 0 and r.used_fp64 == 1
            b = ctx.basis(m, n)
            h = b.hamiltonian(*bhx.lattice("chain", m=m))
            want = h.lanczos(J, 1.0, k=4, prec=bhx.PREC_FP64).evals[0]
            assert abs(r.evals[0] - want[0]) <= 1e-9 * abs(want[0])
            h.close()
            b.close()


def test_lanes_at_different_passes(ctx):
    """Chain m=12, n=10, sixteen requests at J/U in [0.3, 1] on one engine: the lanes of large J finish their
    first pass and start the next (deflation) while the others still iterate and restart.  A thick restart
    writes rows 0..p of the recurrence back for every lane; those rows are read after the speculative steps,
    so a lane early in its pass keeps its own.  Every request is certified in mixed precision (no FP64 rerun)
    and E0 matches the batch solve within 1e-9."""
    rng = np.random.default_rng(5)
    Js = np.sort(10 ** rng.uniform(np.log10(0.3), 0, 16))
    with bhx.Server(ctx, lanes=16, k=4, ncv=160) as srv:
        ids = [srv.submit("chain", 10, float(J), 1.0, 0.0, m=12) for J in Js]
        assert srv.drain(300)
        res = [srv.wait(i) for i in ids]
    off, sites = O.chain_neighbours(12, True)
    b = ctx.basis(12, 10, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    ref = h.lanczos(Js, 1.0, k=4, prec=bhx.PREC_MIXED, ncv=160)
    h.close()
    b.close()
    assert all(r.status == 0 for r in res)
    assert [int(r.used_fp64) for r in res] == [0] * 16
    E0 = np.array([r.evals[0] for r in res])
    assert np.all(np.abs(E0 - ref.evals[:, 0]) <= 1e-9 * np.abs(ref.evals[:, 0]))
