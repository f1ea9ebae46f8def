"""GPU serving loop: continuous batching over sector engines.  Results against individual FP64 solves
(E0 within 1e-9 relative, levels include -mu n), alone-vs-mix equality, budget rejection."""
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
        assert r.status == 0 and np.allclose(r.evals[:4], np.array([0.0, 1.0, 2.0, 3.0]) - 0.5 * 6)
        big = srv.submit("chain", 13, 0.1, 1.0, 0.0, m=13)  # D = C(24, 12): the CSR alone exceeds 64 MiB
        small = srv.submit("chain", 6, 0.1, 1.0, 0.0, m=6)
        rs = srv.wait(small)
        rb = srv.wait(big, 120.0)
        assert rs.status == 0
        assert rb.status == bhx.EBUDGET
