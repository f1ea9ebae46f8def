"""Lanczos steps replayed from captured HIP graphs against the same solves issued launch by launch.

Within a solve the captured step graphs are replayed hundreds of times, interleaved with the host's own
eager work (analysis syncs, lock and restart kernels, copies); between solves eager SpMV calls run on the
same context.  Every replayed solve must equal the eager one bitwise (the kernels reduce in a fixed order)."""
import os

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


def solve(h, J, prec, graphs, **kw):
    os.environ["BHX_GRAPH"] = "1" if graphs else "0"
    try:
        return h.lanczos(J, 1.0, k=4, prec=prec, **kw)
    finally:
        os.environ.pop("BHX_GRAPH", None)


@pytest.mark.parametrize("m,n,prec,kw", [(10, 9, bhx.PREC_MIXED, {}),            # FP32 storage, selective orthogonalisation
                                         (10, 9, bhx.PREC_MIXED, {"ncv": 48}),   # with thick restarts
                                         (8, 8, bhx.PREC_FP64, {}),              # small sector: full reorthogonalisation
                                         (12, 12, bhx.PREC_MIXED, {})])          # D B >= 4M: speculative analyses
def test_graph_replays_equal_eager(ctx, m, n, prec, kw):
    print(f"graph m={m} n={n} prec={prec} {kw}", flush=True)
    off, sites = O.chain_neighbours(m, True)
    b = ctx.basis(m, n, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    J = np.array([0.05, 0.1, 0.2, 0.4])
    eager = solve(h, J, prec, False, **kw)
    assert np.all(eager.status == 0)
    x = np.random.default_rng(1).standard_normal((h.D, 2))
    for rep in range(3):
        r = solve(h, J, prec, True, **kw)
        assert np.array_equal(r.evals, eager.evals), rep
        assert np.array_equal(r.iters, eager.iters)
        h.spmv([0.3, 0.7], 1.0, x)  # eager work on the same context between replays
    h.close()
    b.close()
