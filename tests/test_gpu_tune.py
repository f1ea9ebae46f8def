"""Operator-kernel variants and their autotuner.

Every variant the autotuner can choose (LDS-staged row blocks; B / V threads per row with V vectors each,
V = 1, 2, 4 and, in the Lanczos pass, 8, or V = 1, 2 with 8 or 16 entries of a row in flight; G threads per
(row, vector)) is pinned in turn and checked against the oracle's product, with the
tolerances of the operator tests: FP64 within 1e-13, FP32-stored vectors within 2e-6 relative.  The
row-lane variants sum every (row, vector) over its entries in the same order with the same arithmetic,
so their products must agree bitwise with each other.  In a Lanczos solve the variants also partition the
rows differently among workgroups, which changes the order of the alpha partial sums: whole solves with
each variant pinned must give the same levels to the solver's accuracy (E0 within 1e-9 relative, the levels
above it within 1e-9 in FP64 and 1e-5 with FP32-stored vectors).  The autotuner's record must name the fastest of the candidates it timed."""
import numpy as np
import pytest

import bhx
import oracle as O

pytestmark = pytest.mark.gpu

SPMV, K1 = 1, 0


@pytest.fixture(scope="module")
def ctx():
    c = bhx.Context(0)
    yield c
    c.close()


def variants(h, kind, prec, B):
    """every variant bh_ham_set_tuned accepts for this sector and width"""
    out = []
    for v in (0, 1, 2, 4) + ((102, 104, 108, 116) if kind == SPMV else (8, 17, 33, 18, 34, 20)):
        try:
            h.set_tuned(kind, prec, B, v)
            out.append(v)
        except bhx.BHError:
            pass
    return out


@pytest.mark.parametrize("geom,n", [(("chain", 10, True), 10), (("square", 3, 3, True), 6)])
@pytest.mark.parametrize("B", [1, 4, 16])
def test_every_spmv_variant_matches_oracle(ctx, geom, n, B):
    print(f"spmv variants {geom} n={n} B={B}", flush=True)
    if geom[0] == "chain":
        m, (off, sites) = geom[1], O.chain_neighbours(geom[1], geom[2])
    else:
        m, (off, sites) = geom[1] * geom[2], O.square_neighbours(geom[1], geom[2], geom[3])
    ho = O.hamiltonian(m, n, off, sites)
    b = ctx.basis(m, n)
    h = b.hamiltonian(off, sites)
    rng = np.random.default_rng(7 + B)
    X = rng.standard_normal((h.D, B))
    J = rng.uniform(0.01, 1.0, B)
    U = rng.uniform(0.5, 2.0, B)
    Xf = X.astype(np.float32).astype(np.float64)
    want = {bhx.PREC_FP64: np.stack([O.spmv(ho, J[i], U[i], X[:, i].copy()) for i in range(B)], axis=1),
            bhx.PREC_MIXED: np.stack([O.spmv(ho, J[i], U[i], Xf[:, i].copy()) for i in range(B)], axis=1)}
    tol = {bhx.PREC_FP64: 1e-13, bhx.PREC_MIXED: 2e-6}
    for prec in (bhx.PREC_FP64, bhx.PREC_MIXED):
        vs = variants(h, SPMV, prec, B)
        assert 1 in vs and (B < 2 or 2 in vs)
        assert (4 in vs) == (B >= 4 and prec == bhx.PREC_MIXED)
        rowlane = None
        for v in vs:
            h.set_tuned(SPMV, prec, B, v)
            got = h.spmv(J, U, X, prec=prec)
            w = want[prec]
            assert np.allclose(got, w, rtol=tol[prec], atol=tol[prec] * np.abs(w).max()), (prec, v)
            if v in (1, 2, 4):
                if rowlane is None:
                    rowlane = got
                else:
                    assert np.array_equal(got, rowlane), (prec, v)
            assert h.tuned(SPMV, prec, B)["variant"] == v
    h.close()
    b.close()


def test_autotuner_keeps_the_fastest_candidate(ctx):
    print("autotuner record", flush=True)
    off, sites = O.chain_neighbours(10, True)
    b = ctx.basis(10, 10, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    assert h.tuned(SPMV, bhx.PREC_MIXED, 8) is None
    X = np.random.default_rng(3).standard_normal((h.D, 8))
    h.spmv(np.full(8, 0.2), np.ones(8), X, prec=bhx.PREC_MIXED)
    t = h.tuned(SPMV, bhx.PREC_MIXED, 8)
    cand = t["candidates"]
    assert {0, 1, 2, 4} <= set(cand)          # staged blocks fit (rows of <= 20 entries); V = 1, 2, 4
    assert t["variant"] == min(cand, key=cand.get)
    assert all(us > 0 for us in cand.values())
    # the Lanczos operator pass is tuned on the engine's first steps, per precision and width
    h.lanczos(np.array([0.1, 0.2, 0.3, 0.4]), 1.0, k=2, prec=bhx.PREC_MIXED)
    t1 = h.tuned(K1, bhx.PREC_MIXED, 4)
    assert t1 is not None and t1["variant"] == min(t1["candidates"], key=t1["candidates"].get)
    assert h.tuned(K1, bhx.PREC_MIXED, 8) is None
    h.close()
    b.close()


@pytest.mark.parametrize("prec", [bhx.PREC_MIXED, bhx.PREC_FP64])
def test_lanczos_k1_variants(ctx, prec):
    """whole solves with each K1 variant pinned give the same levels"""
    print(f"k1 variants prec={prec}", flush=True)
    off, sites = O.chain_neighbours(10, True)
    b = ctx.basis(10, 9, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    J = np.array([0.02, 0.05, 0.1, 0.2, 0.3, 0.5, 0.7, 1.0])
    rc, ref, _, _ = O.lanczos_mult(O.hamiltonian(10, 9, off, sites), J[3], 1.0, 4)
    assert rc == 0
    opts = {"fullre_d": -1} if prec == bhx.PREC_MIXED else {}  # D = 48,620: keep the selective path
    res = {}
    for v in variants(h, K1, prec, 8):
        h.set_tuned(K1, prec, 8, v)
        r = h.lanczos(J, 1.0, k=4, prec=prec, **opts)
        assert h.tuned(K1, prec, 8)["variant"] == v
        assert np.all(r.status == 0), v
        res[v] = r
    assert {1, 2} <= set(res) and ((4 in res) == (8 in res) == (prec == bhx.PREC_MIXED))
    # the accuracy bar of the Lanczos tests: E0 within 1e-9 relative; the levels above it within 1e-9 in FP64
    # and 1e-5 with FP32-stored vectors (lanes rerun in FP64 are held to 1e-9)
    def close(e, want, fp64):
        err = np.abs(e - want) / np.abs(want)
        return err[0] <= 1e-9 and np.all(err[1:] <= (1e-9 if fp64 else 1e-5))

    for v, r in res.items():
        for i in range(len(J)):
            both64 = prec == bhx.PREC_FP64 or (r.used_fp64[i] and res[1].used_fp64[i])
            assert close(r.evals[i], res[1].evals[i], both64), (v, i)
    assert close(res[1].evals[3], ref, prec == bhx.PREC_FP64 or res[1].used_fp64[3])
    h.close()
    b.close()
