"""Caller-chosen basis orders and the band order (reverse Cuthill-McKee of the hopping graph).
No reference counterpart: the reference has one basis order and its spectra do not depend on it.
Checked here: the permutation is one, the states at each position are the oracle's lexicographic states
at the recorded ranks, the Hamiltonian built in that order equals the oracle's Hamiltonian over those
states bit for bit, the bandwidth shrinks, and the spectrum is the lexicographic order's (1e-12 relative
in FP64; mixed E0 within 1e-9)."""
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


def lat(geom):
    if geom[0] == "chain":
        return geom[1], O.chain_neighbours(geom[1], geom[2])
    return geom[1] * geom[2], O.square_neighbours(geom[1], geom[2], geom[3])


def lex_states(m, n):
    """the oracle's states indexed by lexicographic rank"""
    st0, _, lex0 = O.basis(m, n, O.ORDER_ASCENDING)
    st = np.zeros_like(st0)
    st[lex0] = st0
    return st


def bandwidth(rp, col):
    rows = np.repeat(np.arange(rp.size - 1), np.diff(rp))
    return int(np.max(np.abs(col.astype(np.int64) - rows))) if col.size else 0


@pytest.mark.parametrize("geom,n", [(("chain", 8, True), 8), (("chain", 10, False), 7), (("square", 3, 3, True), 6),
                                    (("chain", 2, True), 4), (("chain", 6, True), 0)])
def test_band_order_is_exact(ctx, geom, n):
    m, (off, sites) = lat(geom)
    print(f"band order {geom} n={n}", flush=True)
    st_lex = lex_states(m, n)
    b = ctx.basis(m, n, bhx.ORDER_LEX)
    h_lex = b.hamiltonian(off, sites)
    rp0, col0, _, _ = h_lex.csr()
    h_lex.close()
    b.reorder_band(off, sites)
    assert b.order == bhx.ORDER_BAND
    lop, pol = b.perm()
    assert np.array_equal(np.sort(lop), np.arange(b.D))
    assert np.array_equal(pol[lop], np.arange(b.D))
    st = b.states()
    assert np.array_equal(st, st_lex[lop])
    ho = O.hamiltonian(m, n, off, sites, states=st)
    h = b.hamiltonian(off, sites)
    rp, col, val, diag = h.csr()
    assert np.array_equal(rp, ho.row_ptr) and np.array_equal(col, ho.col)
    assert np.array_equal(val, ho.val) and np.array_equal(diag, ho.diag)
    if b.D > 100:
        assert bandwidth(rp, col) < bandwidth(rp0, col0)
    h.close()
    b.close()


def test_permute_validates_and_applies(ctx):
    m, n = 5, 4
    b = ctx.basis(m, n, bhx.ORDER_LEX)
    rng = np.random.default_rng(3)
    p = rng.permutation(b.D).astype(np.int32)
    b.permute(p)
    lop, pol = b.perm()
    assert np.array_equal(lop, p) and np.array_equal(pol[p], np.arange(b.D))
    assert np.array_equal(b.states(), lex_states(m, n)[p])
    bad = p.copy()
    bad[0] = bad[1]
    with pytest.raises(bhx.BHError):
        b.permute(bad)
    with pytest.raises(bhx.BHError):
        b.permute(np.full(b.D, b.D, np.int32))
    b.close()


def test_band_order_spectrum(ctx):
    m, n = 10, 10
    off, sites = O.chain_neighbours(m, True)
    Js = np.array([0.05, 0.1, 0.2, 0.3])
    out = {}
    for order in ("lex", "band"):
        b = ctx.basis(m, n, bhx.ORDER_LEX)
        if order == "band":
            b.reorder_band(off, sites)
        h = b.hamiltonian(off, sites)
        out[order] = (h.lanczos(Js, 1.0, k=4, prec=bhx.PREC_FP64), h.lanczos(Js, 1.0, k=4, prec=bhx.PREC_MIXED))
        rp, col, _, _ = h.csr()
        out[order + "_bw"] = bandwidth(rp, col)
        h.close()
        b.close()
    f64l, mixl = out["lex"]
    f64b, mixb = out["band"]
    assert np.all(f64l.status == 0) and np.all(f64b.status == 0)
    assert np.allclose(f64b.evals, f64l.evals, rtol=1e-12, atol=0)
    assert np.all(np.abs(mixb.evals[:, 0] - f64l.evals[:, 0]) <= 1e-9 * np.abs(f64l.evals[:, 0]))
    assert out["band_bw"] * 4 < out["lex_bw"]
