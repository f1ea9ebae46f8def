"""GPU parity of the basis, the CSR build and the operator application against the oracle.
Integer work is compared bit-exactly; FP64 products within 1e-13 relative, FP32-stored vectors
within 1e-6 relative (tolerances stated per test)."""
import math

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
    Lx, Ly, per = geom[1], geom[2], geom[3]
    return Lx * Ly, O.square_neighbours(Lx, Ly, per)


BASIS_CASES = [(1, 0), (1, 5), (2, 7), (3, 2), (5, 5), (8, 8), (6, 4), (10, 9), (10, 11), (12, 5),
               (4, 30),   # n > 15: 8-bit site fields; key 7^30 > 2^64: 128-bit keys
               (20, 3),   # m > 16: two 64-bit words
               (9, 16),   # n > 15 and m > 8: both words of 8-bit fields in use
               (9, 9)]


@pytest.mark.parametrize("order", [bhx.ORDER_REFERENCE, bhx.ORDER_ASCENDING, bhx.ORDER_LEX])
@pytest.mark.parametrize("m,n", BASIS_CASES)
def test_basis_bit_exact(ctx, m, n, order):
    print(f"basis m={m} n={n} order={order}", flush=True)
    D = O.dimension(m, n)
    if D > 3_000_000:
        pytest.skip("oracle size")
    if order == bhx.ORDER_LEX:
        st0, _, lex0 = O.basis(m, n, O.ORDER_ASCENDING)
        st = np.zeros_like(st0)
        st[lex0] = st0  # position = lexicographic rank
        lex = np.arange(D)
        tags = np.array([sum(float(v) * np.log(p) for v, p in zip(row, [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127, 131])) for row in st]) if D < 5000 else None
    else:
        st, tags, lex = O.basis(m, n, order)
    b = ctx.basis(m, n, order)
    assert b.D == D
    assert np.array_equal(b.states(), st)
    lop, pol = b.perm()
    assert np.array_equal(lop.astype(np.int64), lex)
    assert np.array_equal(pol[lop], np.arange(D))
    if tags is not None:
        assert np.allclose(b.tags(), tags, rtol=0, atol=1e-12) if order == bhx.ORDER_LEX else np.array_equal(b.tags(), tags)
    b.close()


HAM_CASES = [(("chain", 8, True), 8, 0), (("chain", 8, False), 8, 0), (("chain", 5, True), 5, 3),
             (("chain", 2, True), 4, 0), (("square", 3, 3, True), 6, 0), (("square", 4, 3, True), 5, 0),
             (("square", 2, 2, True), 4, 0), (("square", 3, 2, False), 5, 1), (("chain", 10, True), 11, 0),
             (("chain", 6, True), 20, 2), (("square", 4, 4, True), 4, 0)]


@pytest.mark.parametrize("geom,n,flags", HAM_CASES)
def test_hamiltonian_bit_exact(ctx, geom, n, flags):
    m, (off, sites) = lat(geom)
    print(f"hamiltonian {geom} n={n} flags={flags}", flush=True)
    st, _, _ = O.basis(m, n)
    ho = O.hamiltonian(m, n, off, sites, flags=flags, states=st)
    b = ctx.basis(m, n)
    h = b.hamiltonian(off, sites, flags)
    rp, col, val, diag = h.csr()
    assert h.nnz == ho.nnz
    assert np.array_equal(rp, ho.row_ptr)
    assert np.array_equal(col, ho.col)
    assert np.array_equal(val, ho.val)  # -sqrt(code) equals the summed FP64 amplitudes
    assert np.array_equal(diag, ho.diag)
    h.close()
    b.close()


def test_invalid_neighbours_rejected(ctx):
    b = ctx.basis(4, 2)
    off = np.array([0, 1, 1, 1, 1], np.int32)
    with pytest.raises(bhx.BHError):
        b.hamiltonian(off, np.array([1], np.int32))  # not symmetric
    with pytest.raises(bhx.BHError):
        b.hamiltonian(np.array([0, 1, 1, 1, 1], np.int32), np.array([0], np.int32))  # self loop
    b.close()


@pytest.mark.parametrize("geom,n", [(("chain", 8, True), 8), (("square", 3, 3, True), 6), (("chain", 10, True), 10)])
@pytest.mark.parametrize("B", [1, 2, 8, 32])
def test_spmv_matches_oracle(ctx, geom, n, B):
    m, (off, sites) = lat(geom)
    print(f"spmv {geom} n={n} B={B}", flush=True)
    ho = O.hamiltonian(m, n, off, sites)
    b = ctx.basis(m, n)
    h = b.hamiltonian(off, sites)
    rng = np.random.default_rng(B)
    X = rng.standard_normal((h.D, B))
    J = rng.uniform(0.01, 1.0, B)
    U = rng.uniform(0.5, 2.0, B)
    want = np.stack([O.spmv(ho, J[i], U[i], X[:, i].copy()) for i in range(B)], axis=1)
    got = h.spmv(J, U, X, prec=bhx.PREC_FP64)
    assert np.allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())
    Xf = X.astype(np.float32).astype(np.float64)
    want32 = np.stack([O.spmv(ho, J[i], U[i], Xf[:, i].copy()) for i in range(B)], axis=1)
    got32 = h.spmv(J, U, X, prec=bhx.PREC_MIXED)
    assert np.allclose(got32, want32, rtol=2e-6, atol=2e-6 * np.abs(want32).max())
    h.close()
    b.close()


@pytest.mark.parametrize("geom,n", [(("chain", 12, True), 12), (("square", 4, 4, True), 10)])
def test_full_size_build_properties(ctx, geom, n):
    """Configurations of the benchmark: D and nnz closed forms, and T = T^T on sampled rows."""
    m, (off, sites) = lat(geom)
    print(f"full-size {geom} n={n}", flush=True)
    deg = (off[1:] - off[:-1])[0]
    b = ctx.basis(m, n)
    assert b.D == math.comb(m + n - 1, n)
    h = b.hamiltonian(off, sites)
    assert h.nnz == m * deg * math.comb(n + m - 2, m - 1)
    rp, col, val, diag = h.csr()
    rng = np.random.default_rng(0)
    for r in rng.integers(0, h.D, 2000):
        for a in range(rp[r], rp[r + 1]):
            c = col[a]
            seg = slice(rp[c], rp[c + 1])
            hit = np.nonzero(col[seg] == r)[0]
            assert hit.size == 1 and val[seg][hit[0]] == val[a]
        assert np.all(np.diff(col[rp[r]:rp[r + 1]]) > 0)
    h.close()
    b.close()
