"""GPU Lanczos against the oracle (dense FP64 eigh of the oracle's H, the oracle's FP64 Lanczos with
multiplicities), the reference's own Spectra solver run on the same H (oracle/_ref), and closed forms.
Eigenvalues are compared with their multiplicity, as the reference's Op::IRLM_eigen returns them.
Tolerances: FP64 path 1e-9 relative on every value; mixed path 1e-9 relative on E0 (the acceptance
target) and 1e-5 relative on the other values (FP32-stored Krylov vectors leave residuals of
~eps32 ||H|| sqrt(steps); excited levels in near-degenerate clusters keep ~||r||^2/gap of that)."""
MIXED_EXCITED = 1e-5
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


def build(ctx, geom, n):
    if geom[0] == "chain":
        m = geom[1]
        off, sites = O.chain_neighbours(m, geom[2])
    else:
        m = geom[1] * geom[2]
        off, sites = O.square_neighbours(geom[1], geom[2], geom[3])
    b = ctx.basis(m, n)
    return m, off, sites, b, b.hamiltonian(off, sites)


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)) / np.maximum(np.abs(b), 1e-300)


# (precision, extra options): sectors up to 16384 states run FP64 with full reorthogonalisation by default;
# "mixed_so" forces the FP32-stored path with selective orthogonalisation onto these small sectors
MODES = {"fp64": (bhx.PREC_FP64, {}), "mixed": (bhx.PREC_MIXED, {}), "mixed_so": (bhx.PREC_MIXED, {"fullre_d": -1})}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("geom,n,J,U,k", [(("chain", 8, True), 8, 1.0, 1.0, 10), (("chain", 6, True), 6, 0.1, 1.0, 4),
                                          (("chain", 7, False), 5, 0.3, 1.0, 6), (("square", 3, 3, True), 6, 0.2, 1.0, 4),
                                          (("chain", 10, True), 9, 0.05, 1.0, 4)])
def test_lanczos_matches_dense(ctx, geom, n, J, U, k, mode):
    prec, kw = MODES[mode]
    print(f"lanczos dense {geom} n={n} J={J} mode={mode}", flush=True)
    m, off, sites, b, h = build(ctx, geom, n)
    ho = O.hamiltonian(m, n, off, sites)
    if ho.D <= 7000:
        want = np.linalg.eigvalsh(ho.dense(J, U))[:k]
    else:
        rc, want, _, _ = O.lanczos_mult(ho, J, U, k)
        assert rc == 0
    r = h.lanczos(J, U, k=k, prec=prec, **kw)
    assert r.status[0] == 0
    e = rel(r.evals[0], want)
    assert e[0] <= 1e-9
    # FP32-stored Krylov vectors can miss one of two levels closer than ~||r|| (a single start vector, basis
    # orthogonality ~eps32): the forced mixed path is held to the first four levels, as the sweep and the
    # server use them; the default path serves these small sectors in FP64 with full reorthogonalisation
    nchk = min(k, 4) if mode == "mixed_so" else k
    assert np.all(e[:nchk] <= (1e-9 if r.used_fp64[0] else MIXED_EXCITED)), (r.evals[0], want)
    h.close()
    b.close()


@pytest.mark.parametrize("mode", list(MODES))
def test_batched_equals_individual(ctx, mode):
    prec, kw = MODES[mode]
    m, off, sites, b, h = build(ctx, ("chain", 8, True), 8)
    ho = O.hamiltonian(m, n := 8, off, sites)
    Js = np.array([0.02, 0.07, 0.15, 0.3, 0.6])
    r = h.lanczos(Js, 1.0, k=4, prec=prec, **kw)
    for i, J in enumerate(Js):
        want = np.linalg.eigvalsh(ho.dense(J, 1.0))[:4]
        assert rel(r.evals[i, 0], want[0]) <= 1e-9
        assert np.all(rel(r.evals[i], want) <= (1e-9 if prec == bhx.PREC_FP64 else MIXED_EXCITED))
        single = h.lanczos(J, 1.0, k=4, prec=prec, **kw)
        assert rel(single.evals[0, 0], r.evals[i, 0]) <= 1e-9
    h.close()
    b.close()


def test_closed_forms(ctx):
    # U = 0, periodic chain: E0 = -2 J N (all bosons in the k = 0 orbital)
    m, off, sites, b, h = build(ctx, ("chain", 10, True), 10)
    r = h.lanczos(0.7, 0.0, k=2, prec=bhx.PREC_FP64)
    assert abs(r.evals[0, 0] - (-2 * 0.7 * 10)) <= 1e-9 * 14
    # U = 0: the next two values = one boson moved to k = +2 pi / m or -2 pi / m (a degenerate pair)
    e1 = -2 * 0.7 * 9 - 2 * 0.7 * math.cos(2 * math.pi / 10)
    assert abs(r.evals[0, 1] - e1) <= 1e-8
    r = h.lanczos(0.7, 0.0, k=3, prec=bhx.PREC_FP64)
    assert np.allclose(r.evals[0, 1:], [e1, e1], atol=1e-8)
    # J = 0, unit filling: 0 once (the Mott state), then U with multiplicity m (m - 1) (one doubly occupied
    # and one empty site)
    r = h.lanczos(0.0, 1.0, k=4, prec=bhx.PREC_MIXED)
    assert np.allclose(r.evals[0], [0.0, 1.0, 1.0, 1.0], atol=1e-9)
    r = h.lanczos(0.0, 1.0, k=4, prec=bhx.PREC_MIXED, distinct=1)
    assert np.allclose(r.evals[0], [0.0, 1.0, 2.0, 3.0], atol=1e-9)
    h.close()
    b.close()


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("J,k", [(1.0, 10), (0.1, 10), (1.0, 20), (0.1, 25)])
def test_matches_reference_spectra(ctx, J, k):
    """The reference's Op::IRLM_eigen (Spectra GenEigs, nev=k, ncv=2k+1) on the same Hamiltonian: the same k
    values, degenerate pairs included, and the same gap ratio by the reference's own Analysis::gap_ratios."""
    m, n = 8, 8
    _, off, sites, b, h = build(ctx, ("chain", m, True), n)
    ho = O.hamiltonian(m, n, off, sites)
    ir = np.zeros(k)
    assert O.ref().ref_irlm_eigen(ho.D, ho.row_ptr, ho.col, ho.val, ho.diag, J, 1.0, k, ir) == 0
    for prec in (bhx.PREC_FP64, bhx.PREC_MIXED):
        r = h.lanczos(J, 1.0, k=k, prec=prec)
        assert r.status[0] == 0
        assert np.all(rel(r.evals[0], ir) <= 1e-9), (r.evals[0], ir)
        assert O.ref_gap_ratio(r.evals[0]) == pytest.approx(O.ref_gap_ratio(ir), abs=1e-6)
    h.close()
    b.close()


def test_forced_fallback_and_nonconvergence(ctx):
    m, off, sites, b, h = build(ctx, ("chain", 8, True), 8)
    ref = h.lanczos(0.3, 1.0, k=4, prec=bhx.PREC_FP64)
    # an impossible acceptance bound forces the FP64 rerun
    r = h.lanczos(0.3, 1.0, k=4, prec=bhx.PREC_MIXED, accept_rel=1e-300, fullre_d=-1)
    assert r.used_fp64[0] == 1
    assert np.allclose(r.evals[0], ref.evals[0], rtol=1e-12)
    # fallback disabled: reported as is
    r = h.lanczos(0.3, 1.0, k=4, prec=bhx.PREC_MIXED, accept_rel=1e-300, fallback=-1, fullre_d=-1)
    assert r.used_fp64[0] == 0
    # a small sector asked in mixed precision runs FP64 with full reorthogonalisation and says so
    r = h.lanczos(0.3, 1.0, k=4, prec=bhx.PREC_MIXED)
    assert r.used_fp64[0] == 1 and np.allclose(r.evals[0], ref.evals[0], rtol=1e-12)
    # too few steps: not converged status
    r = h.lanczos(0.3, 1.0, k=4, prec=bhx.PREC_FP64, max_iter=12)
    assert r.status[0] == bhx.ENOCONV
    h.close()
    b.close()


def test_refinement_certifies_small_hopping(ctx):
    """Chain m=10, n=9, J/U in [0.01, 0.023]: |E0| is 0.02-0.07 while the FP32-stored Krylov vectors leave Ritz
    residuals near 3e-7, so ||r||^2/gap of E0 sits above 1e-9 |E0| after the Lanczos run.  The FP64 Davidson
    refinement of the finish certifies every lane with the rerun switched off: E0 within 1e-9 of the FP64 solve
    and of the oracle's FP64 Lanczos, every level within 1e-6."""
    m, off, sites, b, h = build(ctx, ("chain", 10, True), 9)
    Js = 0.01 * 40.0 ** (np.arange(8) / 31.0)
    r = h.lanczos(Js, 1.0, k=4, prec=bhx.PREC_MIXED, fallback=-1)
    ref = h.lanczos(Js, 1.0, k=4, prec=bhx.PREC_FP64)
    assert np.all(r.status == 0) and np.all(r.used_fp64 == 0)
    assert np.all(r.errest[:, 0] <= 1e-9 * np.abs(r.evals[:, 0]))
    assert np.all(rel(r.evals[:, 0], ref.evals[:, 0]) <= 1e-9)
    assert np.all(rel(r.evals, ref.evals) <= 1e-6)
    oh = O.hamiltonian(m, 9, off, sites)
    rc, ev, _, _ = O.lanczos_mult(oh, Js[0], 1.0, 4, tol=1e-12)
    assert rc == 0 and rel(r.evals[0, 0], ev[0]) <= 1e-9
    h.close()
    b.close()


def test_deflation_without_barrier(ctx, monkeypatch):
    """Chain m=12, n=10, J/U in [0.3, 1], sixteen lanes with the deflation barrier off (BHX_DEFL_BARRIER=0, as
    in the server): the lanes of large J start their second pass while the others still run and restart.
    Every lane converges and agrees with the barrier run (E0 within 1e-9, levels within 1e-6)."""
    off, sites = O.chain_neighbours(12, True)
    b = ctx.basis(12, 10, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    rng = np.random.default_rng(5)
    Js = np.sort(10 ** rng.uniform(np.log10(0.3), 0, 16))
    ref = h.lanczos(Js, 1.0, k=4, prec=bhx.PREC_MIXED, fallback=-1, ncv=160)
    monkeypatch.setenv("BHX_DEFL_BARRIER", "0")
    r = h.lanczos(Js, 1.0, k=4, prec=bhx.PREC_MIXED, fallback=-1, ncv=160)
    assert np.all(ref.status == 0) and np.all(r.status == 0)
    assert np.all(rel(r.evals[:, 0], ref.evals[:, 0]) <= 1e-9)
    assert np.all(rel(r.evals, ref.evals) <= 1e-6)
    h.close()
    b.close()


@pytest.mark.parametrize("prec,bound", [(bhx.PREC_FP64, 1e-8), (bhx.PREC_MIXED, 3e-4)])
def test_ritz_vectors(ctx, prec, bound):
    """Returned Ritz vectors are unit vectors with small residuals against the oracle's dense H
    (bound: the Lanczos tolerance in FP64; the FP32 storage rounding, ~eps32 ||H|| sqrt(steps), mixed)."""
    m, off, sites, b, h = build(ctx, ("chain", 6, True), 6)
    ho = O.hamiltonian(m, 6, off, sites)
    r = h.lanczos(0.2, 1.0, k=3, prec=prec, want_vectors=1)
    v = h.lanczos_vectors(1, 3)[0]
    H = ho.dense(0.2, 1.0)
    for i in range(3):
        y = v[i]
        assert abs(np.linalg.norm(y) - 1.0) < 1e-12
        assert np.linalg.norm(H @ y - r.evals[0, i] * y) < bound
        assert abs(y @ H @ y - r.evals[0, i]) < 1e-10


@pytest.mark.parametrize("geom,n,J", [(("chain", 12, True), 12, 0.1)])
def test_full_size_mixed_vs_fp64(ctx, geom, n, J):
    """The single large solve of the benchmark: FP32-mixed E0 within 1e-9 of the FP64 run."""
    print(f"full size {geom} n={n}", flush=True)
    m, off, sites, b, h = build(ctx, geom, n)
    r64 = h.lanczos(J, 1.0, k=4, prec=bhx.PREC_FP64)
    r32 = h.lanczos(J, 1.0, k=4, prec=bhx.PREC_MIXED)
    assert r64.status[0] == 0 and r32.status[0] == 0
    assert rel(r32.evals[0, 0], r64.evals[0, 0]) <= 1e-9
    print("fp64", r64.evals[0], r64.iters, "mixed", r32.evals[0], r32.iters, r32.used_fp64, flush=True)
    h.close()
    b.close()


@pytest.mark.parametrize("mode", list(MODES))
def test_thick_restart_small_basis(ctx, mode):
    """A Krylov basis of 24 vectors forces thick restarts; levels still equal the dense spectrum."""
    prec, kw = MODES[mode]
    m, off, sites, b, h = build(ctx, ("chain", 8, True), 8)
    ho = O.hamiltonian(m, 8, off, sites)
    want = np.linalg.eigvalsh(ho.dense(0.3, 1.0))[:4]
    r = h.lanczos(0.3, 1.0, k=4, prec=prec, ncv=24, **kw)
    print("restarted", r.iters, r.used_fp64, r.evals[0], flush=True)
    assert r.status[0] == 0 and r.iters[0] > 24
    assert rel(r.evals[0, 0], want[0]) <= 1e-9
    assert np.all(rel(r.evals[0], want) <= (1e-9 if prec == bhx.PREC_FP64 else MIXED_EXCITED))
    h.close()
    b.close()


@pytest.mark.parametrize("mode", list(MODES))
def test_small_hopping_needs_restarts(ctx, mode):
    """Deep in the Mott regime the low band (width ~4J) is tiny against ||H|| ~ U n^2 / 2; with a basis of
    32 vectors the solve runs through several thick restarts (chain m=8, n=7, J/U=0.004)."""
    prec, kw = MODES[mode]
    m, off, sites, b, h = build(ctx, ("chain", 8, True), 7)
    ho = O.hamiltonian(m, 7, off, sites)
    want = np.linalg.eigvalsh(ho.dense(0.004, 1.0))[:4]
    r = h.lanczos(0.004, 1.0, k=4, prec=prec, ncv=32, **kw)
    print("small J", r.iters, r.used_fp64, r.evals[0], want, flush=True)
    assert r.status[0] == 0 and r.iters[0] > 32
    assert rel(r.evals[0, 0], want[0]) <= 1e-9
    # the band holds a level at ~J^2 / U: errors are measured against the band scale J
    scale = np.maximum(np.abs(want), 0.004)
    assert np.all(np.abs(r.evals[0] - want) <= (1e-9 if prec == bhx.PREC_FP64 else MIXED_EXCITED) * scale)
    h.close()
    b.close()


def test_speculative_analyses_full_size(ctx):
    """A batch large enough for the host analyses to overlap speculative steps (D B >= 4M: chain m=12 n=12,
    four lanes), with the lanes locked and rolled back to their scheduled steps: E0 of every lane within
    1e-9 of its own FP64 solve, the levels above within 1e-5."""
    print("speculative full size", flush=True)
    m, off, sites, b, h = build(ctx, ("chain", 12, True), 12)
    J = np.array([0.08, 0.1, 0.13, 0.2])
    r = h.lanczos(J, 1.0, k=4, prec=bhx.PREC_MIXED)
    assert np.all(r.status == 0)
    for i, Jv in enumerate(J):
        r64 = h.lanczos(Jv, 1.0, k=4, prec=bhx.PREC_FP64)
        e = rel(r.evals[i], r64.evals[0])
        assert e[0] <= 1e-9 and np.all(e <= (1e-9 if r.used_fp64[i] else MIXED_EXCITED)), (Jv, r.evals[i], r64.evals[0])
    h.close()
    b.close()


@pytest.mark.parametrize("prec", [bhx.PREC_FP64, bhx.PREC_MIXED])
def test_distinct_levels_option(ctx, prec):
    """distinct=1 keeps the one-pass solve: the lowest k distinct levels."""
    m, off, sites, b, h = build(ctx, ("chain", 8, True), 8)
    ho = O.hamiltonian(m, 8, off, sites)
    want = O.distinct_levels(np.linalg.eigvalsh(ho.dense(1.0, 1.0)), 6)
    r = h.lanczos(1.0, 1.0, k=6, prec=prec, distinct=1)
    assert r.status[0] == 0 and np.all(rel(r.evals[0], want) <= 1e-9)
    h.close()
    b.close()


@pytest.mark.parametrize("geom,n,J,k", [(("square", 3, 3, True), 7, 0.15, 6), (("chain", 9, True), 8, 0.0, 5),
                                        (("chain", 9, True), 8, 0.4, 6)])
def test_multiplicities_of_symmetric_lattices(ctx, geom, n, J, k):
    """The 3x3 torus has levels of multiplicity up to 4 and more: every copy among the lowest k is returned,
    in the FP64 path and in the FP32-stored path with selective orthogonalisation (mixed_so)."""
    print(f"multiplicities {geom} n={n} J={J} k={k}", flush=True)
    m, off, sites, b, h = build(ctx, geom, n)
    ho = O.hamiltonian(m, n, off, sites)
    want = np.linalg.eigvalsh(ho.dense(J, 1.0))[:k]
    for mode in ("fp64", "mixed_so"):
        prec, kw = MODES[mode]
        r = h.lanczos(J, 1.0, k=k, prec=prec, **kw)
        assert r.status[0] == 0
        tol = 1e-9 if r.used_fp64[0] else MIXED_EXCITED
        assert rel(r.evals[0, 0], want[0]) <= 1e-9 and np.all(np.abs(r.evals[0] - want) <= tol * np.maximum(np.abs(want), 1.0)), (
            mode, r.evals[0], want)
    h.close()
    b.close()


def test_dense_eigen_matches_numpy(ctx):
    """bh_dense_eigen (rocSOLVER dsyevd on the device) at the size of the reference's small exact runs
    (chain m=8, n=8: D = 6435): every eigenvalue against LAPACK, orthonormal eigenvectors with small residuals."""
    ho = O.hamiltonian(8, 8, *O.chain_neighbours(8, True))
    H = ho.dense(1.0, 1.0)
    rows = np.repeat(np.arange(ho.D), np.diff(ho.row_ptr))
    # the CSR of J T + U V: off-diagonal entries then the diagonal
    rp = np.zeros(ho.D + 1, np.int64)
    np.add.at(rp, rows + 1, 1)
    rp[1:] += 1
    rp = np.cumsum(rp)
    col = np.zeros(rp[-1], np.int32)
    val = np.zeros(rp[-1])
    for r in range(ho.D):
        a, e = ho.row_ptr[r], ho.row_ptr[r + 1]
        q = rp[r]
        col[q:q + e - a] = ho.col[a:e]
        val[q:q + e - a] = ho.val[a:e]
        col[q + e - a] = r
        val[q + e - a] = ho.diag[r]
    w, V = ctx.dense_eigen(rp, col, val, vectors=True)
    want = np.linalg.eigvalsh(H)
    assert np.allclose(w, want, rtol=0, atol=1e-11)
    for c in (0, 1, 2, 100, ho.D - 1):
        assert np.linalg.norm(H @ V[:, c] - w[c] * V[:, c]) < 1e-10
    assert abs(V[:, 0] @ V[:, 1]) < 1e-12


def test_tridiag_folm(ctx):
    """bh_lanczos_tridiag: T is the projection V^T H V of H on the returned orthonormal Lanczos basis, and an
    invariant subspace stops the recurrence (n = 1: H = J T on the ring, a space of m states)."""
    m, off, sites, b, h = build(ctx, ("chain", 8, True), 6)
    ho = O.hamiltonian(m, 6, off, sites)
    H = ho.dense(0.3, 1.0)
    a, bb, V = h.tridiag(0.3, 1.0, 40, basis=True)
    assert a.size == 40 and bb[0] == 0.0
    assert np.allclose(V @ V.T, np.eye(40), atol=1e-12)
    T = V @ H @ V.T
    Tw = np.diag(a) + np.diag(bb[1:], 1) + np.diag(bb[1:], -1)
    assert np.allclose(T, Tw, atol=1e-10)
    h.close()
    b.close()
    m, off, sites, b, h = build(ctx, ("chain", 8, True), 1)
    a, bb, _ = h.tridiag(1.0, 0.0, 30)
    # the single-particle ring has 5 distinct levels: the recurrence stops at an invariant space of size 5
    assert a.size == 5
    want = np.unique(np.round(-2 * np.cos(2 * np.pi * np.arange(8) / 8), 12))
    Tw = np.diag(a) + np.diag(bb[1:], 1) + np.diag(bb[1:], -1)
    assert np.allclose(np.linalg.eigvalsh(Tw), want, atol=1e-10)
    h.close()
    b.close()
