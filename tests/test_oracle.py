"""CPU tests of the oracle (the checker), pinned against the reference's own compiled functions
(oracle/_ref, when built), brute force and closed forms.  No GPU."""
import itertools
import math
import os

import numpy as np
import pytest

import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PRIMES = [2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97]


def brute_states(m, n):
    """All occupations with sum n, descending lexicographic (src/hamiltonian.cpp:60-85 order)."""
    out = [s for s in itertools.product(range(n, -1, -1), repeat=m) if sum(s) == n]
    return out  # product over descending ranges is already descending lexicographic


def key(s):
    k = 1
    for i, v in enumerate(s):
        k *= PRIMES[i] ** v
    return k


@pytest.mark.parametrize("m,n", [(1, 0), (1, 4), (2, 3), (3, 3), (4, 2), (5, 5), (6, 3), (3, 7)])
def test_dimension_and_enumeration(m, n):
    lex = brute_states(m, n)
    assert O.dimension(m, n) == len(lex) == math.comb(m + n - 1, n)
    st, tags, lexrank = O.basis(m, n, O.ORDER_ASCENDING)
    # ascending-key order = sorted by the exact integer key
    want = sorted(lex, key=key)
    assert [tuple(r) for r in st] == want
    assert [lex[i] for i in lexrank] == want
    # reference order: the state of lex rank p sits at position sortperm[p]
    sortperm = sorted(range(len(lex)), key=lambda p: key(lex[p]))
    ref = [None] * len(lex)
    for p, q in enumerate(sortperm):
        ref[q] = lex[p]
    st0, tags0, _ = O.basis(m, n, O.ORDER_REFERENCE)
    assert [tuple(r) for r in st0] == ref
    # tags accumulate n_i * log(p_i) in site order
    for r, t in zip(st0, tags0):
        acc = 0.0
        for i, v in enumerate(r):
            acc += float(v) * math.log(PRIMES[i])
        assert acc == t


def test_log_tag_order_equals_integer_order():
    # the reference sorts double tags; the integer key gives the same order wherever both are defined
    for m, n in [(8, 8), (10, 9), (6, 12)]:
        st, tags, _ = O.basis(m, n, O.ORDER_ASCENDING)
        assert np.all(np.diff(tags) > 0)


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("m,n", [(2, 5), (3, 2), (5, 5), (8, 8), (6, 4), (10, 9), (4, 12)])
def test_basis_matches_reference_build(m, n):
    R = O.ref()
    st, tags, _ = O.basis(m, n)
    D = st.shape[0]
    rt = np.zeros(D)
    rs = np.zeros(D * m)
    assert R.ref_fixed_set_basis(m, n, rt, rs) == D
    assert np.array_equal(rs.reshape(D, m).astype(np.uint8), st)
    assert np.array_equal(rt, tags)


@pytest.mark.parametrize("m,n", [(3, 3), (4, 3), (5, 2)])
def test_max_basis_matches_reference_fixture(m, n):
    """variable boson number: BH::max_set_basis as the reference build wrote it (tests/golden/maxbasis_*)"""
    z = np.load(os.path.join(GOLDEN, f"maxbasis_m{m}_n{n}.npz"))
    st, tags = O.max_basis(m, n)
    assert np.array_equal(st, z["states"]) and np.array_equal(tags, z["tags"])
    # brute force: each fixed-N basis in the reference order (lex rank p at position sortperm[p]), the
    # bases concatenated for N = 1..n, and the same placement once more over the concatenation
    def placed(seq):
        out = [None] * len(seq)
        for p, q in enumerate(sorted(range(len(seq)), key=lambda p: key(seq[p]))):
            out[q] = seq[p]
        return out

    concat = [s for N in range(1, n + 1) for s in placed(brute_states(m, N))]
    assert [tuple(r) for r in st] == placed(concat)
    assert not np.all(np.diff(tags) > 0)  # the placement is the inverse of the sort: not ascending


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("m,n", [(2, 3), (3, 3), (4, 4), (6, 3)])
def test_max_basis_matches_reference_build(m, n):
    R = O.ref()
    st, tags = O.max_basis(m, n)
    D = st.shape[0]
    rt = np.zeros(D)
    rs = np.zeros(D * m)
    assert R.ref_max_set_basis(m, n, rt, rs) == D
    assert np.array_equal(rs.reshape(D, m).astype(np.uint8), st)
    assert np.array_equal(rt, tags)


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("m,n", [(3, 3), (5, 5), (8, 8), (7, 9)])
def test_diagonal_terms_match_reference_build(m, n):
    R = O.ref()
    st, _, _ = O.basis(m, n)
    D = st.shape[0]
    h = O.hamiltonian(m, n, *O.chain_neighbours(m), flags=O.CONV_REF_INTERACTION, states=st)
    d = np.zeros(D)
    R.ref_diagonal_term(m, n, 1, d)
    assert np.array_equal(d, h.diag)  # fill_interaction: sum n(n+1)
    R.ref_diagonal_term(m, n, 2, d)
    assert np.all(d == -float(n))  # fill_chemical: -N on the diagonal
    hc = O.hamiltonian(m, n, *O.chain_neighbours(m), states=st)
    # canonical V = sum n(n-1)/2 is an affine image of the reference's: sum n(n+1) = 2 V + 2 N
    assert np.array_equal(h.diag, 2.0 * hc.diag + 2.0 * n)


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
def test_neighbours_match_reference_build():
    R = O.ref()
    for m in (2, 3, 5, 8, 12):
        for closed in (0, 1):
            off = np.zeros(m + 1, np.int32)
            s = np.zeros(2 * m + 2, np.int32)
            t = R.ref_chain_neighbours(m, closed, off, s)
            o2, s2 = O.chain_neighbours(m, bool(closed))
            assert np.array_equal(off, o2) and np.array_equal(s[:t], s2)
    # the reference's square lattice fills a shadowing local (src/neighbours.cpp:45): empty lists
    off = np.zeros(17, np.int32)
    s = np.zeros(80, np.int32)
    assert R.ref_square_neighbours(16, 1, off, s) == 0
    assert R.ref_square_neighbours(15, 1, off, s) == -1  # non-square m throws


def test_square_neighbours_hand_written():
    off, s = O.square_neighbours(3, 3, True)
    lists = [sorted(s[off[i]:off[i + 1]]) for i in range(9)]
    assert lists[0] == [1, 2, 3, 6] and lists[4] == [1, 3, 5, 7] and lists[8] == [2, 5, 6, 7]
    off, s = O.square_neighbours(3, 3, False)
    lists = [sorted(s[off[i]:off[i + 1]]) for i in range(9)]
    assert lists[0] == [1, 3] and lists[4] == [1, 3, 5, 7]
    off, s = O.square_neighbours(4, 3, True)  # 4 columns x 3 rows
    lists = [sorted(s[off[i]:off[i + 1]]) for i in range(12)]
    assert lists[0] == [1, 3, 4, 8] and lists[5] == [1, 4, 6, 9]


def adjacency(m, off, sites):
    A = np.zeros((m, m))
    for i in range(m):
        for a in range(off[i], off[i + 1]):
            A[i, sites[a]] += 1
    return A


def free_boson_spectrum(eps, n):
    """All many-body energies sum_k eps_k n_k with sum n_k = n (multiset sums)."""
    return np.sort([sum(c) for c in itertools.combinations_with_replacement(eps, n)])


@pytest.mark.parametrize("geom", [("chain", 5, True), ("chain", 6, False), ("square", (3, 3), True),
                                  ("square", (2, 3), False), ("chain", 2, True)])
def test_hopping_spectrum_closed_form(geom):
    kind, size, per = geom
    if kind == "chain":
        m = size
        off, sites = O.chain_neighbours(m, per)
    else:
        m = size[0] * size[1]
        off, sites = O.square_neighbours(size[0], size[1], per)
    n = 3
    h = O.hamiltonian(m, n, off, sites)
    H = h.dense(J=1.0, U=0.0)
    assert np.array_equal(H, H.T)
    eps = np.linalg.eigvalsh(-adjacency(m, off, sites))
    want = free_boson_spectrum(eps, n)
    got = np.linalg.eigvalsh(H)
    assert np.allclose(got, want, atol=1e-10)
    # reference hop convention doubles T
    h2 = O.hamiltonian(m, n, off, sites, flags=O.CONV_REF_HOP)
    assert np.array_equal(h2.val, 2.0 * h.val)


def test_nnz_closed_form_and_traces():
    for (m, n) in [(6, 6), (8, 8), (10, 9)]:
        off, sites = O.chain_neighbours(m, True)
        st, _, _ = O.basis(m, n)
        h = O.hamiltonian(m, n, off, sites, states=st)
        assert h.nnz == m * 2 * math.comb(n + m - 2, m - 1)
        # Tr T^2 = sum over states and hops of (n_i + 1) n_j
        tr2 = sum(float(v) ** 2 for v in h.val)
        want = 0.0
        for s in st:
            for i in range(m):
                for a in range(off[i], off[i + 1]):
                    j = sites[a]
                    if s[j] >= 1:
                        want += (s[i] + 1) * s[j]
        assert abs(tr2 - want) < 1e-9 * want
    off, sites = O.square_neighbours(4, 4, True)
    h = O.hamiltonian(16, 3, off, sites)
    assert h.nnz == 16 * 4 * math.comb(3 + 16 - 2, 15)


def test_single_particle_is_minus_adjacency():
    m = 7
    off, sites = O.chain_neighbours(m, True)
    st, _, _ = O.basis(m, 1)
    h = O.hamiltonian(m, 1, off, sites, states=st)
    site = [int(np.argmax(s)) for s in st]
    A = adjacency(m, off, sites)
    H = h.dense(1.0, 0.0)
    for r in range(m):
        for c in range(m):
            assert H[r, c] == -A[site[r], site[c]]


def test_atomic_limit_spectrum():
    m, n = 6, 6
    h = O.hamiltonian(m, n, *O.chain_neighbours(m))
    ev = np.linalg.eigvalsh(h.dense(J=0.0, U=1.0))
    assert np.array_equal(ev, np.sort(h.diag))
    assert ev[0] == 0.0 and ev[1] == 1.0  # unit filling: Mott state, first excitation U
    assert np.sum(ev == 1.0) == m * (m - 1)


@pytest.mark.parametrize("m,n,J,U", [(5, 5, 1.0, 1.0), (6, 6, 0.1, 1.0), (8, 8, 1.0, 1.0), (7, 5, 0.3, 1.0)])
def test_oracle_lanczos_matches_dense(m, n, J, U):
    h = O.hamiltonian(m, n, *O.chain_neighbours(m))
    w = np.linalg.eigvalsh(h.dense(J, U))
    k = 6
    rc, ev, bd, it = O.lanczos(h, J, U, k)
    assert rc == 0
    want = O.distinct_levels(w, k)
    assert np.allclose(ev, want, rtol=0, atol=1e-9 * max(1.0, abs(want[0])))


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("m,n,J,U", [(5, 5, 1.0, 1.0), (8, 8, 1.0, 1.0), (8, 8, 0.1, 1.0)])
def test_reference_spectra_and_eigen_on_oracle_h(m, n, J, U):
    """The reference's own solvers (Spectra GenEigs via Op::IRLM_eigen, Eigen dense via
    Op::exact_eigen) on the oracle's Hamiltonian agree with numpy and with the oracle Lanczos."""
    R = O.ref()
    h = O.hamiltonian(m, n, *O.chain_neighbours(m))
    w = np.linalg.eigvalsh(h.dense(J, U))
    if h.D <= 2000:
        ex = np.zeros(h.D)
        assert R.ref_exact_eigen(h.D, h.row_ptr, h.col, h.val, h.diag, J, U, ex) == 0
        assert np.allclose(ex, w, atol=1e-10)
    ir = np.zeros(10)
    assert R.ref_irlm_eigen(h.D, h.row_ptr, h.col, h.val, h.diag, J, U, 10, ir) == 0
    assert np.allclose(ir, w[:10], atol=1e-9)  # Spectra returns levels with multiplicity
    rc, ev, rs, _ = O.lanczos_mult(h, J, U, 10)
    assert rc == 0 and np.allclose(ev, ir, rtol=0, atol=1e-9) and np.all(rs < 1e-8)


def test_gap_ratio():
    assert O.gap_ratio([0.0, 1.0, 3.0]) == pytest.approx(0.5)
    assert O.gap_ratio([0.0, 1.0, 1.0, 2.0]) == pytest.approx(0.0)  # zero gaps give 0
    ev = np.array([0.1, 0.7, 0.9, 2.0, 2.1, 3.5])
    r = O.gap_ratio(ev)
    assert O.gap_ratio(3.0 * ev - 7.0) == pytest.approx(r)  # positive affine invariance
    assert O.gap_ratio(ev[::-1].copy()) == pytest.approx(r)  # sorted internally


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
def test_gap_ratios_and_coherence_match_reference_build():
    """Analysis::gap_ratios and Analysis::coherence as the reference computes them"""
    R = O.ref()
    rng = np.random.default_rng(5)
    for n in (3, 4, 7, 10, 20):
        for deg in (False, True):
            ev = rng.standard_normal(n)
            if deg:
                ev[1] = ev[0]                      # a zero gap
                ev[-1] = ev[-2]
            want = np.zeros(n - 2)
            R.ref_gap_ratios(np.ascontiguousarray(ev), n, want)
            got = np.zeros(n - 2)
            mean = O.lib().orc_gap_ratio(n, np.ascontiguousarray(ev), got.ctypes.data_as(O.C.c_void_p))
            assert np.array_equal(got, want)
            assert mean == pytest.approx(float(np.mean(want)), rel=1e-15, abs=1e-300)
    for m in (2, 5, 12):
        a = rng.standard_normal((m, m))
        rho = a @ a.T
        assert O.coherence(rho) == pytest.approx(R.ref_coherence(np.ascontiguousarray(rho).reshape(-1), m), rel=1e-14)


def test_start_vector_park_miller():
    r = O.start_vector(3, 0)
    # seed 0 -> 1; x1 = 16807, x2 = 16807^2 mod (2^31-1), x3 = 16807^3 mod (2^31-1)
    p = 2 ** 31 - 1
    xs = [16807 % p, 16807 ** 2 % p, 16807 ** 3 % p]
    assert np.array_equal(r, np.array([x / p - 0.5 for x in xs]))


def test_tridiag_ql():
    rng = np.random.default_rng(1)
    d = rng.standard_normal(30)
    e = rng.standard_normal(29)
    T = np.diag(d) + np.diag(e, 1) + np.diag(e, -1)
    w, z = O.tridiag_eig(d, e, vectors=True)
    assert np.allclose(w, np.linalg.eigvalsh(T), atol=1e-12)
    assert np.allclose(T @ z, z * w, atol=1e-10)


def brute_hop_operator(states, i, j):
    """b_i^+ b_j as a dense matrix on the basis `states` (integer occupations, explicit ladder amplitudes)"""
    idx = {tuple(int(x) for x in s): k for k, s in enumerate(states)}
    D = len(states)
    op = np.zeros((D, D))
    for k, s in enumerate(states):
        t = [int(x) for x in s]
        if t[j] == 0:
            continue
        amp = math.sqrt(t[j])
        t[j] -= 1
        amp *= math.sqrt(t[i] + 1)
        t[i] += 1
        op[idx[tuple(t)], k] += amp
    return op


@pytest.mark.parametrize("m,n", [(3, 2), (4, 3), (5, 2)])
def test_spdm_matches_ladder_operators(m, n):
    states, _, _ = O.basis(m, n)
    rng = np.random.default_rng(m * 10 + n)
    phi = rng.standard_normal(states.shape[0])
    phi /= np.linalg.norm(phi)
    rho = O.spdm(states, phi)
    want = np.array([[phi @ brute_hop_operator(states, i, j) @ phi for j in range(m)] for i in range(m)])
    assert np.allclose(rho, want, atol=1e-13)
    assert np.trace(rho) == pytest.approx(n)  # sum_i <n_i> = n for a normalised state


def test_spdm_closed_forms():
    # U = 0: the ground state of the periodic chain puts every boson in the k = 0 mode, rho_ij = n / m:
    # condensate fraction 1, coherence 1 - 1/m
    m, n = 6, 4
    off, sites = O.chain_neighbours(m, True)
    states, _, _ = O.basis(m, n)
    h = O.hamiltonian(m, n, off, sites, states=states)
    _, v = np.linalg.eigh(h.dense(1.0, 0.0))
    rho = O.spdm(states, v[:, 0])
    assert np.allclose(rho, n / m, atol=1e-10)
    assert O.condensate_fraction(rho) == pytest.approx(1.0, abs=1e-10)
    assert O.coherence(rho) == pytest.approx(1.0 - 1.0 / m, abs=1e-10)
    # J = 0 at unit filling: the Mott state |1,...,1>, rho = identity, coherence 0, condensate fraction 1/m
    m = n = 5
    states, _, _ = O.basis(m, n)
    phi = np.array([1.0 if np.all(s == 1) else 0.0 for s in states])
    rho = O.spdm(states, phi)
    assert np.allclose(rho, np.eye(m)) and O.coherence(rho) == 0.0
    assert O.condensate_fraction(rho) == pytest.approx(1.0 / m)


# ---- mean-field single-site problem (Analysis::SCMF) --------------------------------------------------------

def mott_lobe_jc(mu):
    """Mean-field Mott-lobe boundary for U = 1 (Fisher et al. 1989; van Oosten et al. 2001; the report's
    eq. 10): in lobe g (g - 1 < mu < g), qJ_c = (g - mu)(mu - g + 1) / (mu + 1)"""
    g = np.floor(mu) + 1.0
    return (g - mu) * (mu - g + 1.0) / (mu + 1.0)


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("p", [1, 2, 5, 12])
def test_mf_ground_state_of_reference_h_mf(p):
    """the oracle's ground state of h_MF against LAPACK on the matrix BH::h_MF fills"""
    R = O.ref()
    rng = np.random.default_rng(p)
    for _ in range(20):
        psi, mu, J = rng.uniform(0, 2), rng.uniform(0, 4), rng.uniform(0, 0.3)
        N = 2 * p + 1
        h = np.zeros(N * N)
        R.ref_h_mf(psi, p, mu, J, 1, h)
        h = h.reshape(N, N)
        w, v = np.linalg.eigh(h)
        e, sf = O.mf_ground(psi, p, mu, J, 1)
        assert e == pytest.approx(w[0], rel=1e-13, abs=1e-13)
        # <a> (Hellmann-Feynman) against the reference's SF_density of LAPACK's ground state
        want = R.ref_sf_density(np.ascontiguousarray(v[:, 0]), p)
        assert sf == pytest.approx(want, rel=1e-11, abs=1e-13)
        # a start below the eigenvalue, and one above it (rejected), give the same pair
        for hint in (w[0] - 1e-3 * (abs(w[0]) + 1), w[0] + 0.5, w[0]):
            eh, sfh = O.mf_ground(psi, p, mu, J, 1, hint)
            assert eh == pytest.approx(w[0], rel=1e-13, abs=1e-13)
            assert sfh == pytest.approx(want, rel=1e-11, abs=1e-13)


def test_mf_ground_state_closed_forms():
    # psi = 0: h is diagonal, the ground state the lowest -mu n + n(n-1)/2, <a> = 0
    for mu in (0.3, 1.5, 2.7):
        e, sf = O.mf_ground(0.0, 4, mu, 0.1)
        nstar = int(np.floor(mu)) + 1 if mu > 0 else 0
        assert e == pytest.approx(-mu * nstar + nstar * (nstar - 1) / 2, abs=1e-15) and sf == 0.0
    # mu = -1/2, weak coupling c = -qJ psi from |0> to |1> (gap 1/2): second order E0 = qJ psi^2 - c^2 / (1/2)
    # and first order <a> = phi_0 phi_1 = -c / (1/2)
    c = 1e-3
    e, sf = O.mf_ground(1.0, 6, -0.5, c)
    assert e == pytest.approx(c - c * c / 0.5, rel=1e-6)
    assert sf == pytest.approx(c / 0.5, rel=1e-5)


@pytest.mark.skipif(not O.ref_available(), reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("precision,q", [(5, 1), (6, 1), (8, 1), (6, 4)])
def test_scmf_matches_reference_build(precision, q):
    """psi of Analysis::SCMF (Spectra ground states) against the oracle at random points; J > 0 so that the
    ground state is not degenerate"""
    R = O.ref()
    rng = np.random.default_rng(precision + 10 * q)
    for _ in range(25):
        mu, J, psi0 = rng.uniform(0, 4), rng.uniform(0.002, 0.25) / q, rng.uniform(0.01, 1)
        st, psi, it, p = O.scmf(mu, J, q, psi0, 10.0 ** -precision)
        assert st == O.MF_CONVERGED
        assert R.ref_scmf(mu, J, q, psi0, precision) == pytest.approx(psi, abs=1e-8), (mu, J, psi0)


def test_scmf_mott_lobes():
    """psi vanishes inside the mean-field Mott lobes and not outside them (closed-form boundary)"""
    eps = 1e-6
    rng = np.random.default_rng(11)
    inside = outside = 0
    for _ in range(300):
        mu, J = rng.uniform(0.05, 3.95), rng.uniform(0.0, 0.25)
        if abs(mu - round(mu)) < 0.05:
            continue
        jc = mott_lobe_jc(mu)
        st, psi, it, p = O.scmf(mu, J, 1, rng.uniform(0.05, 1.0), eps)
        assert st == O.MF_CONVERGED
        if J < 0.8 * jc:
            inside += 1
            assert abs(psi) <= 10 * eps, (mu, J, psi)
        elif J > 1.25 * jc:
            outside += 1
            assert psi > 1e-3, (mu, J, psi)
    assert inside > 20 and outside > 20


def test_scmf_atomic_limit_and_caps():
    # J = 0: h is diagonal, psi = 0 after one self-consistency step for a non-integer mu
    st, psi, it, p = O.scmf(1.5, 0.0, 1, 0.7, 1e-6)
    assert (st, psi, it) == (O.MF_CONVERGED, 0.0, 2)
    # caps: one outer iteration, or p <= 2, are reported
    assert O.scmf(0.5, 0.2, 1, 0.3, 1e-10, max_outer=1)[0] == O.MF_MAX_OUTER
    assert O.scmf(0.5, 0.2, 1, 0.3, 1e-12, pmax=2)[0] == O.MF_MAX_P


def test_mf_start_values_splitmix64():
    # SplitMix64 with seed 0: first outputs 0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4 (published test vector)
    z = O.mf_psi0(0, [0, 1])
    assert int(z[0] * 2.0 ** 53) == 0xE220A8397B1DCDAF >> 11
    assert int(z[1] * 2.0 ** 53) == 0x6E789E6AA1B965F4 >> 11


@pytest.mark.parametrize("m,n,J,U,k", [(8, 8, 1.0, 1.0, 20), (8, 8, 0.1, 1.0, 25), (6, 6, 0.3, 1.0, 12),
                                       (7, 5, 0.0, 1.0, 8), (8, 8, 0.5, 0.0, 15)])
def test_oracle_lanczos_mult_matches_dense(m, n, J, U, k):
    """The oracle's eigenvalues with multiplicity against LAPACK on the same H: degenerate pairs of the periodic
    chain, the atomic limit (J = 0: U with multiplicity m (m - 1) at unit filling) and free bosons (U = 0)."""
    h = O.hamiltonian(m, n, *O.chain_neighbours(m))
    w = np.linalg.eigvalsh(h.dense(J, U))
    rc, ev, rs, _ = O.lanczos_mult(h, J, U, k)
    assert rc == 0
    assert np.allclose(ev, w[:k], rtol=0, atol=1e-9 * max(1.0, abs(w[0])))


def test_oracle_lanczos_mult_free_bosons_closed_form():
    """U = 0 on the periodic chain: many-body values are sums of single-particle energies -2J cos(2 pi q / m);
    the lowest k with their multiplicity, at a size dense eigh does not reach (m = 10, n = 10: D = 92378)."""
    m, n, J, k = 10, 10, 0.5, 6
    h = O.hamiltonian(m, n, *O.chain_neighbours(m))
    eps = -2 * J * np.cos(2 * np.pi * np.arange(m) / m)
    want = np.sort(free_boson_spectrum(eps, n))[:k]
    rc, ev, rs, _ = O.lanczos_mult(h, J, 0.0, k)
    assert rc == 0 and np.allclose(ev, want, rtol=0, atol=1e-9)


def test_start_vector_continuation():
    """The pass start vectors continue the Park-Miller sequence of the first one."""
    r = O.start_vector(40, 7)
    assert np.array_equal(O.start_vector_at(25, 7, 15), r[15:])
