"""Golden fixtures written by tests/golden/make_golden.py from the reference's own compiled functions.

CPU: the oracle against the fixtures (basis order and tags, neighbour lists, reference-convention
diagonal terms), and the fixture spectra of the reference's Spectra solver against the dense spectrum of
the oracle's operator.  GPU: the device basis, diagonal and eigensolver against the same fixtures."""
import json
import os

import numpy as np
import pytest

import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BASES = [(3, 4), (5, 5), (6, 6), (8, 8)]


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.mark.parametrize("m,n", BASES)
def test_oracle_basis_matches_reference_fixture(m, n):
    g = load(f"basis_m{m}_n{n}.npz")
    st, tags, _ = O.basis(m, n, O.ORDER_REFERENCE)
    assert np.array_equal(st, g["states"])
    assert np.allclose(tags, g["tags"], rtol=0, atol=1e-13)


def test_oracle_neighbours_match_reference_fixture():
    nb = json.load(open(os.path.join(GOLD, "neighbours.json")))
    for key, lists in nb["chain"].items():
        m, closed = map(int, key.split("_"))
        off, sites = O.chain_neighbours(m, bool(closed))
        assert [sites[off[i]:off[i + 1]].tolist() for i in range(m)] == lists
    # the reference's square lattice: empty lists for a perfect square (shadowing local,
    # src/neighbours.cpp:45), an exception otherwise (:41-44); the oracle fills the lists instead
    assert nb["square"]["4_1"] == [[], [], [], []] and nb["square"]["6_1"] is None
    off, sites = O.square_neighbours(2, 2, True)
    assert off[-1] > 0


@pytest.mark.parametrize("m,n", BASES)
def test_oracle_diagonals_match_reference_fixture(m, n):
    g = load(f"diag_m{m}_n{n}.npz")
    st, _, _ = O.basis(m, n, O.ORDER_REFERENCE)
    off, sites = O.chain_neighbours(m, True)
    h = O.hamiltonian(m, n, off, sites, flags=O.CONV_REF_INTERACTION)
    assert np.array_equal(h.diag, g["interaction"])
    assert np.all(g["chemical"] == -float(n))
    # the canonical interaction U/2 sum n(n-1) is the reference's sum n(n+1) less 2N, halved
    hc = O.hamiltonian(m, n, off, sites)
    assert np.allclose(hc.diag, (g["interaction"] - 2 * n) / 2, rtol=0, atol=1e-12)


@pytest.mark.parametrize("key,J,U", [("J1.0_U1.0", 1.0, 1.0), ("J0.1_U1.0", 0.1, 1.0)])
def test_reference_spectra_fixture_matches_dense(key, J, U):
    ev = load("spectra_m8_n8.npz")[key]
    off, sites = O.chain_neighbours(8, True)
    h = O.hamiltonian(8, 8, off, sites)
    dense = np.linalg.eigvalsh(h.dense(J, U))[:10]
    assert np.allclose(ev, dense, rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
@pytest.mark.parametrize("m,n", BASES)
def test_device_basis_and_diagonal_match_reference_fixture(m, n):
    import bhx

    print(f"golden device m={m} n={n}", flush=True)
    g = load(f"basis_m{m}_n{n}.npz")
    gd = load(f"diag_m{m}_n{n}.npz")
    ctx = bhx.Context(0)
    b = ctx.basis(m, n, bhx.ORDER_REFERENCE)
    assert np.array_equal(b.states(), g["states"])
    assert np.array_equal(b.tags(), g["tags"])
    off, sites = O.chain_neighbours(m, True)
    h = b.hamiltonian(off, sites, flags=bhx.CONV_REF_INTERACTION)
    assert np.array_equal(h.csr()[3], gd["interaction"])
    h.close()
    b.close()
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("key,J,U", [("J1.0_U1.0", 1.0, 1.0), ("J0.1_U1.0", 0.1, 1.0)])
def test_device_lanczos_matches_reference_spectra_fixture(key, J, U):
    import bhx

    print(f"golden spectra {key}", flush=True)
    want = load("spectra_m8_n8.npz")[key]  # the reference's 10 values, degenerate pairs included
    ctx = bhx.Context(0)
    b = ctx.basis(8, 8, bhx.ORDER_REFERENCE)
    off, sites = O.chain_neighbours(8, True)
    h = b.hamiltonian(off, sites)
    for prec in (bhx.PREC_FP64, bhx.PREC_MIXED):
        r = h.lanczos(J, U, k=10, prec=prec)
        assert r.status[0] == 0
        assert np.all(np.abs(r.evals[0] - want) <= 1e-9 * np.maximum(np.abs(want), 1.0)), (r.evals[0], want)
        assert O.gap_ratio(r.evals[0]) == pytest.approx(O.gap_ratio(want), abs=1e-6)
    h.close()
    b.close()
    ctx.close()
