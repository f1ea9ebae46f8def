"""Writes the golden fixtures of tests/golden/ from the reference's own code.

The reference's functions are compiled from its sources into oracle/_ref/libbhref.so (oracle/Makefile,
target `ref`; needs the reference tree) and called through oracle/ref_driver.cpp.  Only what the reference
computes deterministically is recorded (its hopping matrix is not: see the oracle's header):

  basis_m{m}_n{n}.npz     BH::fixed_set_basis (src/hamiltonian.cpp:143-149): states[D, m] (uint8, row k =
                          the reference's basis column k) and tags[D] in the reference's order
  diag_m{m}_n{n}.npz      diagonals of BH::fixed_bosons_hamiltonian with (J, U, mu) = (0, 1, 0), the
                          interaction term U sum n(n+1) (:194-211), and (0, 0, 1), the chemical term -mu N
                          (:214-232)
  neighbours.json         Neighbours::chain_neighbours (src/neighbours.cpp:21-34), open and periodic,
                          m = 2..9, and whether square_neighbours throws / what it returns (:40-71)
  maxbasis_m{m}_n{n}.npz  BH::max_set_basis (src/hamiltonian.cpp:151-166): states[D, m] and tags[D] of every
                          boson number 1..n after the reference's sort
  spectra_m8_n8.npz       Op::IRLM_eigen (src/operator.cpp:22-33, Spectra GenEigs nev=10) run on the
                          intended operator J T + U V of the chain m=8, n=8 at (J, U) = (1, 1) and (0.1, 1)

Run from the repository root:  python tests/golden/make_golden.py
"""
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

BASES = [(3, 4), (5, 5), (6, 6), (8, 8)]
MAX_BASES = [(3, 3), (4, 3), (5, 2)]


def main():
    R = O.ref()
    if R is None:
        raise SystemExit("oracle/_ref/libbhref.so is not built (make -C oracle ref)")
    for m, n in BASES:
        D = O.dimension(m, n)
        tags = np.zeros(D)
        st = np.zeros(D * m)
        assert R.ref_fixed_set_basis(m, n, tags, st) == D
        np.savez_compressed(os.path.join(HERE, f"basis_m{m}_n{n}.npz"), states=st.reshape(D, m).astype(np.uint8),
                            tags=tags)
        dU = np.zeros(D)
        dmu = np.zeros(D)
        assert R.ref_diagonal_term(m, n, 1, dU) == D
        assert R.ref_diagonal_term(m, n, 2, dmu) == D
        np.savez_compressed(os.path.join(HERE, f"diag_m{m}_n{n}.npz"), interaction=dU, chemical=dmu)
    for m, n in MAX_BASES:
        D = sum(O.dimension(m, k) for k in range(1, n + 1))
        tags = np.zeros(D)
        st = np.zeros(D * m)
        assert R.ref_max_set_basis(m, n, tags, st) == D
        np.savez_compressed(os.path.join(HERE, f"maxbasis_m{m}_n{n}.npz"), states=st.reshape(D, m).astype(np.uint8),
                            tags=tags)
    nb = {"chain": {}, "square": {}}
    for m in range(2, 10):
        for closed in (0, 1):
            off = np.zeros(m + 1, np.int32)
            sites = np.zeros(4 * m + 4, np.int32)
            ne = R.ref_chain_neighbours(m, closed, off, sites)
            nb["chain"][f"{m}_{closed}"] = [sites[off[i]:off[i + 1]].tolist() for i in range(m)]
            assert ne == off[m]
    for m in (4, 6, 9):
        for closed in (0, 1):
            off = np.zeros(m + 1, np.int32)
            sites = np.zeros(4 * m + 4, np.int32)
            ne = R.ref_square_neighbours(m, closed, off, sites)
            nb["square"][f"{m}_{closed}"] = None if ne < 0 else [sites[off[i]:off[i + 1]].tolist() for i in range(m)]
    with open(os.path.join(HERE, "neighbours.json"), "w") as f:
        json.dump(nb, f, indent=0)
    off, sites = O.chain_neighbours(8, True)
    h = O.hamiltonian(8, 8, off, sites)
    out = {}
    for J, U in ((1.0, 1.0), (0.1, 1.0)):
        ev = np.zeros(10)
        assert R.ref_irlm_eigen(h.D, h.row_ptr, h.col, h.val, h.diag, J, U, 10, ev) == 0
        out[f"J{J}_U{U}"] = ev
    np.savez_compressed(os.path.join(HERE, "spectra_m8_n8.npz"), **out)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
