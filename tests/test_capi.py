"""CPU checks of the C ABI: libbhx.so loads, exports every function include/bh_capi.h declares, and
the host-only entry points agree with the oracle.  No device compute."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import bhx
import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            txt = open(os.path.join(inc, f)).read()
            names |= set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(bh_[a-z_0-9]+)\s*\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(bhx.LIB_PATH)
    names = declared_functions()
    assert len(names) > 20
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


@pytest.mark.parametrize("m,per", [(2, True), (3, False), (8, True), (12, True), (31, False)])
def test_chain_neighbours_match_oracle(m, per):
    off, s = bhx.lattice("chain", m=m, periodic=per)
    o2, s2 = O.chain_neighbours(m, per)
    assert np.array_equal(off, o2) and np.array_equal(s, s2)


@pytest.mark.parametrize("Lx,Ly,per", [(4, 4, True), (3, 4, True), (4, 3, False), (2, 2, True), (3, 3, True)])
def test_square_neighbours_match_oracle(Lx, Ly, per):
    off, s = bhx.lattice("square", Lx=Lx, Ly=Ly, periodic=per)
    o2, s2 = O.square_neighbours(Lx, Ly, per)
    assert np.array_equal(off, o2) and np.array_equal(s, s2)


def test_invalid_lattice_rejected():
    with pytest.raises(bhx.BHError):
        bhx.lattice("chain", m=0)
    with pytest.raises(bhx.BHError):
        bhx.lattice("square", Lx=8, Ly=8)  # more than 32 sites


def test_tridiag_eigvals_host_entry():
    L = bhx.lib()
    if not hasattr(L, "bh_tridiag_eigvals"):
        pytest.skip("not built")
    rng = np.random.default_rng(3)
    for n in (1, 2, 7, 60):
        d = rng.standard_normal(n)
        e = rng.standard_normal(max(n - 1, 1))
        out = np.zeros(n)
        assert L.bh_tridiag_eigvals(n, d.ctypes.data, e.ctypes.data, out.ctypes.data) == 0
        T = np.diag(d) + (np.diag(e[: n - 1], 1) + np.diag(e[: n - 1], -1) if n > 1 else 0)
        assert np.allclose(out, np.linalg.eigvalsh(T), atol=1e-12)
