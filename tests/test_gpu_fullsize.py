"""The benchmark sectors at full size: chain m=12 n=12 (D = 1,352,078, nnz 16.9M) and the 4x4 torus with
n=10 (D = 3,268,760, nnz 83.7M).  Basis and CSR bit-exact against the oracle's (the reference's basis
order, the intended hopping operator), and the 4x4 mixed-precision solve against the FP64 GPU solve and the
oracle's FP64 Lanczos (E0 within 1e-9 relative; the other three values, with their multiplicity, within 1e-5
of the FP64 run, 1e-9 when the mixed lane was rerun in FP64)."""
import math

import numpy as np
import pytest

import bhx
import oracle as O

pytestmark = pytest.mark.gpu

SECTORS = {"chain12": ("chain", 12, 12), "square4x4": ("square", 16, 10)}


def lattice(kind, m):
    return O.chain_neighbours(m, True) if kind == "chain" else O.square_neighbours(4, 4, True)


@pytest.fixture(scope="module")
def ctx():
    c = bhx.Context(0)
    yield c
    c.close()


@pytest.fixture(scope="module")
def oracle_sectors():
    cache = {}

    def get(name):
        if name not in cache:
            kind, m, n = SECTORS[name]
            off, sites = lattice(kind, m)
            print(f"oracle build {name}", flush=True)
            st, _, _ = O.basis(m, n)
            cache[name] = (st, O.hamiltonian(m, n, off, sites, states=st))
        return cache[name]

    return get


@pytest.mark.parametrize("name", list(SECTORS))
def test_full_size_basis_and_csr_bit_exact(ctx, oracle_sectors, name):
    kind, m, n = SECTORS[name]
    off, sites = lattice(kind, m)
    st, ho = oracle_sectors(name)
    print(f"full-size bit-exact {name} D={ho.D} nnz={ho.nnz}", flush=True)
    b = ctx.basis(m, n)
    assert b.D == math.comb(m + n - 1, n) == ho.D
    assert np.array_equal(b.states(), st)
    h = b.hamiltonian(off, sites)
    rp, col, val, diag = h.csr()
    assert h.nnz == ho.nnz
    assert np.array_equal(rp, ho.row_ptr)
    assert np.array_equal(col, ho.col)
    assert np.array_equal(val, ho.val)
    assert np.array_equal(diag, ho.diag)
    h.close()
    b.close()


@pytest.mark.parametrize("J", [0.05, 0.2])
def test_square4x4_mixed_solve(ctx, oracle_sectors, J):
    kind, m, n = SECTORS["square4x4"]
    off, sites = lattice(kind, m)
    print(f"4x4 n=10 solve J={J}", flush=True)
    b = ctx.basis(m, n, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    r32 = h.lanczos(J, 1.0, k=4, prec=bhx.PREC_MIXED)
    r64 = h.lanczos(J, 1.0, k=4, prec=bhx.PREC_FP64)
    print("mixed", r32.evals[0].tolist(), r32.iters, r32.used_fp64, "fp64", r64.evals[0].tolist(), r64.iters, flush=True)
    assert r32.status[0] == 0 and r64.status[0] == 0
    e = np.abs(r32.evals[0] - r64.evals[0]) / np.abs(r64.evals[0])
    assert e[0] <= 1e-9
    assert np.all(e <= (1e-9 if r32.used_fp64[0] else 1e-5)), e
    # the oracle's FP64 Lanczos (full reorthogonalisation) for E0, on its own CSR in the reference order
    _, ho = oracle_sectors("square4x4")
    rc, ev, _, it = O.lanczos(ho, J, 1.0, 1, max_iter=400)
    assert rc == 0
    assert abs(r32.evals[0, 0] - ev[0]) <= 1e-9 * abs(ev[0])
    h.close()
    b.close()
