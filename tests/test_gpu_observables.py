"""Single-particle density matrix on the device (bh_spdm) against the oracle's, and the observables the
sweep writes from it (condensate fraction, coherence).  Random real states are compared entry by entry
within 1e-12 of the largest entry (FP64 sums in a different order); at full size the ground state of the
periodic chain is checked through properties that do not depend on the size: symmetry, trace n, and
translation invariance rho_ij = rho_0,(j-i) mod m."""
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


@pytest.mark.parametrize("order", [bhx.ORDER_REFERENCE, bhx.ORDER_LEX])
@pytest.mark.parametrize("m,n", [(8, 8), (9, 6), (20, 3), (5, 16), (4, 30), (1, 5)])
def test_spdm_matches_oracle(ctx, m, n, order):
    print(f"spdm m={m} n={n} order={order}", flush=True)
    b = ctx.basis(m, n, order)
    states = b.states()
    phi = np.random.default_rng(m * 100 + n).standard_normal(b.D)
    phi /= np.linalg.norm(phi)
    got = b.spdm(phi)
    want = O.spdm(states, phi)
    assert np.allclose(got, want, rtol=0, atol=1e-12 * np.abs(want).max())
    assert np.trace(got) == pytest.approx(n, rel=1e-12)
    b.close()


def test_spdm_rejects_wrong_length(ctx):
    b = ctx.basis(5, 3)
    with pytest.raises(bhx.BHError):
        b.spdm(np.ones(b.D + 1))
    b.close()


def test_spdm_ground_state_full_size(ctx):
    print("spdm chain m=12 n=12", flush=True)
    m, n = 12, 12
    off, sites = bhx.lattice("chain", m=m, periodic=True)
    b = ctx.basis(m, n, bhx.ORDER_LEX)
    h = b.hamiltonian(off, sites)
    r = h.lanczos(0.1, 1.0, k=2, prec=bhx.PREC_FP64, want_vectors=1)
    assert r.status[0] == 0
    phi = h.lanczos_vectors(1, 2)[0, 0]
    rho = b.spdm(phi)
    assert np.allclose(rho, rho.T, atol=1e-12)
    assert np.trace(rho) == pytest.approx(n, rel=1e-9)
    for i in range(m):
        assert np.allclose(rho[i], np.roll(rho[0], i), atol=1e-8)
    cf, K = O.condensate_fraction(rho), O.coherence(rho)
    assert 1.0 / m < cf < 1.0 and 0.0 < K < 1.0
    h.close()
    b.close()
