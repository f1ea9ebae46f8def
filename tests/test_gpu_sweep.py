"""GPU phase-diagram sweep (sectors n-1, n, n+1 solved as batches over J; mu as the shift -mu N)
against the oracle's dense spectra and atomic-limit / free-boson closed forms."""
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


def test_sweep_matches_dense(ctx):
    m, n, U = 6, 6, 1.0
    J = np.array([0.0, 0.03, 0.08, 0.2, 0.4])
    mu = np.linspace(-0.5, 2.0, 11)
    res = ctx.sweep("chain", n, J, U, mu, m=m, k=4)
    off, sites = O.chain_neighbours(m, True)
    E0 = np.zeros((3, J.size))
    for q in range(3):
        h = O.hamiltonian(m, n - 1 + q, off, sites)
        for i, Jv in enumerate(J):
            lv = np.linalg.eigvalsh(h.dense(Jv, U))[:4]  # with multiplicity
            E0[q, i] = lv[0]
            assert abs(res["E"][q, i, 0] - lv[0]) <= 1e-9 * max(1.0, abs(lv[0]))
            assert np.allclose(res["E"][q, i], lv, rtol=1e-5, atol=1e-9)
            # gap ratio of the returned values by the reference's formula; a degenerate pair gives a zero gap
            assert res["gap_ratio"][q, i] == pytest.approx(O.gap_ratio(res["E"][q, i]), abs=1e-12)
            assert res["gap_ratio"][q, i] == pytest.approx(O.gap_ratio(lv), abs=1e-3)
    # density: argmin over sectors of E0(N) - mu N; compressibility on the central plateau
    for i in range(J.size):
        mup, mum = E0[2, i] - E0[1, i], E0[1, i] - E0[0, i]
        assert res["mu_plus"][i] == pytest.approx(mup, abs=1e-8)
        assert res["mu_minus"][i] == pytest.approx(mum, abs=1e-8)
        for c, mv in enumerate(mu):
            om = [E0[q, i] - mv * (n - 1 + q) for q in range(3)]
            q = int(np.argmin(om))
            assert res["density"][i, c] == pytest.approx((n - 1 + q) / m)
            if q == 1:
                assert res["compressibility"][i, c] == pytest.approx(1.0 / (m * (mup - mum)), rel=1e-6)
            else:
                assert np.isnan(res["compressibility"][i, c])


def test_sweep_atomic_limit_and_free_bosons(ctx):
    m, n = 8, 8
    mu = np.array([-0.3, 0.5, 1.5])
    res = ctx.sweep("chain", n, [0.0], 1.0, mu, m=m, k=4)
    # J = 0, unit filling: E0(n) = 0, E0(n+1) = U, E0(n-1) = 0 -> Mott plateau mu in (0, U)
    assert res["mu_plus"][0] == pytest.approx(1.0) and res["mu_minus"][0] == pytest.approx(0.0)
    assert np.allclose(res["density"][0], [(n - 1) / m, 1.0, (n + 1) / m])
    assert res["compressibility"][0, 1] == pytest.approx(1.0 / m)
    # U = 0: E0(N) = -2 J N, no charge gap (infinite compressibility at the degenerate point)
    res = ctx.sweep("chain", n, [0.5], 0.0, [-1.0], m=m, k=2, prec=bhx.PREC_FP64)
    for q in range(3):
        assert res["E"][q, 0, 0] == pytest.approx(-2 * 0.5 * (n - 1 + q), abs=1e-9)
    assert res["mu_plus"][0] == pytest.approx(-1.0) and res["mu_minus"][0] == pytest.approx(-1.0)


def test_sweep_full_size(ctx):
    """The batched sweep configuration: chain m=10, sectors 9/10/11, 32 log-spaced J/U in [0.01, 0.4],
    64 mu/U in [0, 3]; E0 of sampled points equals individual FP64 solves within 1e-9."""
    m, n = 10, 10
    J = 0.01 * 40.0 ** (np.arange(32) / 31.0)
    mu = 3.0 * np.arange(64) / 63.0
    res = ctx.sweep("chain", n, J, 1.0, mu, m=m, k=4)
    assert np.all(res["status"] == 0)
    assert all(w in (8, 16, 32) for w in res["width"])  # block width chosen per sector by the tuner
    off, sites = O.chain_neighbours(m, True)
    for q in (0, 2):
        b = ctx.basis(m, n - 1 + q, bhx.ORDER_LEX)
        h = b.hamiltonian(off, sites)
        for i in (0, 15, 31):
            r = h.lanczos(J[i], 1.0, k=4, prec=bhx.PREC_FP64)
            assert abs(res["E"][q, i, 0] - r.evals[0, 0]) <= 1e-9 * abs(r.evals[0, 0])
        h.close()
        b.close()
    # Mott plateau at small J: density exactly 1 for mu inside (mu_minus, mu_plus)
    inside = (mu > res["mu_minus"][0]) & (mu < res["mu_plus"][0])
    assert inside.sum() > 10 and np.all(res["density"][0, inside] == 1.0)
    # the charge gap closes as J/U grows (1D Mott transition near J/U ~ 0.3)
    gap = res["mu_plus"] - res["mu_minus"]
    assert gap[0] > 0.9 and gap[-1] < gap[0] / 3


def test_sweep_fixed_width_equals_tuned(ctx):
    """Batches of a fixed width 8 give the levels of the tuned width (E0 within 1e-9, the levels above
    within 1e-5: FP32-stored solves) and report the width used per sector."""
    m, n = 10, 10
    J = 0.01 * 40.0 ** (np.arange(20) / 19.0)
    mu = np.array([0.5, 1.5])
    a = ctx.sweep("chain", n, J, 1.0, mu, m=m, k=4)
    b = ctx.sweep("chain", n, J, 1.0, mu, m=m, k=4, width=8)
    assert list(b["width"]) == [8, 8, 8] and all(w in (8, 16, 32) for w in a["width"])
    assert np.all(a["status"] == 0) and np.all(b["status"] == 0)
    e0a, e0b = a["E"][:, :, 0], b["E"][:, :, 0]
    assert np.all(np.abs(e0a - e0b) <= 1e-9 * np.abs(e0b))
    assert np.allclose(a["E"], b["E"], rtol=1e-5, atol=1e-9)
    with pytest.raises(bhx.BHError):
        ctx.sweep("chain", n, J, 1.0, mu, m=m, k=4, width=12)
