"""Batched self-consistent mean-field solves on the device (bh_mean_field) against the oracle.

The kernel and the oracle evaluate the same sequence of correctly rounded IEEE operations (no contraction,
the sqrt(i) table from the host), so psi, the status, the number of self-consistency iterations and the last
truncation p must agree exactly at every point, degenerate ones included.  At the reference's full size (a
1000 x 1000 grid at precision 1e-5) every point is checked against the closed-form Mott-lobe boundary and a
sample against the oracle."""
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


def grid(n):
    def rng(lo, hi, d):
        out, x = [], lo
        while x < hi:
            out.append(x)
            x += d
        return out
    mus, Js = rng(0.0, 4.0, 4.0 / n), rng(0.0, 0.25, 0.25 / n)
    return np.repeat(mus, len(Js)), np.tile(Js, len(mus))


def oracle_all(mu, J, psi0, eps, **kw):
    out = np.array([O.scmf(mu[k], J[k], 1, psi0[k], eps, **kw) for k in range(mu.size)], dtype=object)
    return (out[:, 0].astype(np.int32), out[:, 1].astype(np.float64), out[:, 2].astype(np.int32),
            out[:, 3].astype(np.int32))


@pytest.mark.parametrize("n,prec", [(40, 6), (30, 8)])
def test_scmf_grid_bitwise(ctx, n, prec):
    print(f"scmf grid n={n} prec={prec}", flush=True)
    mu, J = grid(n)
    psi0 = O.mf_psi0(5, np.arange(mu.size))
    r = ctx.mean_field(mu, J, psi0, eps=10.0 ** -prec)
    st, psi, it, p = oracle_all(mu, J, psi0, 10.0 ** -prec)
    assert np.array_equal(r["status"], st)
    assert np.array_equal(r["iters"], it) and np.array_equal(r["p"], p)
    assert np.array_equal(r["psi"], psi)


@pytest.mark.parametrize("q", [2, 4, 6])
def test_scmf_coordination_bitwise(ctx, q):
    """q neighbours per site (square and cubic lattices): the lobes shrink to qJ_c; still bit for bit"""
    print(f"scmf q={q}", flush=True)
    rng = np.random.default_rng(q)
    mu = rng.uniform(0, 4, 600)
    J = rng.uniform(0, 0.25 / q, 600)
    psi0 = rng.uniform(0, 1, 600)
    r = ctx.mean_field(mu, J, psi0, q=q, eps=1e-7)
    out = np.array([O.scmf(mu[k], J[k], q, psi0[k], 1e-7) for k in range(mu.size)], dtype=object)
    assert np.array_equal(r["status"], out[:, 0].astype(np.int32))
    assert np.array_equal(r["psi"], out[:, 1].astype(np.float64))
    assert np.array_equal(r["iters"], out[:, 2].astype(np.int32))


def test_scmf_caps_and_edge_points(ctx):
    print("scmf caps", flush=True)
    rng = np.random.default_rng(2)
    mu = np.concatenate([rng.uniform(0, 4, 200), [0.0, 1.0, 2.0, 1.5, 0.5]])
    J = np.concatenate([rng.uniform(0, 0.25, 200), [0.0, 0.0, 0.1, 0.0, 0.2]])
    psi0 = np.concatenate([rng.uniform(0, 1, 200), [0.5, 0.0, 0.0, 0.9, 1.0]])
    for kw in ({"max_outer": 3}, {"pmax": 2, "eps": 1e-12}, {"pmax": 3}, {"eps": 0.0, "max_outer": 50}):
        eps = kw.pop("eps", 1e-6)
        r = ctx.mean_field(mu, J, psi0, eps=eps, **kw)
        st, psi, it, p = oracle_all(mu, J, psi0, eps, **kw)
        assert np.array_equal(r["status"], st) and np.array_equal(r["psi"], psi), kw
        assert np.array_equal(r["iters"], it) and np.array_equal(r["p"], p), kw
        assert (st != 0).any(), kw


def test_scmf_rejects_bad_arguments(ctx):
    one = np.ones(1)
    for kw in ({"pmax": 33}, {"pmax": 1}, {"max_outer": 0}, {"eps": -1.0}, {"eps": np.inf}):
        with pytest.raises(bhx.BHError):
            ctx.mean_field(one, one, one, **kw)
    with pytest.raises(ValueError):
        ctx.mean_field(np.ones(2), one, one)
    r = ctx.mean_field(np.zeros(0), np.zeros(0), np.zeros(0))
    assert r["psi"].size == 0


def test_scmf_full_size_phase_diagram(ctx):
    """the reference's published run: 1000 x 1000 points, precision 1e-5"""
    print("scmf full size 1000x1000", flush=True)
    eps = 1e-5
    mu, J = grid(1000)
    psi0 = O.mf_psi0(0, np.arange(mu.size))
    r = ctx.mean_field(mu, J, psi0, eps=eps)
    assert np.all(r["status"] == 0)
    # Mott lobes: psi ~ 0 well inside, clearly nonzero well outside (away from integer mu)
    g = np.floor(mu) + 1.0
    jc = (g - mu) * (mu - g + 1.0) / (mu + 1.0)
    off_int = np.abs(mu - np.round(mu)) > 0.05
    inside = off_int & (J < 0.8 * jc)
    outside = off_int & (J > 1.25 * jc)
    assert inside.sum() > 10000 and outside.sum() > 10000
    assert np.all(np.abs(r["psi"][inside]) <= 10 * eps)
    # outside, psi = 0 is an unstable fixed point; the reference's rule |psi' - psi| <= eps still stops on the
    # first step when the start value is that close to it (psi0 (qJ/qJ_c - 1) <= eps): 7 points of this grid
    early = psi0 < 1e-4
    assert np.all(r["psi"][outside & ~early] > 1e-3)
    assert np.all(r["iters"][outside & (r["psi"] <= 1e-3)] == 1)
    # psi grows with J at fixed mu in the superfluid (monotone up to the convergence threshold).  Near the lobe
    # boundary qJ/qJ_c - 1 is small and the same first-step stop leaves larger start values unchanged, so the
    # check takes pairs of neighbours that both lie well outside the lobes
    P = r["psi"].reshape(1000, 1000)
    W = (outside & ~early).reshape(1000, 1000)
    pair = W[:, 1:] & W[:, :-1]
    assert pair.sum() > 100000 and np.all(np.diff(P, axis=1)[pair] > -10 * eps)
    # a sample against the oracle, exactly
    idx = np.random.default_rng(9).choice(mu.size, 1500, replace=False)
    st, psi, it, p = oracle_all(mu[idx], J[idx], psi0[idx], eps)
    assert np.array_equal(r["psi"][idx], psi) and np.array_equal(r["iters"][idx], it)
