"""bh_spmv_noalloc, the SpMV entry point that allocates nothing (its per-lane J, U go to a library-owned
device array), against bh_spmv on the same device buffers: bit-identical when both use the same
variant, on two calls in a row with the second on a smaller sector and fewer lanes (the array keeps the
first call's contents), in both precisions; within rounding of bh_spmv when the handle was never tuned;
and the host-side refusals (B beyond the array, FP64 before bh_ham_prepare)."""
import ctypes as C

import numpy as np
import pytest

import bhx
import oracle as O

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """Device buffers for the test from the HIP runtime libbhx.so itself is linked to."""

    def __init__(self):
        bhx.lib()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), a.nbytes) == 0
        self.bufs.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D) == 0
        return p.value

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(p), out.nbytes, D2H) == 0
        return out

    def close(self):
        for p in self.bufs:
            self.L.hipFree(p)
        self.bufs = []


@pytest.fixture(scope="module")
def env():
    ctx = bhx.Context(0)
    hip = Hip()
    yield ctx, hip
    hip.close()
    ctx.close()


def sector(ctx, m, n):
    off, sites = O.chain_neighbours(m, True)
    b = ctx.basis(m, n)
    return b, b.hamiltonian(off, sites)


def case(hip, h, B, prec, seed):
    rng = np.random.default_rng(seed)
    dt = np.float64 if prec == bhx.PREC_FP64 else np.float32
    x = rng.standard_normal((h.D, B)).astype(dt)
    J = rng.uniform(0.02, 0.3, B)
    U = rng.uniform(0.5, 1.5, B)
    nan = np.full_like(x, np.nan)
    return dict(x=x, J=J, U=U, B=B, prec=prec, px=hip.upload(x), pref=hip.upload(nan), pout=hip.upload(nan))


@pytest.mark.parametrize("prec", [bhx.PREC_FP64, bhx.PREC_MIXED])
def test_noalloc_matches_spmv_two_calls(env, prec):
    ctx, hip = env
    big_b, big = sector(ctx, 10, 10)    # D = 92378
    small_b, small = sector(ctx, 8, 6)  # D = 1716
    cases = [(big, case(hip, big, 8, prec, 1)), (small, case(hip, small, 2, prec, 2))]
    for h, c in cases:  # bh_spmv first: it prepares the handle and tunes the variant for (prec, B)
        h.spmv_device(c["J"], c["U"], c["px"], c["pref"], c["B"], prec)
    for h, c in cases:  # then the two entry-point calls back to back, the second smaller
        print(f"noalloc prec={prec} D={h.D} B={c['B']}", flush=True)
        h.spmv_device(c["J"], c["U"], c["px"], c["pout"], c["B"], prec, noalloc=True)
    for h, c in cases:
        ref = hip.download(c["pref"], c["x"])
        out = hip.download(c["pout"], c["x"])
        assert np.all(np.isfinite(ref))
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8)), (h.D, c["B"])
        # the small sector's product against the oracle's
        if h.D < 5000:
            off, sites = O.chain_neighbours(8, True)
            ho = O.hamiltonian(8, 6, off, sites)
            x = c["x"].astype(np.float64)
            want = np.stack([O.spmv(ho, c["J"][b], c["U"][b], x[:, b].copy()) for b in range(c["B"])], axis=1)
            tol = 1e-13 if prec == bhx.PREC_FP64 else 2e-6
            assert np.allclose(out, want, rtol=tol, atol=tol * np.abs(want).max())
    for b_, h in ((big_b, big), (small_b, small)):
        h.close()
        b_.close()


def test_noalloc_untuned_handle(env):
    ctx, hip = env
    b, h = sector(ctx, 9, 7)  # D = 6435, never applied before: the default variant
    h.prepare(bhx.PREC_FP64)
    c = case(hip, h, 4, bhx.PREC_FP64, 3)
    h.spmv_device(c["J"], c["U"], c["px"], c["pout"], 4, bhx.PREC_FP64, noalloc=True)
    h.spmv_device(c["J"], c["U"], c["px"], c["pref"], 4, bhx.PREC_FP64)
    ref = hip.download(c["pref"], c["x"])
    out = hip.download(c["pout"], c["x"])
    assert np.max(np.abs(out - ref)) <= 1e-13 * np.max(np.abs(ref))
    h.close()
    b.close()


def test_noalloc_refusals(env):
    ctx, hip = env
    b, h = sector(ctx, 6, 4)
    x = np.zeros((h.D, 1))
    px, py = hip.upload(x), hip.upload(x)
    with pytest.raises(bhx.BHError, match="bh_ham_prepare"):  # FP64 CSR not materialised yet
        h.spmv_device(1.0, 1.0, px, py, 1, bhx.PREC_FP64, noalloc=True)
    with pytest.raises(bhx.BHError, match="power of two"):
        h.spmv_device(1.0, 1.0, px, py, 128, bhx.PREC_MIXED, noalloc=True)
    with pytest.raises(bhx.BHError, match="power of two"):
        h.spmv_device(1.0, 1.0, px, py, 3, bhx.PREC_MIXED, noalloc=True)
    h.close()
    b.close()
