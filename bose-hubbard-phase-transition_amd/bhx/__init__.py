"""Python front end of libbhx.so (the gfx950 exact-diagonalisation engine), over its C ABI.

Mirrors the reference's surface for the exact path:
    Neighbours / lattice()         Neighbours::chain_neighbours, square_neighbours (src/neighbours.cpp:21-71)
    Basis  (fixed_set_basis)       BH::fixed_set_basis (src/hamiltonian.cpp:143-149)
    Hamiltonian                    BH::fixed_bosons_hamiltonian, J and U terms (src/hamiltonian.cpp:170-256)
    Hamiltonian.spmv               SparseGenMatProd::perform_op (Spectra MatOp/SparseGenMatProd.h:81-86)
    Hamiltonian.lanczos / IRLM_eigen  Op::IRLM_eigen (src/operator.cpp:22-33)
Errors raise BHError with the library's message.  There is no CPU fallback: without libbhx.so or a
GPU every compute call fails.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_PKG = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG, "libbhx.so")

OK, EINVAL, ENOMEM, EBUDGET, ENOCONV, EHIP, ERANGE, EAGAIN = 0, 1, 2, 3, 4, 5, 6, 7
CONV_REF_HOP = 1
CONV_REF_INTERACTION = 2
ORDER_REFERENCE = 0
ORDER_ASCENDING = 1
ORDER_LEX = 2
ORDER_CUSTOM = 3
ORDER_BAND = 4
PREC_FP64 = 0
PREC_MIXED = 1

_STATUS = {OK: "ok", EINVAL: "invalid argument", ENOMEM: "out of memory", EBUDGET: "over HBM budget",
           ENOCONV: "not converged", EHIP: "HIP error", ERANGE: "out of range", EAGAIN: "not ready"}


class BHError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_STATUS.get(code, code)}: {msg}")
        self.code = code


class LanczosOpts(C.Structure):
    _fields_ = [("k", C.c_int), ("prec", C.c_int), ("max_iter", C.c_int), ("chunk", C.c_int),
                ("tol", C.c_double), ("dedup_rel", C.c_double), ("accept_rel", C.c_double),
                ("seed", C.c_uint64), ("fallback", C.c_int), ("want_vectors", C.c_int), ("ncv", C.c_int),
                ("fullre_d", C.c_int), ("distinct", C.c_int)]


class SweepSpec(C.Structure):
    _fields_ = [("kind", C.c_int), ("Lx", C.c_int), ("Ly", C.c_int), ("periodic", C.c_int), ("n", C.c_int),
                ("nJ", C.c_int), ("J", C.c_void_p), ("U", C.c_double), ("nmu", C.c_int), ("mu", C.c_void_p),
                ("flags", C.c_uint), ("order", C.c_int), ("width", C.c_int)]


class SweepOut(C.Structure):
    _fields_ = [("E", C.c_void_p), ("gap_ratio", C.c_void_p), ("density", C.c_void_p),
                ("compressibility", C.c_void_p), ("mu_plus", C.c_void_p), ("mu_minus", C.c_void_p),
                ("status", C.c_void_p), ("used_fp64", C.c_void_p), ("width", C.c_void_p)]


class Request(C.Structure):
    _fields_ = [("kind", C.c_int), ("Lx", C.c_int), ("Ly", C.c_int), ("periodic", C.c_int), ("n", C.c_int),
                ("J", C.c_double), ("U", C.c_double), ("mu", C.c_double)]


class ServerOpts(C.Structure):
    _fields_ = [("lanes", C.c_int), ("prec", C.c_int), ("hbm_budget", C.c_size_t), ("engine_bytes", C.c_size_t),
                ("flags", C.c_uint), ("lanczos", LanczosOpts)]


class Result(C.Structure):
    _fields_ = [("id", C.c_uint64), ("status", C.c_int), ("k", C.c_int), ("evals", C.c_double * 16),
                ("resid", C.c_double * 16), ("iters", C.c_int), ("used_fp64", C.c_int), ("t_submit", C.c_double),
                ("t_start", C.c_double), ("t_done", C.c_double)]


class ServerStats(C.Structure):
    _fields_ = [("submitted", C.c_int64), ("completed", C.c_int64), ("queued", C.c_int64), ("in_flight", C.c_int64),
                ("engines", C.c_int), ("hbm_used", C.c_size_t), ("p50_latency", C.c_double),
                ("p99_latency", C.c_double), ("mean_latency", C.c_double), ("busy_seconds", C.c_double)]


class LanczosResult(C.Structure):
    _fields_ = [("evals", C.c_void_p), ("resid", C.c_void_p), ("errest", C.c_void_p), ("iters", C.c_void_p),
                ("used_fp64", C.c_void_p), ("status", C.c_void_p)]


_lib = None
_vp = C.c_void_p


def lib():
    """Loads libbhx.so (raises when it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise BHError(EHIP, f"libbhx.so not built at {LIB_PATH}")
        L = C.CDLL(LIB_PATH)
        sig = {
            "bh_ctx_create": (C.c_int, [C.c_int, C.c_size_t, C.POINTER(_vp)]),
            "bh_ctx_destroy": (C.c_int, [_vp]),
            "bh_last_error": (C.c_char_p, [_vp]),
            "bh_ctx_memory": (C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]),
            "bh_ctx_set_timing": (C.c_int, [_vp, C.c_int]),
            "bh_ctx_kernel_stats": (C.c_int, [_vp, C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_double), C.c_int]),
            "bh_device_name": (C.c_int, [C.c_int, C.c_char_p, C.c_int]),
            "bh_lattice_neighbours": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp]),
            "bh_basis_build": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(_vp)]),
            "bh_basis_destroy": (C.c_int, [_vp]),
            "bh_basis_dim": (C.c_int64, [_vp]),
            "bh_basis_copy_states": (C.c_int, [_vp, _vp]),
            "bh_basis_copy_tags": (C.c_int, [_vp, _vp]),
            "bh_basis_copy_perm": (C.c_int, [_vp, _vp, _vp]),
            "bh_basis_permute": (C.c_int, [_vp, _vp, _vp]),
            "bh_basis_reorder_band": (C.c_int, [_vp, _vp, _vp, _vp]),
            "bh_hamiltonian_build": (C.c_int, [_vp, _vp, _vp, _vp, C.c_uint, C.POINTER(_vp)]),
            "bh_ham_destroy": (C.c_int, [_vp]),
            "bh_ham_dim": (C.c_int64, [_vp]),
            "bh_ham_nnz": (C.c_int64, [_vp]),
            "bh_ham_device_bytes": (C.c_int64, [_vp]),
            "bh_ham_copy_csr": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
            "bh_ham_copy_packed": (C.c_int, [_vp, _vp, C.POINTER(C.c_int)]),
            "bh_ham_prepare": (C.c_int, [_vp, C.c_int]),
            "bh_spmv": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
            "bh_spmv_noalloc": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp, _vp]),
            "bh_spmv_host": (C.c_int, [_vp, _vp, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
            "bh_spmv_bench": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_double),
                                        C.POINTER(C.c_double)]),
            "bh_ham_tuned": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                       C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_int]),
            "bh_ham_set_tuned": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
            "bh_spdm": (C.c_int, [_vp, _vp, _vp, _vp]),
            "bh_mean_field": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, C.c_int, C.c_double, C.c_int, C.c_int, _vp, _vp,
                                        _vp, _vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
        _bind_optional(L)
    return _lib


def _bind_optional(L):
    opt = {
        "bh_lanczos_default_opts": (None, [C.POINTER(LanczosOpts)]),
        "bh_lanczos": (C.c_int, [_vp, _vp, C.c_int, _vp, _vp, C.POINTER(LanczosOpts), C.POINTER(LanczosResult)]),
        "bh_lanczos_copy_vectors": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
        "bh_lanczos_tridiag": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_int, C.c_uint64, _vp, _vp,
                                         C.POINTER(C.c_int), _vp]),
        "bh_dense_eigen": (C.c_int, [_vp, C.c_int64, _vp, _vp, _vp, _vp, _vp]),
        "bh_tridiag_eigvals": (C.c_int, [C.c_int, _vp, _vp, _vp]),
        "bh_sweep": (C.c_int, [_vp, C.POINTER(SweepSpec), C.POINTER(LanczosOpts), C.POINTER(SweepOut)]),
        "bh_server_create": (C.c_int, [_vp, C.POINTER(ServerOpts), C.POINTER(_vp)]),
        "bh_server_submit": (C.c_int, [_vp, C.POINTER(Request), C.POINTER(C.c_uint64)]),
        "bh_server_poll": (C.c_int, [_vp, C.c_uint64, C.POINTER(Result)]),
        "bh_server_wait": (C.c_int, [_vp, C.c_uint64, C.c_double, C.POINTER(Result)]),
        "bh_server_drain": (C.c_int, [_vp, C.c_double]),
        "bh_server_stats_get": (C.c_int, [_vp, C.POINTER(ServerStats)]),
        "bh_server_destroy": (C.c_int, [_vp]),
    }
    for name, (res, args) in opt.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def lattice(geometry: str, m: int = 0, Lx: int = 0, Ly: int = 0, periodic: bool = True):
    """Neighbour lists as (off[m+1], sites) for a chain of m sites or an Lx x Ly square lattice."""
    if geometry == "chain":
        kind, a, b, msites = 0, m, 1, m
    elif geometry == "square":
        kind, a, b, msites = 1, Lx, Ly, Lx * Ly
    else:
        raise ValueError(geometry)
    off = np.zeros(msites + 1, np.int32)
    sites = np.zeros(4 * msites + 4, np.int32)
    t = lib().bh_lattice_neighbours(kind, a, b, int(periodic), _ptr(off), _ptr(sites))
    if t < 0:
        raise BHError(-t, "invalid lattice")
    return off, sites[:t].copy()


class Context:
    """A device context (one HIP stream, an HBM budget)."""

    def __init__(self, device: int = 0, hbm_budget: int = 0):
        self.h = _vp()
        rc = lib().bh_ctx_create(device, hbm_budget, C.byref(self.h))
        if rc != OK:
            raise BHError(rc, f"bh_ctx_create(device={device}) failed")

    def check(self, rc):
        if rc != OK:
            raise BHError(rc, lib().bh_last_error(self.h).decode())
        return rc

    def close(self):
        if self.h:
            lib().bh_ctx_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def memory(self):
        u, b = C.c_size_t(), C.c_size_t()
        self.check(lib().bh_ctx_memory(self.h, C.byref(u), C.byref(b)))
        return u.value, b.value

    def set_timing(self, on: bool):
        self.check(lib().bh_ctx_set_timing(self.h, int(on)))

    def kernel_stats(self, name: str, reset: bool = False):
        c, t = C.c_int64(), C.c_double()
        self.check(lib().bh_ctx_kernel_stats(self.h, name.encode(), C.byref(c), C.byref(t), int(reset)))
        return c.value, t.value

    def basis(self, m: int, n: int, order: int = ORDER_REFERENCE) -> "Basis":
        return Basis(self, m, n, order)

    def sweep(self, geometry: str, n: int, J, U: float, mu, m: int = 0, Lx: int = 0, Ly: int = 0,
              periodic: bool = True, k: int = 4, prec: int = PREC_MIXED, flags: int = 0, order: int = ORDER_LEX,
              width: int = 0, **kw):
        """Batched phase-diagram sweep over the sectors n-1, n, n+1 (see bh_sweep in bh_capi.h); width 0
        lets the block-width tuner choose the batch width per sector (reported in res["width"])."""
        J = np.ascontiguousarray(np.atleast_1d(J), np.float64)
        mu = np.ascontiguousarray(np.atleast_1d(mu), np.float64)
        kind, a, b = (0, m, 1) if geometry == "chain" else (1, Lx, Ly)
        sp = SweepSpec(kind, a, b, int(periodic), n, J.size, _ptr(J).value, U, mu.size, _ptr(mu).value, flags, order,
                       width)
        o = LanczosOpts()
        lib().bh_lanczos_default_opts(C.byref(o))
        o.k, o.prec = k, prec
        for key, v in kw.items():
            setattr(o, key, v)
        nJ, nmu = J.size, mu.size
        res = {"E": np.full((3, nJ, k), np.nan), "gap_ratio": np.full((3, nJ), np.nan),
               "density": np.full((nJ, nmu), np.nan), "compressibility": np.full((nJ, nmu), np.nan),
               "mu_plus": np.full(nJ, np.nan), "mu_minus": np.full(nJ, np.nan),
               "status": np.zeros((3, nJ), np.int32), "used_fp64": np.zeros((3, nJ), np.int32),
               "width": np.zeros(3, np.int32)}
        out = SweepOut(*[_ptr(res[f]).value for f in ("E", "gap_ratio", "density", "compressibility", "mu_plus",
                                                      "mu_minus", "status", "used_fp64", "width")])
        self.check(lib().bh_sweep(self.h, C.byref(sp), C.byref(o), C.byref(out)))
        return res


    def dense_eigen(self, row_ptr, col, val, vectors: bool = False):
        """All eigenvalues (and the column eigenvectors) of a symmetric CSR matrix by the device dense solve
        (bh_dense_eigen)."""
        rp = np.ascontiguousarray(row_ptr, np.int64)
        n = rp.size - 1
        col = np.ascontiguousarray(col, np.int32) if len(col) else np.zeros(1, np.int32)
        val = np.ascontiguousarray(val, np.float64) if len(val) else np.zeros(1, np.float64)
        w = np.zeros(n)
        V = np.zeros((n, n)) if vectors else None
        self.check(lib().bh_dense_eigen(self.h, n, _ptr(rp), _ptr(col), _ptr(val), _ptr(w), _ptr(V) if vectors else None))
        return (w, V.T.copy()) if vectors else w  # V holds the eigenvectors as columns in column-major order

    def mean_field(self, mu, J, psi0, q: int = 1, eps: float = 1e-6, max_outer: int = 100000, pmax: int = 32):
        """Self-consistent mean-field order parameter at the points (mu[k], J[k]) from psi0[k] (see
        bh_mean_field in bh_capi.h).  Returns a dict of psi, status (0 converged, 1 iteration cap, 2 p cap),
        iters (self-consistency iterations) and p (last truncation)."""
        mu = np.ascontiguousarray(np.atleast_1d(mu), np.float64)
        J = np.ascontiguousarray(np.atleast_1d(J), np.float64)
        psi0 = np.ascontiguousarray(np.atleast_1d(psi0), np.float64)
        if not (mu.shape == J.shape == psi0.shape) or mu.ndim != 1:
            raise ValueError("mu, J and psi0 must be 1-D arrays of one length")
        n = mu.size
        res = {"psi": np.zeros(n), "status": np.zeros(n, np.int32), "iters": np.zeros(n, np.int32),
               "p": np.zeros(n, np.int32)}
        self.check(lib().bh_mean_field(self.h, n, _ptr(mu), _ptr(J), _ptr(psi0), q, eps, max_outer, pmax,
                                       _ptr(res["psi"]), _ptr(res["status"]), _ptr(res["iters"]), _ptr(res["p"])))
        return res


class Basis:
    def __init__(self, ctx: Context, m: int, n: int, order: int = ORDER_REFERENCE):
        self.ctx, self.m, self.n, self.order = ctx, m, n, order
        self.h = _vp()
        ctx.check(lib().bh_basis_build(ctx.h, m, n, order, C.byref(self.h)))
        self.D = int(lib().bh_basis_dim(self.h))

    def states(self):
        out = np.zeros((self.D, self.m), np.uint8)
        self.ctx.check(lib().bh_basis_copy_states(self.h, _ptr(out)))
        return out

    def tags(self):
        out = np.zeros(self.D, np.float64)
        self.ctx.check(lib().bh_basis_copy_tags(self.h, _ptr(out)))
        return out

    def spdm(self, phi):
        """Single-particle density matrix <phi| b_i^+ b_j |phi> (m x m) of a real state in this basis' order."""
        phi = np.ascontiguousarray(phi, np.float64)
        if phi.shape != (self.D,):
            raise BHError(EINVAL, f"spdm: phi must have {self.D} entries")
        rho = np.zeros((self.m, self.m))
        self.ctx.check(lib().bh_spdm(self.ctx.h, self.h, _ptr(phi), _ptr(rho)))
        return rho

    def perm(self):
        a = np.zeros(self.D, np.int32)
        b = np.zeros(self.D, np.int32)
        self.ctx.check(lib().bh_basis_copy_perm(self.h, _ptr(a), _ptr(b)))
        return a, b

    def permute(self, lex_of_pos):
        """Re-orders the basis: position k holds the state of lexicographic rank lex_of_pos[k] (bh_basis_permute).
        Build Hamiltonians after this call."""
        lp = np.ascontiguousarray(lex_of_pos, np.int32)
        if lp.shape != (self.D,):
            raise BHError(EINVAL, f"permute: lex_of_pos must have {self.D} entries")
        self.ctx.check(lib().bh_basis_permute(self.ctx.h, self.h, _ptr(lp)))
        self.order = ORDER_CUSTOM

    def reorder_band(self, off, sites):
        """Re-orders the basis into the reverse Cuthill-McKee order of the lattice's hopping graph
        (bh_basis_reorder_band). Build Hamiltonians after this call."""
        off = np.ascontiguousarray(off, np.int32)
        sites = np.ascontiguousarray(sites, np.int32)
        if off.shape != (self.m + 1,):
            raise BHError(EINVAL, f"reorder_band: off must have {self.m + 1} entries")
        if sites.size == 0:
            sites = np.zeros(1, np.int32)
        self.ctx.check(lib().bh_basis_reorder_band(self.ctx.h, self.h, _ptr(off), _ptr(sites)))
        self.order = ORDER_BAND

    def close(self):
        if self.h:
            lib().bh_basis_destroy(self.h)
            self.h = _vp()

    def hamiltonian(self, off, sites, flags: int = 0) -> "Hamiltonian":
        return Hamiltonian(self, off, sites, flags)


@dataclass
class LanczosOut:
    evals: np.ndarray
    resid: np.ndarray
    errest: np.ndarray
    iters: np.ndarray
    used_fp64: np.ndarray
    status: np.ndarray


class Hamiltonian:
    def __init__(self, basis: Basis, off, sites, flags: int = 0):
        self.basis, self.ctx = basis, basis.ctx
        off = np.ascontiguousarray(off, np.int32)
        sites = np.ascontiguousarray(sites, np.int32)
        if sites.size == 0:
            sites = np.zeros(1, np.int32)
        self.h = _vp()
        self.ctx.check(lib().bh_hamiltonian_build(self.ctx.h, basis.h, _ptr(off), _ptr(sites), flags, C.byref(self.h)))
        self.D = int(lib().bh_ham_dim(self.h))
        self.nnz = int(lib().bh_ham_nnz(self.h))

    def close(self):
        if self.h:
            lib().bh_ham_destroy(self.h)
            self.h = _vp()

    def device_bytes(self):
        return int(lib().bh_ham_device_bytes(self.h))

    def csr(self):
        rp = np.zeros(self.D + 1, np.int64)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.float64)
        diag = np.zeros(self.D, np.float64)
        self.ctx.check(lib().bh_ham_copy_csr(self.h, _ptr(rp), _ptr(col), _ptr(val), _ptr(diag)))
        return rp, col[: self.nnz], val[: self.nnz], diag

    def packed(self):
        w = np.zeros(max(self.nnz, 1), np.uint32)
        cb = C.c_int()
        self.ctx.check(lib().bh_ham_copy_packed(self.h, _ptr(w), C.byref(cb)))
        return w[: self.nnz], cb.value

    def spmv(self, J, U, x, prec: int = PREC_FP64):
        """y_b = J_b T x_b + U_b V x_b for x of shape (D,) or (D, B)."""
        x = np.asarray(x, np.float64)
        one = x.ndim == 1
        X = np.ascontiguousarray(x.reshape(self.D, -1))
        B = X.shape[1]
        Jv = np.ascontiguousarray(np.broadcast_to(np.asarray(J, np.float64), (B,)))
        Uv = np.ascontiguousarray(np.broadcast_to(np.asarray(U, np.float64), (B,)))
        Y = np.zeros_like(X)
        self.ctx.check(lib().bh_spmv_host(self.ctx.h, self.h, prec, B, _ptr(Jv), _ptr(Uv), _ptr(X), _ptr(Y)))
        return Y[:, 0] if one else Y

    def prepare(self, prec: int):
        """Materialises what `prec` needs (the FP64 CSR of the FP64 SpMV) ahead of the first operator application."""
        self.ctx.check(lib().bh_ham_prepare(self.h, prec))

    def spmv_device(self, J, U, x_ptr: int, y_ptr: int, B: int, prec: int = PREC_FP64, noalloc: bool = False,
                    stream: int = 0):
        """y = H x on device buffers [D][B] (double for FP64, float for mixed) given by address; noalloc
        selects the entry point that allocates nothing (the handle must be prepared for `prec`)."""
        Jv = np.ascontiguousarray(np.broadcast_to(np.asarray(J, np.float64), (B,)))
        Uv = np.ascontiguousarray(np.broadcast_to(np.asarray(U, np.float64), (B,)))
        f = lib().bh_spmv_noalloc if noalloc else lib().bh_spmv
        self.ctx.check(f(self.ctx.h, self.h, prec, B, _ptr(Jv), _ptr(Uv), _vp(x_ptr), _vp(y_ptr),
                         _vp(stream) if stream else None))

    def spmv_bench(self, prec: int, B: int = 1, reps: int = 20):
        ms, by = C.c_double(), C.c_double()
        self.ctx.check(lib().bh_spmv_bench(self.ctx.h, self.h, prec, B, reps, C.byref(ms), C.byref(by)))
        return ms.value, by.value

    def tuned(self, kind: int, prec: int, B: int):
        """Autotuner choice for (kind 0 = Lanczos operator pass, 1 = spmv; precision; width B):
        {"variant": v, "candidates": {variant: us}}, or None before the sector was used that way."""
        v, n = C.c_int(), C.c_int()
        cap = 32
        cv, cu = (C.c_int * cap)(), (C.c_double * cap)()
        if lib().bh_ham_tuned(self.h, kind, prec, B, C.byref(v), C.byref(n), cv, cu, cap) != OK:
            return None
        return {"variant": v.value, "candidates": {cv[i]: cu[i] for i in range(min(n.value, cap))}}

    def set_tuned(self, kind: int, prec: int, B: int, variant: int):
        """Pins the operator variant for (kind, precision, width) instead of the autotuner's choice."""
        self.ctx.check(lib().bh_ham_set_tuned(self.h, kind, prec, B, variant))

    def lanczos(self, J, U, k: int = 4, prec: int = PREC_MIXED, **kw) -> LanczosOut:
        J = np.atleast_1d(np.asarray(J, np.float64))
        U = np.atleast_1d(np.asarray(U, np.float64))
        B = max(J.size, U.size)
        J = np.ascontiguousarray(np.broadcast_to(J, (B,)))
        U = np.ascontiguousarray(np.broadcast_to(U, (B,)))
        o = LanczosOpts()
        lib().bh_lanczos_default_opts(C.byref(o))
        o.k, o.prec = k, prec
        for key, v in kw.items():
            setattr(o, key, v)
        ev = np.full((B, k), np.nan)
        rs = np.full((B, k), np.nan)
        ee = np.full((B, k), np.nan)
        it = np.zeros(B, np.int32)
        uf = np.zeros(B, np.int32)
        st = np.zeros(B, np.int32)
        r = LanczosResult(_ptr(ev).value, _ptr(rs).value, _ptr(ee).value, _ptr(it).value, _ptr(uf).value,
                          _ptr(st).value)
        rc = lib().bh_lanczos(self.ctx.h, self.h, B, _ptr(J), _ptr(U), C.byref(o), C.byref(r))
        if rc not in (OK, ENOCONV):
            self.ctx.check(rc)
        return LanczosOut(ev, rs, ee, it, uf, st)

    def lanczos_vectors(self, B: int, k: int):
        v = np.zeros((B, k, self.D))
        self.ctx.check(lib().bh_lanczos_copy_vectors(self.ctx.h, B, k, _ptr(v)))
        return v

    def tridiag(self, J: float, U: float, nb_iter: int, seed: int = 0, basis: bool = False):
        """(alpha, beta, basis or None): nb_iter FP64 Lanczos steps with full reorthogonalisation
        (bh_lanczos_tridiag), cut where beta < 1e-6; beta[0] = 0, beta[t] couples t-1 and t."""
        n = min(nb_iter, self.D)
        a, b = np.zeros(n), np.zeros(n)
        steps = C.c_int()
        V = np.zeros((n, self.D)) if basis else None
        self.ctx.check(lib().bh_lanczos_tridiag(self.ctx.h, self.h, float(J), float(U), nb_iter, seed, _ptr(a), _ptr(b),
                                                C.byref(steps), _ptr(V) if basis else None))
        s = steps.value
        return a[:s], b[:s], (V[:s] if basis else None)


def device_name(device: int = 0) -> str:
    buf = C.create_string_buffer(256)
    rc = lib().bh_device_name(device, buf, 256)
    if rc != OK:
        raise BHError(rc, "bh_device_name failed")
    return buf.value.decode()


def gap_ratio(levels) -> float:
    """Mean of min(d_i, d_{i+1}) / max(d_i, d_{i+1}) over ascending levels, 0 when the max is 0
    (Analysis::gap_ratios, src/analysis.cpp:433-454, mean at :317-318)."""
    ev = np.sort(np.asarray(levels, np.float64))
    if ev.size < 3:
        return 0.0
    d = np.diff(ev)
    a, b = d[:-1], d[1:]
    mn, mx = np.minimum(a, b), np.maximum(a, b)
    r = np.where(mx != 0, mn / np.where(mx != 0, mx, 1.0), 0.0)
    return float(r.mean())


class Server:
    """Continuous-batching request server (bh_server_* of the C ABI)."""

    def __init__(self, ctx: Context, lanes: int = 16, prec: int = PREC_MIXED, k: int = 4, hbm_budget: int = 0,
                 engine_bytes: int = 0, flags: int = 0, **lanczos_kw):
        self.ctx = ctx
        o = ServerOpts()
        lib().bh_lanczos_default_opts(C.byref(o.lanczos))
        o.lanes, o.prec, o.hbm_budget, o.engine_bytes, o.flags = lanes, prec, hbm_budget, engine_bytes, flags
        o.lanczos.k = k
        for key, v in lanczos_kw.items():
            setattr(o.lanczos, key, v)
        self.h = _vp()
        ctx.check(lib().bh_server_create(ctx.h, C.byref(o), C.byref(self.h)))

    def submit(self, geometry: str, n: int, J: float, U: float, mu: float = 0.0, m: int = 0, Lx: int = 0, Ly: int = 0,
               periodic: bool = True) -> int:
        kind, a, b = (0, m, 1) if geometry == "chain" else (1, Lx, Ly)
        rid = C.c_uint64()
        rc = lib().bh_server_submit(self.h, C.byref(Request(kind, a, b, int(periodic), n, J, U, mu)), C.byref(rid))
        if rc != OK:
            raise BHError(rc, "bh_server_submit failed")
        return rid.value

    def wait(self, rid: int, timeout: float = 300.0) -> Result:
        r = Result()
        rc = lib().bh_server_wait(self.h, rid, timeout, C.byref(r))
        if rc != OK:
            raise BHError(rc, f"request {rid} not done within {timeout} s")
        return r

    def poll(self, rid: int):
        r = Result()
        return r if lib().bh_server_poll(self.h, rid, C.byref(r)) == OK else None

    def drain(self, timeout: float = 600.0) -> bool:
        return lib().bh_server_drain(self.h, timeout) == OK

    def stats(self) -> ServerStats:
        st = ServerStats()
        lib().bh_server_stats_get(self.h, C.byref(st))
        return st

    def close(self):
        if self.h:
            lib().bh_server_destroy(self.h)
            self.h = _vp()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
