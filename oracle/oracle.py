"""ctypes front end of the CPU oracle (liboracle.so) and of the reference checker (_ref/libbhref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the cpu_baseline leg of bench.py.
The product library never imports this module.

Function-to-reference map (reference file:line):
    dimension            src/hamiltonian.cpp:39-54
    chain_neighbours     src/neighbours.cpp:21-34
    square_neighbours    src/neighbours.cpp:40-71 (shadowing local at :45 fixed; Lx x Ly extension)
    basis                src/hamiltonian.cpp:60-123, 143-149
    hamiltonian          src/hamiltonian.cpp:170-211 (intended operator, exact target search)
    lanczos              src/operator.cpp:22-33 (+ Spectra start vector / convergence rule), distinct levels
    lanczos_mult         src/operator.cpp:22-33, eigenvalues with multiplicity as Spectra returns them
    gap_ratio            src/analysis.cpp:433-454
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

CONV_REF_HOP = 1
CONV_REF_INTERACTION = 2

_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        L.orc_binom.restype = C.c_int64
        L.orc_binom.argtypes = [C.c_int, C.c_int]
        L.orc_dimension.restype = C.c_int64
        L.orc_dimension.argtypes = [C.c_int, C.c_int]
        L.orc_key_fits.restype = C.c_int
        L.orc_key_fits.argtypes = [C.c_int, C.c_int]
        L.orc_chain_neighbours.restype = C.c_int
        L.orc_chain_neighbours.argtypes = [C.c_int, C.c_int, _i32p, _i32p]
        L.orc_square_neighbours.restype = C.c_int
        L.orc_square_neighbours.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p]
        L.orc_basis.restype = C.c_int
        L.orc_basis.argtypes = [C.c_int, C.c_int, C.c_int, _u8p, _f64p, _i64p]
        L.orc_ham_count.restype = C.c_int64
        L.orc_ham_count.argtypes = [C.c_int, C.c_int, _u8p, _i32p, _i32p, _i64p]
        L.orc_ham_fill.restype = C.c_int
        L.orc_ham_fill.argtypes = [C.c_int, C.c_int, _u8p, _i32p, _i32p, C.c_uint, _i64p, _i32p, _f64p, _f64p]
        L.orc_spmv.restype = None
        L.orc_spmv.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_double, C.c_double, _f64p, _f64p]
        L.orc_start_vector.restype = None
        L.orc_start_vector.argtypes = [C.c_int64, C.c_uint64, _f64p]
        L.orc_tridiag_eig.restype = C.c_int
        L.orc_tridiag_eig.argtypes = [C.c_int, _f64p, _f64p, C.c_void_p]
        L.orc_lanczos.restype = C.c_int
        L.orc_lanczos.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_double, C.c_double, C.c_int,
                                  C.c_double, C.c_int, C.c_uint64, C.c_double, _f64p, _f64p,
                                  C.POINTER(C.c_int)]
        L.orc_lanczos_mult.restype = C.c_int
        L.orc_lanczos_mult.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_double, C.c_double, C.c_int,
                                       C.c_double, C.c_int, C.c_uint64, _f64p, _f64p, C.POINTER(C.c_int)]
        L.orc_start_vector_at.restype = None
        L.orc_start_vector_at.argtypes = [C.c_int64, C.c_uint64, C.c_uint64, _f64p]
        L.orc_gap_ratio.restype = C.c_double
        L.orc_gap_ratio.argtypes = [C.c_int, _f64p, C.c_void_p]
        L.orc_spdm.restype = C.c_int
        L.orc_spdm.argtypes = [C.c_int, C.c_int64, _u8p, _f64p, _f64p]
        L.orc_mf_ground.restype = C.c_double
        L.orc_mf_ground.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, _f64p, C.c_double,
                                    C.POINTER(C.c_double)]
        L.orc_scmf.restype = C.c_int
        L.orc_scmf.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int,
                               C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_int)]
        _LIB = L
    return _LIB


def ref_available() -> bool:
    return os.path.exists(os.path.join(_HERE, "_ref", "libbhref.so"))


def ref():
    """The reference's own functions (compiled from its sources); None when not built."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libbhref.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_chain_neighbours.restype = C.c_int
        R.ref_chain_neighbours.argtypes = [C.c_int, C.c_int, _i32p, _i32p]
        R.ref_square_neighbours.restype = C.c_int
        R.ref_square_neighbours.argtypes = [C.c_int, C.c_int, _i32p, _i32p]
        R.ref_fixed_set_basis.restype = C.c_int
        R.ref_fixed_set_basis.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
        R.ref_max_set_basis.restype = C.c_int
        R.ref_max_set_basis.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
        R.ref_scmf.restype = C.c_double
        R.ref_scmf.argtypes = [C.c_double, C.c_double, C.c_int, C.c_double, C.c_int]
        R.ref_h_mf.restype = None
        R.ref_h_mf.argtypes = [C.c_double, C.c_int, C.c_double, C.c_double, C.c_int, _f64p]
        R.ref_sf_density.restype = C.c_double
        R.ref_sf_density.argtypes = [_f64p, C.c_int]
        R.ref_irlm_sweep.restype = C.c_int
        R.ref_irlm_sweep.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_int, _f64p, _f64p, C.c_int, _i32p,
                                     _f64p, _f64p, C.c_int]
        R.ref_gap_ratios.restype = None
        R.ref_gap_ratios.argtypes = [_f64p, C.c_int, _f64p]
        R.ref_coherence.restype = C.c_double
        R.ref_coherence.argtypes = [_f64p, C.c_int]
        R.ref_diagonal_term.restype = C.c_int
        R.ref_diagonal_term.argtypes = [C.c_int, C.c_int, C.c_int, _f64p]
        R.ref_irlm_eigen.restype = C.c_int
        R.ref_irlm_eigen.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_double, C.c_double, C.c_int,
                                     _f64p]
        R.ref_exact_eigen.restype = C.c_int
        R.ref_exact_eigen.argtypes = [C.c_int64, _i64p, _i32p, _f64p, _f64p, C.c_double, C.c_double, _f64p]
        _REF = R
    return _REF


# ---- neighbours -------------------------------------------------------------------------------

def chain_neighbours(m: int, periodic: bool = True):
    off = np.zeros(m + 1, np.int32)
    sites = np.zeros(2 * m + 2, np.int32)
    t = lib().orc_chain_neighbours(m, int(periodic), off, sites)
    if t < 0:
        raise ValueError("invalid chain")
    return off, sites[:t].copy()


def square_neighbours(Lx: int, Ly: int, periodic: bool = True):
    m = Lx * Ly
    off = np.zeros(m + 1, np.int32)
    sites = np.zeros(4 * m + 4, np.int32)
    t = lib().orc_square_neighbours(Lx, Ly, int(periodic), off, sites)
    if t < 0:
        raise ValueError("invalid square lattice")
    return off, sites[:t].copy()


def lattice(geometry: str, m: int = 0, Lx: int = 0, Ly: int = 0, periodic: bool = True):
    if geometry == "chain":
        return chain_neighbours(m, periodic)
    if geometry == "square":
        return square_neighbours(Lx, Ly, periodic)
    raise ValueError(geometry)


# ---- basis and Hamiltonian --------------------------------------------------------------------

def dimension(m: int, n: int) -> int:
    return int(lib().orc_dimension(m, n))


ORDER_REFERENCE = 0
ORDER_ASCENDING = 1


def basis(m: int, n: int, order: int = ORDER_REFERENCE):
    """(states[D,m] uint8, tags[D], lex_rank[D]) in the reference's order (order 0) or in
    ascending prod p_i^n_i order (order 1)."""
    D = dimension(m, n)
    st = np.zeros((D, m), np.uint8)
    tags = np.zeros(D, np.float64)
    lex = np.zeros(D, np.int64)
    rc = lib().orc_basis(m, n, order, st.reshape(-1), tags, lex)
    if rc != 0:
        raise ValueError(f"orc_basis failed rc={rc}")
    return st, tags, lex


_PRIMES = (2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103,
           107, 109, 113, 127, 131)


def max_basis(m: int, n: int):
    """(states[D,m] uint8, tags[D]) of every boson number 1..n as BH::max_set_basis returns them
    (src/hamiltonian.cpp:151-166): the fixed-N bases in the reference order, one after the other, then
    sort_basis (src/hamiltonian.cpp:108-123) over the concatenation.  As in the fixed-N case, its cycle-swap
    loop places entry p of the concatenation at position sortperm[p], sortperm being the argsort of the tags
    (equal to the order of the exact keys prod p_i^n_i, which are sorted here)."""
    parts = [basis(m, N, ORDER_REFERENCE) for N in range(1, n + 1)]
    st = np.concatenate([q[0] for q in parts]) if parts else np.zeros((0, m), np.uint8)
    tags = np.concatenate([q[1] for q in parts]) if parts else np.zeros(0)
    keys = []
    for r in st:
        k = 1
        for i, v in enumerate(r):
            k *= _PRIMES[i] ** int(v)
        keys.append(k)
    perm = np.array(sorted(range(len(keys)), key=keys.__getitem__), dtype=np.int64)
    out_st, out_tags = np.empty_like(st), np.empty_like(tags)
    out_st[perm] = st
    out_tags[perm] = tags
    return out_st, out_tags


class CSR:
    def __init__(self, D, row_ptr, col, val, diag):
        self.D, self.row_ptr, self.col, self.val, self.diag = D, row_ptr, col, val, diag

    @property
    def nnz(self):
        return int(self.row_ptr[-1])

    def dense(self, J=1.0, U=1.0):
        H = np.zeros((self.D, self.D))
        rows = np.repeat(np.arange(self.D), np.diff(self.row_ptr))
        np.add.at(H, (rows, self.col), J * self.val)
        H[np.arange(self.D), np.arange(self.D)] += U * self.diag
        return H


def hamiltonian(m: int, n: int, off, sites, flags: int = 0, states=None, order: int = ORDER_REFERENCE) -> CSR:
    if states is None:
        states, _, _ = basis(m, n, order)
    D = states.shape[0]
    off = np.ascontiguousarray(off, np.int32)
    sites = np.ascontiguousarray(sites, np.int32)
    st = np.ascontiguousarray(states.reshape(-1))
    row_ptr = np.zeros(D + 1, np.int64)
    nnz = lib().orc_ham_count(m, n, st, off, sites, row_ptr)
    if nnz < 0:
        raise ValueError(f"orc_ham_count failed rc={-nnz}")
    col = np.zeros(max(nnz, 1), np.int32)
    val = np.zeros(max(nnz, 1), np.float64)
    diag = np.zeros(D, np.float64)
    rc = lib().orc_ham_fill(m, n, st, off, sites, flags, row_ptr, col, val, diag)
    if rc != 0:
        raise ValueError(f"orc_ham_fill failed rc={rc}")
    return CSR(D, row_ptr, col[:nnz].copy(), val[:nnz].copy(), diag)


def spmv(h: CSR, J: float, U: float, x):
    x = np.ascontiguousarray(x, np.float64)
    y = np.zeros(h.D)
    lib().orc_spmv(h.D, h.row_ptr, _c32(h.col), _c64(h.val), h.diag, J, U, x, y)
    return y


def _c32(a):
    return a if a.size else np.zeros(1, np.int32)


def _c64(a):
    return a if a.size else np.zeros(1, np.float64)


def start_vector(D: int, seed: int = 0):
    r = np.zeros(D)
    lib().orc_start_vector(D, seed, r)
    return r


def tridiag_eig(d, e, vectors=False):
    d = np.array(d, np.float64)
    e = np.array(e, np.float64) if len(e) else np.zeros(1)
    n = d.size
    z = np.zeros((n, n)) if vectors else None
    rc = lib().orc_tridiag_eig(n, d, e, z.ctypes.data if vectors else None)
    if rc != 0:
        raise RuntimeError("tridiagonal QL did not converge")
    return (d, z.T.copy()) if vectors else d  # z.T: column-major buffer -> columns are vectors


def lanczos(h: CSR, J: float, U: float, k: int, tol: float = 1e-10, max_iter: int = 600, seed: int = 0,
            dedup_rel: float = 1e-10):
    ev = np.zeros(k)
    bd = np.zeros(k)
    it = C.c_int(0)
    rc = lib().orc_lanczos(h.D, h.row_ptr, _c32(h.col), _c64(h.val), h.diag, J, U, k, tol, max_iter, seed,
                           dedup_rel, ev, bd, C.byref(it))
    return rc, ev, bd, it.value


def lanczos_mult(h: CSR, J: float, U: float, k: int, tol: float = 1e-10, max_iter: int = 1000, seed: int = 0):
    """(rc, evals, resid, steps): the lowest k eigenvalues of J T + U V with their multiplicity (orc_lanczos_mult)."""
    ev = np.zeros(k)
    rs = np.zeros(k)
    it = C.c_int(0)
    rc = lib().orc_lanczos_mult(h.D, h.row_ptr, _c32(h.col), _c64(h.val), h.diag, J, U, k, tol, max_iter, seed, ev, rs,
                                C.byref(it))
    return rc, ev, rs, it.value


def start_vector_at(D: int, seed: int, off: int):
    r = np.zeros(D)
    lib().orc_start_vector_at(D, seed, off, r)
    return r


def spdm(states, phi):
    """<phi| b_i^+ b_j |phi> (m x m) for a real state phi in the order of `states` (D x m occupations)."""
    states = np.ascontiguousarray(states, np.uint8)
    D, m = states.shape
    rho = np.zeros(m * m)
    rc = lib().orc_spdm(m, D, states.reshape(-1), np.ascontiguousarray(phi, np.float64), rho)
    if rc != 0:
        raise ValueError(f"orc_spdm failed rc={rc}")
    return rho.reshape(m, m)


def condensate_fraction(rho) -> float:
    """largest eigenvalue of the (real part of the) density matrix over its trace, as src/analysis.cpp:330-333"""
    w = np.linalg.eigvalsh(0.5 * (rho + rho.T))
    return float(abs(w[np.argmax(np.abs(w))]) / np.trace(rho))


def coherence(rho) -> float:
    """(sum_ij rho_ij rho_ji - sum_i rho_ii^2) / sum_ij rho_ij rho_ji, as src/analysis.cpp:541-556"""
    total = float(np.sum(rho * rho.T))
    return (total - float(np.sum(np.diag(rho) ** 2))) / total


def gap_ratio(ev) -> float:
    ev = np.ascontiguousarray(ev, np.float64)
    return float(lib().orc_gap_ratio(ev.size, ev, None))


def ref_irlm_sweep(h: CSR, J, U, nb_eigen: int, progress: bool = False):
    """The reference's OpenMP sweep (src/analysis.cpp:302-314) over the points (J[p], U[p]) of one sector:
    (threads, rc[p], seconds[p], evals[p, nb_eigen]); progress: a line per finished point on stderr."""
    J = np.ascontiguousarray(J, np.float64)
    U = np.ascontiguousarray(np.broadcast_to(U, J.shape), np.float64)
    rc = np.zeros(J.size, np.int32)
    secs = np.zeros(J.size)
    ev = np.zeros((J.size, nb_eigen))
    th = ref().ref_irlm_sweep(h.D, h.row_ptr, _c32(h.col), _c64(h.val), h.diag, J.size, J, U, nb_eigen, rc, secs,
                              ev.reshape(-1), int(progress))
    return th, rc, secs, ev


def ref_gap_ratio(ev) -> float:
    """Mean gap ratio by the reference's own Analysis::gap_ratios (src/analysis.cpp:433-454), compiled in
    oracle/_ref, as src/analysis.cpp:317-318 averages it."""
    ev = np.ascontiguousarray(ev, np.float64)
    n = ev.size
    if n < 3:
        return 0.0
    r = np.zeros(n - 2)
    ref().ref_gap_ratios(ev, n, r)
    return float(r.mean())


def distinct_levels(ev, k: int, rel: float = 1e-9):
    """Lowest k distinct values of an ascending spectrum (values within rel*max|ev| merged)."""
    ev = np.sort(np.asarray(ev))
    scale = max(abs(ev[0]), abs(ev[-1]), 1e-300)
    out = []
    for v in ev:
        if out and abs(v - out[-1]) <= rel * scale:
            continue
        out.append(v)
        if len(out) == k:
            break
    return np.array(out)


# ---- mean-field single-site problem -----------------------------------------------------------

MF_CONVERGED, MF_MAX_OUTER, MF_MAX_P = 0, 1, 2


def mf_ground(psi: float, p: int, mu: float, J: float, q: int = 1, hint: float = -np.inf):
    """(E0, <a>): ground energy of h_MF(psi, p) (src/hamiltonian.cpp:290-309) and SF_density of its ground
    state (src/analysis.cpp:95-105); hint: a start for the eigenvalue iteration (used when below it)."""
    N = 2 * p + 1
    sq = np.sqrt(np.arange(N + 1, dtype=np.float64))
    sf = C.c_double()
    e = lib().orc_mf_ground(psi, p, mu, J, q, sq, hint, C.byref(sf))
    return e, sf.value


def scmf(mu: float, J: float, q: int = 1, psi0: float = 0.5, eps: float = 1e-6, max_outer: int = 100000,
         pmax: int = 32):
    """(status, psi, outer iterations, last p) of Analysis::SCMF (src/analysis.cpp:111-178) at one point."""
    psi, it, p = C.c_double(), C.c_int(), C.c_int()
    st = lib().orc_scmf(mu, J, q, psi0, eps, max_outer, pmax, C.byref(psi), C.byref(it), C.byref(p))
    if st < 0:
        raise ValueError(f"orc_scmf failed rc={st}")
    return st, psi.value, it.value, p.value


def mf_psi0(seed: int, idx):
    """Start values of the mean-field grid: output idx of the SplitMix64 stream seeded with seed
    (z = seed + (idx + 1) * 0x9E3779B97F4A7C15, then its two xor-shift-multiply rounds), top 53 bits as a
    uniform value in [0, 1)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + (np.asarray(idx, np.uint64) + np.uint64(1)) * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
