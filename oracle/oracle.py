"""ctypes front-end of the CPU oracle (liboracle.so) and, when built, of the
reference's own operator (_ref/libdoci_ref.so).

TEST INFRASTRUCTURE ONLY: tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg are the only callers.  Nothing in the doci_amd package
imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libdoci_ref.so")

_lib = None
_ref = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_up = np.ctypeslib.ndpointer(dtype=np.uint64, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build() -> None:
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (make -C oracle)")
        l = C.CDLL(LIB_PATH)
        l.orc_read_fcidump.argtypes = [C.c_char_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                       C.POINTER(C.c_uint), C.POINTER(C.c_double),
                                       C.POINTER(C.POINTER(C.c_double)),
                                       C.POINTER(C.POINTER(C.c_double))]
        l.orc_read_fcidump.restype = C.c_int
        l.orc_free.argtypes = [C.c_void_p]
        l.orc_transform_integrals.argtypes = [_dp, _dp, C.c_uint]
        l.orc_exchange_table.argtypes = [_dp, C.c_uint, _dp]
        l.orc_binom.argtypes = [C.c_uint, C.c_uint]
        l.orc_binom.restype = C.c_uint64
        l.orc_lex_masks.argtypes = [C.c_uint, C.c_uint, C.c_uint64, _up]
        l.orc_rank.argtypes = [C.c_uint64, C.c_uint, C.c_uint]
        l.orc_rank.restype = C.c_uint64
        l.orc_diagonal.argtypes = [_dp, _dp, C.c_double, C.c_uint, C.c_uint, C.c_uint64, _dp, C.c_int]
        l.orc_diagonal_rows.argtypes = [_dp, _dp, C.c_double, C.c_uint, C.c_uint, _up,
                                        C.c_int64, _dp]
        l.orc_connections.argtypes = [C.c_uint64, C.c_uint, C.c_uint, _up, _u32p, C.c_int]
        l.orc_connections.restype = C.c_int
        l.orc_sigma.argtypes = [_dp, _dp, C.c_uint, C.c_uint, C.c_uint64, _dp, _dp, C.c_int]
        l.orc_davidson.argtypes = [_dp, C.POINTER(C.c_double), C.c_int, C.c_int, C.c_double,
                                   _dp, _dp, C.c_uint, C.c_uint, C.c_uint64, C.c_int, C.c_int,
                                   _dp, _dp, C.c_int, C.POINTER(C.c_int)]
        l.orc_davidson.restype = C.c_int
        l.orc_davidson_timed.argtypes = l.orc_davidson.argtypes[:14] + [_dp] + \
            l.orc_davidson.argtypes[14:]
        l.orc_davidson_timed.restype = C.c_int
        l.orc_ray_q.argtypes = [_dp, C.POINTER(C.c_double), C.c_double, C.c_int, C.c_int,
                                _dp, _dp, C.c_uint, C.c_uint, C.c_uint64, C.c_int,
                                _dp, _dp, C.c_int, C.POINTER(C.c_int)]
        l.orc_ray_q.restype = C.c_int
        l.orc_ray_q_planted.argtypes = l.orc_ray_q.argtypes + [C.c_int]
        l.orc_ray_q_planted.restype = C.c_int
        l.orc_jacobi_eig.argtypes = [C.c_int, _dp, C.c_int, _dp, _dp, C.c_int]
        l.orc_sygv2.argtypes = [_dp, _dp, _dp, _dp]
        l.orc_sygv2.restype = C.c_int
        _lib = l
    return _lib


def binom(k: int, n: int) -> int:
    return int(lib().orc_binom(k, n))


class Problem:
    """Integrals of one FCIDUMP, transformed the way src/doci.c:630-683 does."""

    def __init__(self, path: str):
        l = lib()
        L, Na, Nb = C.c_uint(), C.c_uint(), C.c_uint()
        core = C.c_double()
        one, two = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
        rc = l.orc_read_fcidump(path.encode(), C.byref(L), C.byref(Na), C.byref(Nb),
                                C.byref(core), C.byref(one), C.byref(two))
        if rc != 0:
            raise ValueError(f"orc_read_fcidump({path}) -> {rc}")
        self.L = L.value
        self.Na, self.Nb = Na.value, Nb.value
        self.N = self.Na
        self.core = core.value
        n1, n2 = self.L ** 2, self.L ** 4
        self.one_raw = np.ctypeslib.as_array(one, (n1,)).copy()
        self.two_raw = np.ctypeslib.as_array(two, (n2,)).copy()
        l.orc_free(C.cast(one, C.c_void_p))
        l.orc_free(C.cast(two, C.c_void_p))
        self.one = self.one_raw.copy()
        self.two = self.two_raw.copy()
        l.orc_transform_integrals(self.one, self.two, self.L)
        self.K = np.zeros(self.L * self.L)
        l.orc_exchange_table(self.two, self.L, self.K)
        self.n = binom(self.N, self.L)

    def masks(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.uint64)
        lib().orc_lex_masks(self.N, self.L, self.n, out)
        return out

    def diagonal(self, threads: int = 0) -> np.ndarray:
        out = np.zeros(self.n)
        lib().orc_diagonal(self.one, self.two, self.core, self.N, self.L, self.n, out,
                           threads or os.cpu_count())
        return out

    def diagonal_rows(self, ranks) -> np.ndarray:
        """Diagonal entries at the given ranks only (same bits as diagonal())."""
        r = np.ascontiguousarray(ranks, dtype=np.uint64)
        out = np.zeros(r.size)
        lib().orc_diagonal_rows(self.one, self.two, self.core, self.N, self.L, r, r.size, out)
        return out

    def sigma(self, c: np.ndarray, D: np.ndarray | None = None, threads: int = 1) -> np.ndarray:
        D = self.diagonal() if D is None else D
        out = np.zeros(self.n)
        lib().orc_sigma(self.K, np.ascontiguousarray(D), self.N, self.L, self.n,
                        np.ascontiguousarray(c, dtype=np.float64), out, threads)
        return out

    def connections(self, rank: int):
        cap = self.N * (self.L - self.N) + 1
        J = np.zeros(cap, dtype=np.uint64)
        k = np.zeros(cap, dtype=np.uint32)
        cnt = lib().orc_connections(rank, self.N, self.L, J, k, cap)
        return J[:cnt].copy(), k[:cnt].copy()

    def dense_h(self) -> np.ndarray:
        """Dense H from the diagonal plus the connection lists (tiny n only)."""
        D = self.diagonal(1)
        H = np.diag(D)
        for r in range(self.n):
            J, k = self.connections(r)
            H[r, J.astype(np.int64)] = self.K[k]
            H[J.astype(np.int64), r] = self.K[k]
        return H

    def davidson(self, guess, tol=1e-8, max_its=200, keep=2, max_vec=30, threads=0,
                 D=None):
        D = self.diagonal() if D is None else D
        x = np.ascontiguousarray(guess, dtype=np.float64).copy()
        e = C.c_double()
        its = C.c_int()
        tr_r = np.zeros(max_its + 1)
        tr_e = np.zeros(max_its + 1)
        tr_t = np.zeros(max_its + 1)
        rc = lib().orc_davidson_timed(x, C.byref(e), max_vec, keep, tol, self.K, D, self.N,
                                      self.L, self.n, max_its, threads or os.cpu_count(), tr_r,
                                      tr_e, tr_t, max_its + 1, C.byref(its))
        k = min(its.value, max_its)
        return dict(rc=rc, energy=e.value, x=x, its=its.value, res=tr_r[:k], e=tr_e[:k],
                    t=tr_t[:k])

    def ray_q(self, guess, precond: bool, tol=1e-8, max_its=200, threads=0, D=None,
              mistake: int = 0):
        """mistake != 0 plants an error (see orc_ray_q_planted) for tests
        that show the fixture bars reject it."""
        D = self.diagonal() if D is None else D
        x = np.ascontiguousarray(guess, dtype=np.float64).copy()
        e = C.c_double()
        its = C.c_int()
        tr_r = np.zeros(max_its + 1)
        tr_e = np.zeros(max_its + 1)
        rc = lib().orc_ray_q_planted(x, C.byref(e), tol, max_its, int(precond), self.K, D,
                                     self.N, self.L, self.n, threads or os.cpu_count(), tr_r,
                                     tr_e, max_its + 1, C.byref(its), int(mistake))
        k = its.value
        return dict(rc=rc, energy=e.value, x=x, its=k, res=tr_r[:k], e=tr_e[:k])


def hf_guess(D: np.ndarray) -> np.ndarray:
    """executable.c:276-288: unit vector at the first index of min(D)."""
    g = np.zeros(D.size)
    g[int(np.argmin(D))] = 1.0
    return g


def jacobi_eig(A: np.ndarray):
    m = A.shape[0]
    a = np.asfortranarray(A, dtype=np.float64).ravel(order="F").copy()
    w = np.zeros(m)
    Z = np.zeros(m * m)
    lib().orc_jacobi_eig(m, a, m, w, Z, m)
    return w, Z.reshape(m, m, order="F")


def sygv2(A: np.ndarray, B: np.ndarray):
    a = np.asarray(A, dtype=np.float64).ravel(order="F").copy()
    b = np.asarray(B, dtype=np.float64).ravel(order="F").copy()
    w = np.zeros(2)
    Z = np.zeros(4)
    rc = lib().orc_sygv2(a, b, w, Z)
    if rc:
        raise ValueError("B not positive definite")
    return w, Z.reshape(2, 2, order="F")


# ---------------------------------------------------------------------------
# The reference's own operator, built from its sources by oracle/Makefile.
# ---------------------------------------------------------------------------

def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        r = C.CDLL(REF_PATH)
        r.fill_doci_data.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
        r.doci_get_diagonal.argtypes = [C.c_void_p]
        r.doci_get_diagonal.restype = C.POINTER(C.c_double)
        r.doci_get_basis_size.argtypes = [C.c_void_p]
        r.doci_get_basis_size.restype = C.c_int
        r.doci_cleanup_data.argtypes = [C.POINTER(C.c_void_p)]
        r.doci_matvec.argtypes = [_dp, _dp, C.c_void_p]
        r.read_fcidump.argtypes = [C.c_char_p, C.POINTER(C.c_uint), C.POINTER(C.c_uint),
                                   C.POINTER(C.c_uint), C.POINTER(C.c_double),
                                   C.POINTER(C.POINTER(C.c_double)),
                                   C.POINTER(C.POINTER(C.c_double))]
        _ref = r
    return _ref


class RefOperator:
    """fill_doci_data / doci_get_diagonal / doci_matvec of the reference build."""

    def __init__(self, path: str):
        self.h = C.c_void_p()
        ref().fill_doci_data(path.encode(), C.byref(self.h))
        self.n = ref().doci_get_basis_size(self.h)

    def diagonal(self) -> np.ndarray:
        p = ref().doci_get_diagonal(self.h)
        return np.ctypeslib.as_array(p, (self.n,)).copy()

    def matvec(self, c: np.ndarray) -> np.ndarray:
        out = np.zeros(self.n)
        ref().doci_matvec(np.ascontiguousarray(c, dtype=np.float64), out, self.h)
        return out

    def close(self):
        if self.h:
            ref().doci_cleanup_data(C.byref(self.h))

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def ref_read_fcidump(path: str):
    L, Na, Nb = C.c_uint(), C.c_uint(), C.c_uint()
    core = C.c_double()
    one, two = C.POINTER(C.c_double)(), C.POINTER(C.c_double)()
    ref().read_fcidump(path.encode(), C.byref(L), C.byref(Na), C.byref(Nb), C.byref(core),
                       C.byref(one), C.byref(two))
    l = L.value
    o = np.ctypeslib.as_array(one, (l * l,)).copy()
    t = np.ctypeslib.as_array(two, (l ** 4,)).copy()
    return l, Na.value, Nb.value, core.value, o, t
