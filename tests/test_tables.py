"""Host-side DOCI tables (doci_integral_tables) against the oracle's full
integral transform (src/doci.c:630-683 restated): bit-exact, on random
tensors whose stored entries have zeros and whose non-canonical entries
carry unrelated values, so that the reference's in-place unpacking order
(first non-zero canonical value of a permutation orbit wins, an all-zero
orbit keeps its raw entries) is exercised."""
import ctypes as C

import numpy as np
import pytest

import doci_amd as da
import oracle as O

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


def _lib_tables():
    l = da.lib()
    l.doci_integral_tables.argtypes = [C.c_uint, _dp, _dp, _dp, _dp, _dp, _dp]
    l.doci_integral_tables.restype = C.c_int
    return l.doci_integral_tables


def _oracle_tables(one, two, L):
    one, two = one.copy(), two.copy()
    O.lib().orc_transform_integrals(one, two, L)
    K = np.zeros(L * L)
    O.lib().orc_exchange_table(two, L, K)
    J = np.zeros(L * L)
    X = np.zeros(L * L)
    for p in range(L):
        for q in range(L):
            J[p * L + q] = two[p + p * L + q * L * L + q * L ** 3]
            X[p * L + q] = two[p + q * L + q * L * L + p * L ** 3]
    h = np.array([one[p * (L + 1)] for p in range(L)])
    return K, J, X, h


@pytest.mark.parametrize("L,seed", [(3, 1), (5, 2), (8, 3), (11, 4)])
def test_tables_match_full_transform(L, seed):
    rng = np.random.default_rng(seed)
    one = rng.normal(size=L * L)
    two = rng.normal(size=L ** 4)
    two[rng.random(L ** 4) < 0.35] = 0.0  # zeros among stored and other entries
    two[rng.random(L ** 4) < 0.05] = -0.0
    fn = _lib_tables()
    K, J, X, h = (np.zeros(L * L), np.zeros(L * L), np.zeros(L * L), np.zeros(L))
    assert fn(L, one, two, K, J, X, h) == 0
    rK, rJ, rX, rh = _oracle_tables(one, two, L)
    for a, b, name in ((K, rK, "K"), (J, rJ, "J"), (X, rX, "X"), (h, rh, "h")):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), name


def test_tables_canonical_only_input():
    """The usual FCIDUMP case: only canonical entries stored."""
    L = 9
    rng = np.random.default_rng(7)
    one = rng.normal(size=L * L)
    two = np.zeros(L ** 4)
    for i in range(L):
        for j in range(i + 1):
            for k in range(i + 1):
                for l in range(k + 1):
                    if rng.random() < 0.8:
                        two[i + L * j + L * L * k + L ** 3 * l] = rng.normal()
    fn = _lib_tables()
    K, J, X, h = (np.zeros(L * L), np.zeros(L * L), np.zeros(L * L), np.zeros(L))
    assert fn(L, one, two, K, J, X, h) == 0
    rK, rJ, rX, rh = _oracle_tables(one, two, L)
    for a, b, name in ((K, rK, "K"), (J, rJ, "J"), (X, rX, "X"), (h, rh, "h")):
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), name
