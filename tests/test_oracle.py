"""The CPU oracle pinned against the reference's own fixtures, brute force,
closed forms and LAPACK (CPU only)."""
import itertools
import os
from math import comb

import numpy as np
import pytest
import scipy.linalg as sl

import oracle as O
from conftest import LARGE, SMALL, check_golden_array

# ---------------------------------------------------------------- addressing


@pytest.mark.parametrize("L,N", [(1, 1), (4, 2), (6, 3), (7, 1), (8, 7), (10, 4), (12, 6), (12, 1)])
def test_lex_order_matches_brute_force(L, N):
    """Masks in rank order = sorted occupied tuples in lexicographic order
    (src/doci.c:472-548); the closed-form rank inverts it."""
    P_masks = np.zeros(comb(L, N), dtype=np.uint64)
    O.lib().orc_lex_masks(N, L, P_masks.size, P_masks)
    want = [sum(1 << p for p in t) for t in itertools.combinations(range(L), N)]
    assert [int(m) for m in P_masks] == want
    for r, m in enumerate(want):
        assert O.lib().orc_rank(m, N, L) == r


def test_lex_order_small_example():
    """L=4, N=2: masks 3, 5, 9, 6, 10, 12 (not numeric order)."""
    m = np.zeros(6, dtype=np.uint64)
    O.lib().orc_lex_masks(2, 4, 6, m)
    assert list(m) == [3, 5, 9, 6, 10, 12]


def test_rank_at_full_size():
    """Closed form at L=34, N=10 against direct enumeration counts."""
    L, N = 34, 10
    n = comb(L, N)
    assert O.lib().orc_rank(sum(1 << p for p in range(N)), N, L) == 0
    assert O.lib().orc_rank(sum(1 << p for p in range(L - N, L)), N, L) == n - 1
    # the first determinant with orbital 0 empty comes after all C(33, 9) with it occupied
    assert O.lib().orc_rank(sum(1 << p for p in range(1, N + 1)), N, L) == comb(L - 1, N - 1)


# ---------------------------------------------------------------- connections


@pytest.mark.parametrize("name", ["H2O_STO3G", "N2_STO3G", "H2O_631G"])
def test_connections_slater_condon(name, fcidump_dir):
    """Lower-triangle connection lists: exactly the pair hops J = I - {o} + {t}
    with t < o, value K[o][t], in the reference's order (hole index N-1 down
    to 0, target descending; src/doci.c:851-978)."""
    P = O.Problem(str(fcidump_dir / f"{name}.FCIDUMP"))
    masks = P.masks()
    for r in range(P.n):
        m = int(masks[r])
        occ = [p for p in range(P.L) if m >> p & 1]
        want_J, want_k = [], []
        for o in reversed(occ):
            for t in range(o - 1, -1, -1):
                if m >> t & 1:
                    continue
                want_J.append(O.lib().orc_rank(m ^ (1 << o) ^ (1 << t), P.N, P.L))
                want_k.append(o * P.L + t)
        J, k = P.connections(r)
        assert list(J) == want_J and list(k) == want_k
        assert all(j < r for j in J)


# ---------------------------------------------------------------- diagonal


@pytest.mark.parametrize("name", SMALL + LARGE)
def test_diagonal_matches_reference(name, fcidump_dir, golden):
    P = O.Problem(str(fcidump_dir / f"{name}.FCIDUMP"))
    D = P.diagonal()
    assert D.size == golden[name]["n"]
    check_golden_array(D, golden[name], "diag")


def test_diagonal_synthetic_matches_reference(syn24, golden):
    P = O.Problem(str(syn24))
    check_golden_array(P.diagonal(), golden["syn_L24_N8_s1"], "diag")


@pytest.mark.parametrize("name", SMALL)
def test_diagonal_closed_form(name, fcidump_dir):
    """The seniority-zero energy expression with the untransformed h:
    D_I = E_core + sum_{i in I} [2 h_ii + sum_{j in I} (2 J_ij - K_ij)]
    (to 1e-12), which pins the h' = h - 1/2 sum K transform as well."""
    P = O.Problem(str(fcidump_dir / f"{name}.FCIDUMP"))
    L = P.L
    two = P.two.reshape(L, L, L, L)  # two[l,k,j,i] = (ij|kl)
    h_raw = P.one_raw.reshape(L, L)  # file stores h_ij at [j, i] for i >= j
    h = np.tril(h_raw.T) + np.tril(h_raw.T, -1).T
    J = np.array([[two[q, q, p, p] for q in range(L)] for p in range(L)])
    K = np.array([[two[p, q, p, q] for q in range(L)] for p in range(L)])
    D = P.diagonal()
    for r, m in enumerate(P.masks()):
        occ = [p for p in range(L) if int(m) >> p & 1]
        e = P.core + sum(2 * h[i, i] + sum(2 * J[i, j] - K[i, j] for j in occ) for i in occ)
        assert abs(D[r] - e) <= 1e-12 * max(1, abs(e))


# ---------------------------------------------------------------- sigma


@pytest.mark.parametrize("name", SMALL + LARGE)
def test_sigma_matches_reference(name, fcidump_dir, golden):
    """Single-thread walk: bit-identical to the reference's doci_matvec."""
    import doci_amd as da
    P = O.Problem(str(fcidump_dir / f"{name}.FCIDUMP"))
    c = da.gen_vector(P.n, golden[name]["sigma_seed"])
    check_golden_array(P.sigma(c, threads=1), golden[name], "sigma")


def test_sigma_synthetic_matches_reference(syn24, golden):
    import doci_amd as da
    P = O.Problem(str(syn24))
    c = da.gen_vector(P.n, 3)
    check_golden_array(P.sigma(c, threads=1), golden["syn_L24_N8_s1"], "sigma")


@pytest.mark.parametrize("name", SMALL)
def test_dense_h_invariants(name, fcidump_dir, golden, golden_small):
    P = O.Problem(str(fcidump_dir / f"{name}.FCIDUMP"))
    H = P.dense_h()
    assert np.array_equal(H, H.T)
    np.testing.assert_allclose(np.trace(H), golden_small[name + "_diag"].sum(), rtol=1e-14)
    w = np.linalg.eigvalsh(H)
    np.testing.assert_allclose(w[:4], golden[name]["dense_lowest"], rtol=0, atol=1e-10)
    rng = np.random.default_rng(5)
    u, v = rng.standard_normal(P.n), rng.standard_normal(P.n)
    D = P.diagonal()
    Hu, Hv = P.sigma(u, D, threads=3), P.sigma(v, D, threads=3)
    assert abs(u @ Hv - Hu @ v) <= 1e-12 * abs(u @ Hv)
    np.testing.assert_allclose(Hv, H @ v, rtol=0, atol=1e-12 * np.abs(H @ v).max())


def test_sigma_one_pair_is_explicit_matrix():
    """N = 1: H is the L x L matrix with H_ii = E_core + 2h_ii + J_ii (h
    untransformed) and H_ia = K_ia."""
    import doci_amd as da
    import tempfile
    L = 9
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "n1.FCIDUMP")
        core, one, two = da.gen_integrals(L, 77)
        da.write_fcidump(p, L, 1, core, one, two)
        P = O.Problem(p)
    two4 = P.two.reshape(L, L, L, L)
    want = np.array([[P.core + 2 * P.one_raw[i * L + i] + two4[i, i, i, i] if i == a else
                      two4[a, i, a, i] for a in range(L)] for i in range(L)])
    H = P.dense_h()
    # rank r <-> orbital r for N = 1
    np.testing.assert_allclose(H, want, rtol=0, atol=1e-13)


def test_sigma_thread_count_only_changes_rounding(fcidump_dir):
    P = O.Problem(str(fcidump_dir / "lif_631g_1.FCIDUMP"))
    D = P.diagonal()
    c = np.random.default_rng(1).standard_normal(P.n)
    s1, s5 = P.sigma(c, D, 1), P.sigma(c, D, 5)
    assert np.abs(s1 - s5).max() <= 1e-13 * np.abs(s1).max()


# ---------------------------------------------------------------- dense solvers


def test_sygv2_matches_lapack_dsygv():
    """Eigenvalues and eigenvector signs equal LAPACK dsygv's (scipy, 'gv')."""
    rng = np.random.default_rng(3)
    for trial in range(3000):
        A = rng.standard_normal((2, 2))
        A = A + A.T
        if trial % 3 == 0:
            A *= 1e3
        X = rng.standard_normal((2, 2))
        B = X @ X.T + 0.1 * np.eye(2)
        w, Z = O.sygv2(A, B)
        w2, Z2 = sl.eigh(A, B, driver="gv")
        np.testing.assert_allclose(w, w2, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(Z, Z2, rtol=1e-9, atol=1e-12)


def test_jacobi_matches_lapack_dsyev():
    rng = np.random.default_rng(4)
    for m in [1, 2, 3, 7, 16, 30]:
        A = rng.standard_normal((m, m))
        A = A + A.T
        w, Z = O.jacobi_eig(A)
        np.testing.assert_allclose(w, sl.eigh(A, driver="ev", eigvals_only=True),
                                   atol=1e-12 * np.abs(w).max())
        np.testing.assert_allclose(A @ Z, Z * w, atol=1e-12 * np.abs(w).max())


# ---------------------------------------------------------------- solvers


def test_davidson_reproduces_reference_logs(fcidump_dir, golden):
    """HF-guess Davidson trajectories of tests/lifHF.out and tests/coHF.out."""
    for name, log in [("lif_631g_1", "lifHF"), ("co_631g_1", "coHF")]:
        P = O.Problem(str(fcidump_dir / f"{name}.FCIDUMP"))
        D = P.diagonal()
        r = P.davidson(O.hf_guess(D), tol=1e-8, max_its=1000, D=D)
        ref = golden["logs"][log]
        assert r["its"] == ref["iterations"] and r["rc"] == 0
        assert abs(r["energy"] - ref["energy"]) < 1e-9
        for (it, res, e), gres, ge in zip(ref["trajectory"], r["res"], r["e"]):
            assert abs(gres - res) <= 1e-4 * res + 1e-12
            assert abs(ge - e) <= 1e-6
        # the random-guess logs pin only the final energy
        assert abs(r["energy"] - golden["logs"][log.replace("HF", "random")]["energy"]) < 1e-9


@pytest.mark.parametrize("name", SMALL)
def test_solvers_reach_dense_ground_state(name, fcidump_dir, golden):
    P = O.Problem(str(fcidump_dir / f"{name}.FCIDUMP"))
    D = P.diagonal()
    e0 = golden[name]["dense_lowest"][0]
    r = P.davidson(O.hf_guess(D), D=D)
    assert abs(r["energy"] - e0) < 1e-9
    for pc in (False, True):
        q = P.ray_q(O.hf_guess(D), pc, tol=1e-8, max_its=2000, D=D)
        assert q["rc"] == 1
        assert abs(q["energy"] - e0) < 1e-7


def test_davidson_deflation_small_subspace(fcidump_dir, golden):
    """max_vec 4 / keep 2 forces deflations; the energy is unchanged."""
    P = O.Problem(str(fcidump_dir / "H2O_631G.FCIDUMP"))
    D = P.diagonal()
    r = P.davidson(O.hf_guess(D), max_vec=4, keep=2, max_its=500, D=D)
    assert r["rc"] == 0 and r["its"] > 4
    assert abs(r["energy"] - golden["H2O_631G"]["dense_lowest"][0]) < 1e-9


def test_davidson_max_its_code(fcidump_dir):
    P = O.Problem(str(fcidump_dir / "H2O_631G.FCIDUMP"))
    D = P.diagonal()
    r = P.davidson(O.hf_guess(D), max_its=2, D=D)
    assert r["rc"] == 1 and r["its"] == 3  # its++ in the loop test (davidson.c:72)


@pytest.mark.skipif(not O.ref_available(), reason="reference build absent")
@pytest.mark.parametrize("name", SMALL + ["lif_631g_1"])
def test_oracle_equals_reference_build(name, fcidump_dir):
    """Direct comparison with the reference's own operator (oracle/_ref)."""
    p = str(fcidump_dir / f"{name}.FCIDUMP")
    P = O.Problem(p)
    R = O.RefOperator(p)
    assert np.array_equal(P.diagonal().view(np.uint64), R.diagonal().view(np.uint64))
    c = np.random.default_rng(9).standard_normal(P.n)
    assert np.array_equal(P.sigma(c, threads=1).view(np.uint64), R.matvec(c).view(np.uint64))
    l, na, nb, core, one, two = O.ref_read_fcidump(p)
    assert (l, na, nb, core) == (P.L, P.Na, P.Nb, P.core)
    assert np.array_equal(one, P.one_raw) and np.array_equal(two, P.two_raw)
    R.close()
