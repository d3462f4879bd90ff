"""Parity of the HIP path (through the C ABI) with the reference's fixtures
and the CPU oracle.  Bars: diagonal bit-exact; FP64 sigma within 1e-12
relative to max|sigma|; energies within 1e-9 Eh; eigenvector overlap above
1 - 1e-8."""
import ctypes as C
import os
import subprocess
from math import comb

import numpy as np
import pytest

import doci_amd as da
import oracle as O
from conftest import LARGE, SMALL, check_golden_array

pytestmark = pytest.mark.gpu

SIGMA_RTOL = 1e-12


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _checked_case(op, seed=3):
    """Each blocked kernel runs bounds-checked first (no violation allowed);
    then the unchecked kernels must reproduce the checked result exactly and
    agree with the lane kernel to the FP64 bar."""
    c = da.gen_vector(op.n, seed)
    op.set_kernel("lane")
    s_lane = op.sigma(c)
    for kern in ("star", "slab"):
        op.set_kernel(kern + "-checked")
        s_chk = op.sigma(c)
        assert op.check_violations() == 0, kern
        op.set_kernel(kern)
        s = op.sigma(c)
        if kern == "slab":
            assert np.array_equal(s_chk, s), kern
        else:  # the star pass sums a centre's hops with LDS atomics: order varies
            assert rel_err(s, s_chk) <= 1e-14, kern
        assert rel_err(s, s_lane) <= SIGMA_RTOL, kern
    op.set_kernel("star")
    return c, s_lane


@pytest.mark.parametrize("name", SMALL + LARGE + ["syn", "L34"])
def test_blocked_kernels_bounds_then_parity(name, fcidump_dir, syn24, tmp_path_factory):
    """Runs before every other sigma test: the star and slab kernels' index
    arithmetic is proven in bounds on each shape used later (including L=34,
    N=10), in FP64 and FP32."""
    print(f"[case] blocked kernels checked {name}", flush=True)
    if name == "L34":
        op = da.Operator.synthetic(34, 10, 2)
    elif name == "syn":
        op = da.Operator.from_fcidump(str(syn24))
    else:
        op = da.Operator.from_fcidump(str(fcidump_dir / f"{name}.FCIDUMP"))
    with op:
        a, b, g = op.slab_split()
        assert a + b == op.L and g > 0
        sp = op.star_split()
        assert sp["a"] + sp["b"] == op.L and sp["groups"] > 0
        print(f"[case]   star split {sp}", flush=True)
        _checked_case(op)
        c5 = da.gen_vector(op.n, 5)
        for kern in ("star", "slab"):
            op.set_kernel(kern + "-checked")
            s32 = _sigma_fp32(op, c5)
            assert op.check_violations() == 0, kern
            op.set_kernel(kern)
            s = _sigma_fp32(op, c5)
            if kern == "slab":
                assert np.array_equal(s, s32)
            else:
                assert rel_err(s, s32) <= 1e-6, kern
        op.set_kernel("star")


@pytest.mark.parametrize("name", ["co_631g_1", "syn"])
def test_star_passes_and_limits(name, fcidump_dir, syn24):
    """Every selectable star pass (0 column, 1 per-star vector, 2 per-star
    FP64 matrix cores, 3 FP32 column pass on the matrix cores) and rebuilt
    star tables with other limits: bounds-checked first, then the FP64
    sigma to the 1e-12 bar and the FP32-storage sigma to 1e-6 of the lane
    kernel's."""
    print(f"[case] star passes and limits {name}", flush=True)
    path = str(syn24) if name == "syn" else str(fcidump_dir / f"{name}.FCIDUMP")
    with da.Operator.from_fcidump(path) as op:
        c = da.gen_vector(op.n, 7)
        op.set_kernel("lane")
        s_lane = op.sigma(c)
        s_lane32 = _sigma_fp32(op, c)

        def check(tag):
            print(f"[case]   {tag} split {op.star_split()}", flush=True)
            op.set_kernel("star-checked")
            s = op.sigma(c)
            s32 = _sigma_fp32(op, c)
            assert op.check_violations() == 0, tag
            assert rel_err(s, s_lane) <= SIGMA_RTOL, tag
            assert rel_err(s32, s_lane32) <= 1e-6, tag
            op.set_kernel("star")
            assert rel_err(op.sigma(c), s_lane) <= SIGMA_RTOL, tag
            assert rel_err(_sigma_fp32(op, c), s_lane32) <= 1e-6, tag

        for p in (1, 2, 3, 0):
            op.tune("star_pass", p)
            check(f"star_pass {p}")
        for key, v in (("star_sgmax", 8), ("star_lds", 40 * 1024), ("star_emax", 2048)):
            op.tune(key, v)
            check(f"{key} {v}")
            op.tune(key, 0)


def _sigma_fp32(op, c):
    import ctypes
    n = op.n
    hip = ctypes.CDLL("libamdhip64.so")
    c32 = c.astype(np.float32)
    din, dout = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(din), ctypes.c_size_t(4 * n)) == 0
    assert hip.hipMalloc(ctypes.byref(dout), ctypes.c_size_t(4 * n)) == 0
    try:
        assert hip.hipMemcpy(din, c32.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(4 * n), 1) == 0
        op.sigma_dev(din.value, dout.value, prec=1)
        out = np.zeros(n, dtype=np.float32)
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dout, ctypes.c_size_t(4 * n), 2) == 0
    finally:
        hip.hipFree(din)
        hip.hipFree(dout)
    return out


@pytest.mark.parametrize("name", SMALL + LARGE)
def test_diagonal_bit_exact(name, fcidump_dir, golden):
    print(f"[case] diagonal {name}", flush=True)
    with da.Operator.from_fcidump(str(fcidump_dir / f"{name}.FCIDUMP")) as op:
        check_golden_array(op.diagonal(), golden[name], "diag")


def test_diagonal_bit_exact_synthetic(syn24, golden):
    print("[case] diagonal syn L24 N8", flush=True)
    with da.Operator.from_fcidump(str(syn24)) as op:
        assert op.n == comb(24, 8)
        check_golden_array(op.diagonal(), golden["syn_L24_N8_s1"], "diag")


@pytest.mark.parametrize("name", SMALL + LARGE + ["syn"])
def test_sigma_fp64(name, fcidump_dir, syn24, golden, golden_small):
    print(f"[case] sigma fp64 {name}", flush=True)
    path = str(syn24) if name == "syn" else str(fcidump_dir / f"{name}.FCIDUMP")
    key = "syn_L24_N8_s1" if name == "syn" else name
    with da.Operator.from_fcidump(path) as op:
        c = da.gen_vector(op.n, 3)
        s = op.sigma(c)
    if name in SMALL:
        ref = golden_small[name + "_sigma"]
    else:
        P = O.Problem(path)
        ref = P.sigma(c, threads=1)
        check_golden_array(ref, golden[key], "sigma")  # the oracle output is the reference's
    assert rel_err(s, ref) <= SIGMA_RTOL


def test_sigma_random_vectors_and_symmetry(fcidump_dir):
    print("[case] sigma symmetry co", flush=True)
    path = str(fcidump_dir / "co_631g_1.FCIDUMP")
    P = O.Problem(path)
    D = P.diagonal()
    rng = np.random.default_rng(17)
    with da.Operator.from_fcidump(path) as op:
        for _ in range(2):
            u, v = rng.standard_normal(op.n), rng.standard_normal(op.n)
            Hu, Hv = op.sigma(u), op.sigma(v)
            assert rel_err(Hv, P.sigma(v, D, threads=os.cpu_count())) <= SIGMA_RTOL
            assert abs(u @ Hv - Hu @ v) <= 1e-12 * np.abs(Hu).max() * np.abs(v).sum()
        # linearity
        w = op.sigma(2.5 * u - v)
        assert rel_err(w, 2.5 * Hu - Hv) <= 1e-12


def _row_by_brute_force(P, masks_of, r, c, D):
    """sigma[r] from every pair hop of determinant r (closed-form ranks)."""
    m = masks_of(r)
    L, N = P.L, P.N
    acc = D[r] * c[r]
    for o in range(L):
        if not m >> o & 1:
            continue
        for t in range(L):
            if m >> t & 1:
                continue
            J = O.lib().orc_rank(m ^ (1 << o) ^ (1 << t), N, L)
            acc += P.K[o * L + t] * c[J]
    return acc


def test_sigma_full_size_L34_rows(tmp_path):
    """L=34, N=10 (131,128,140 determinants): sampled rows of the GPU sigma
    against brute-force rows built from the oracle's integrals and diagonal
    entries (which the GPU diagonal equals bit for bit), plus symmetry
    <u,Hv> = <Hu,v>."""
    print("[case] sigma L34 N10 sampled rows", flush=True)
    L, N = 34, 10
    path = str(tmp_path / "syn34.FCIDUMP")
    da.gen_fcidump(path, L, N, 2)
    P = O.Problem(path)
    with da.Operator.from_fcidump(path) as op:
        n = op.n
        assert n == comb(L, N)
        D_gpu = op.diagonal()
        c = da.gen_vector(n, 3)
        s = op.sigma(c)
        rng = np.random.default_rng(0)
        rows = np.concatenate([[0, 1, n - 1, n - 2], rng.integers(0, n, 300)])
        # the diagonal of the sampled rows from the oracle (bit-exact on the GPU too)
        D = dict(zip(rows.tolist(), P.diagonal_rows(rows)))
        assert all(np.float64(D_gpu[r]).view(np.uint64) == np.float64(D[r]).view(np.uint64)
                   for r in D)
        unrank = {}

        def masks_of(r):
            if r not in unrank:
                # walk the lexicographic tree with binomial counts
                m, p, rr = 0, 0, int(r)
                for e in range(N):
                    while True:
                        blk = comb(L - 1 - p, N - e - 1)
                        if rr < blk:
                            break
                        rr -= blk
                        p += 1
                    m |= 1 << p
                    p += 1
                assert O.lib().orc_rank(m, N, L) == r
                unrank[r] = m
            return unrank[r]

        scale = np.abs(s[rows]).max()
        for r in rows:
            want = _row_by_brute_force(P, masks_of, int(r), c, D)
            assert abs(s[r] - want) <= 1e-12 * max(scale, abs(want))
        v = da.gen_vector(n, 4)
        Hv = op.sigma(v)
        assert abs(c @ Hv - s @ v) <= 1e-11 * abs(c @ Hv) + 1e-12


def test_davidson_reproduces_reference_logs(fcidump_dir, golden):
    for name, log in [("lif_631g_1", "lifHF"), ("co_631g_1", "coHF")]:
        print(f"[case] davidson {name}", flush=True)
        path = str(fcidump_dir / f"{name}.FCIDUMP")
        with da.Operator.from_fcidump(path) as op:
            r = op.solve("D", tol=1e-8, max_its=1000)
        ref = golden["logs"][log]
        assert r.code == 0 and r.info.iterations == ref["iterations"]
        assert abs(r.energy - ref["energy"]) < 1e-9
        assert abs(np.linalg.norm(r.x) - 1) < 1e-10
        P = O.Problem(path)
        D = P.diagonal()
        orc = P.davidson(O.hf_guess(D), D=D, max_its=1000)
        assert abs(abs(orc["x"] @ r.x) - 1) < 1e-8


@pytest.mark.parametrize("name", SMALL)
def test_davidson_small_inputs(name, fcidump_dir, golden):
    """Davidson on the small inputs against their dense lowest eigenvalue
    (CG / CGP are compared with the reference's own runs in
    test_gpu_solvers.py)."""
    print(f"[case] D {name}", flush=True)
    e0 = golden[name]["dense_lowest"][0]
    with da.Operator.from_fcidump(str(fcidump_dir / f"{name}.FCIDUMP")) as op:
        r = op.solve("D")
    assert r.code == 0 and abs(r.energy - e0) < 1e-9


def test_max_its_return_code(fcidump_dir):
    print("[case] max_its code", flush=True)
    with da.Operator.from_fcidump(str(fcidump_dir / "H2O_631G.FCIDUMP")) as op:
        r = op.solve("D", max_its=2)
    assert r.code == 1 and r.info.iterations == 3


def test_mixed_precision_matches_fp64(fcidump_dir, syn24):
    for path in [str(fcidump_dir / "co_631g_1.FCIDUMP"), str(syn24)]:
        print(f"[case] mixed {os.path.basename(path)}", flush=True)
        with da.Operator.from_fcidump(path) as op:
            r64 = op.solve("D")
            rmx = op.solve("D", precision="mixed")
        assert rmx.code == 0 and rmx.info.iterations_fp32 > 0
        assert abs(rmx.energy - r64.energy) <= 1e-9
        assert abs(abs(rmx.x @ r64.x) - 1) <= 1e-8
        assert rmx.info.residual <= 1e-8


def test_sigma_fp32_storage(fcidump_dir):
    """FP32 vectors with FP64 sums: error at FP32 rounding of the output."""
    print("[case] sigma fp32", flush=True)
    path = str(fcidump_dir / "lif_631g_1.FCIDUMP")
    with da.Operator.from_fcidump(path) as op:
        c = da.gen_vector(op.n, 3)
        ref = op.sigma(c)
        for kern in ("star", "slab", "lane"):
            op.set_kernel(kern)
            out = _sigma_fp32(op, c)
            assert rel_err(out.astype(np.float64), ref) < 3e-7


def test_dropin_c_entry_points(fcidump_dir, golden):
    """fill_doci_data / doci_matvec / sparse_eigensolve as the reference's
    main() calls them (executable.c:62-81)."""
    print("[case] drop-in C ABI", flush=True)
    lib = da.lib()
    path = str(fcidump_dir / "lif_631g_1.FCIDUMP")
    dat = C.c_void_p()
    lib.fill_doci_data.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    lib.fill_doci_data(path.encode(), C.byref(dat))
    lib.doci_get_basis_size.argtypes = [C.c_void_p]
    n = lib.doci_get_basis_size(dat)
    lib.doci_get_diagonal.argtypes = [C.c_void_p]
    lib.doci_get_diagonal.restype = C.POINTER(C.c_double)
    D = np.ctypeslib.as_array(lib.doci_get_diagonal(dat), (n,)).copy()
    check_golden_array(D, golden["lif_631g_1"], "diag")
    c = da.gen_vector(n, 3)
    out = np.zeros(n)
    lib.doci_matvec.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
    lib.doci_matvec(c.ctypes.data, out.ctypes.data, dat)
    P = O.Problem(path)
    assert rel_err(out, P.sigma(c, threads=1)) <= SIGMA_RTOL
    x = O.hf_guess(D)
    e = C.c_double()
    MV = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)
    matvec = C.cast(lib.doci_matvec, C.c_void_p)
    lib.sparse_eigensolve.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_void_p,
                                      C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_char_p,
                                      C.c_int, C.c_int, C.c_void_p]
    diag_ptr = lib.doci_get_diagonal(dat)
    rc = lib.sparse_eigensolve(x.ctypes.data, n, C.byref(e), matvec,
                               C.cast(diag_ptr, C.c_void_p), None, 1e-8, 1000, b"D", 2, 30, dat)
    assert rc == 0 and abs(e.value - golden["logs"]["lifHF"]["energy"]) < 1e-9
    # unknown solver -> -1 (wrapper_solvers.c:14)
    assert lib.sparse_eigensolve(x.ctypes.data, n, C.byref(e), matvec, None, None, 1e-8, 10,
                                 b"LANCZOS", 2, 30, dat) == -1
    # a host callback matvec (generic operator) on the same problem
    H_calls = []

    @MV
    def host_matvec(vin, vout, _):
        H_calls.append(1)
        a = np.ctypeslib.as_array(vin, (n,))
        b = np.ctypeslib.as_array(vout, (n,))
        b[:] = P.sigma(a, D, threads=os.cpu_count())

    x2 = O.hf_guess(D)
    rc = lib.sparse_eigensolve(x2.ctypes.data, n, C.byref(e), C.cast(host_matvec, C.c_void_p),
                               C.cast(diag_ptr, C.c_void_p), None, 1e-8, 1000, b"D", 2, 30, None)
    assert rc == 0 and len(H_calls) == golden["logs"]["lifHF"]["iterations"]
    assert abs(e.value - golden["logs"]["lifHF"]["energy"]) < 1e-9
    lib.doci_cleanup_data.argtypes = [C.POINTER(C.c_void_p)]
    lib.doci_cleanup_data(C.byref(dat))
    assert not dat.value


def test_cli_end_to_end(fcidump_dir, golden, tmp_path):
    print("[case] CLI lif", flush=True)
    (tmp_path / "in.in").write_text(f"FCIDUMP= {fcidump_dir / 'lif_631g_1.FCIDUMP'}\nMAX_ITS=1000\n")
    r = subprocess.run([da.CLI_PATH, str(tmp_path / "in.in")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    line = [l for l in r.stdout.splitlines() if "DOCI ENERGY" in l][0]
    assert abs(float(line.split(":")[1]) - golden["logs"]["lifHF"]["energy"]) < 1e-9
    assert "EXIT CODE SOLVER : 0" in r.stdout
    assert "DAVIDSON ITERATIONS : 10" in r.stdout
