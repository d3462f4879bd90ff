"""Host-side logic of libdoci_amd.so that needs no GPU: the exported C ABI,
the FCIDUMP reader/writer, the generator and the input-file parser."""
import os
import re
import subprocess

import numpy as np
import pytest

import doci_amd as da
import oracle as O
from conftest import GOLDEN, LARGE, ROOT, SMALL


def header_functions():
    src = open(os.path.join(ROOT, "include", "doci_amd.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{]*\)\s*;", src)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 25
    assert sorted(da.EXPORTS) == names
    out = subprocess.run(["nm", "-D", "--defined-only", da.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    lib = da.lib()
    for n in names:
        assert hasattr(lib, n)


def test_kernels_are_gfx950_code_objects():
    """The library embeds code objects for gfx950 (and nothing else)."""
    blob = open(da.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob
    assert b"amdgcn-amd-amdhsa--gfx942" not in blob


@pytest.mark.parametrize("name", SMALL + LARGE)
def test_fcidump_reader_matches_reference_layout(name, fcidump_dir):
    """Product reader == oracle reader (== the reference's, see test_oracle)."""
    p = str(fcidump_dir / f"{name}.FCIDUMP")
    L, na, nb, core, one, two = da.read_fcidump(p)
    P = O.Problem(p)
    assert (L, na, nb, core) == (P.L, P.Na, P.Nb, P.core)
    assert np.array_equal(one, P.one_raw) and np.array_equal(two, P.two_raw)


def test_fcidump_errors(tmp_path):
    good = " &FCI NORB=2,NELEC=2,MS2=0,\n &END\n 0.5 1 1 1 1\n -1.0 1 1 0 0\n 1.0 0 0 0 0\n"
    (tmp_path / "ok").write_text(good)
    L, na, nb, core, one, two = da.read_fcidump(str(tmp_path / "ok"))
    assert (L, na, nb, core) == (2, 1, 1, 1.0) and one[0] == -1.0 and two[0] == 0.5
    # a blank trailing line is fatal, with its line number (read_fcidump.c:61-67)
    (tmp_path / "blank").write_text(good + "\n")
    with pytest.raises(da.DociError) as e:
        da.read_fcidump(str(tmp_path / "blank"))
    assert e.value.code == -3 and e.value.line == 6
    (tmp_path / "short").write_text(good.replace(" 0.5 1 1 1 1", " 0.5 1 1 1"))
    with pytest.raises(da.DociError) as e:
        da.read_fcidump(str(tmp_path / "short"))
    assert e.value.line == 3
    (tmp_path / "hdr").write_text(" &FCI NORB=2\n")
    with pytest.raises(da.DociError) as e:
        da.read_fcidump(str(tmp_path / "hdr"))
    assert e.value.code == -2
    with pytest.raises(da.DociError) as e:
        da.read_fcidump(str(tmp_path / "missing"))
    assert e.value.code == -1
    (tmp_path / "big").write_text(" &FCI NORB=65,NELEC=2,MS2=0,\n &END\n")
    with pytest.raises(da.DociError) as e:
        da.read_fcidump(str(tmp_path / "big"))
    assert e.value.code == -5


def test_fcidump_alternate_terminators(tmp_path):
    for term in ["/", " /END", "&END", "  & END  "]:
        (tmp_path / "f").write_text(f" &FCI NORB=1,NELEC=2,MS2=0,\n ORBSYM=1,\n{term}\n 2.0 1 1 1 1\n")
        L, na, nb, core, one, two = da.read_fcidump(str(tmp_path / "f"))
        assert two[0] == 2.0


def test_generator_is_deterministic_and_round_trips(tmp_path, golden):
    import hashlib
    p = tmp_path / "s.FCIDUMP"
    da.gen_fcidump(str(p), 24, 8, 1)
    assert hashlib.sha256(p.read_bytes()).hexdigest() == golden["syn_L24_N8_s1"]["fcidump_sha256"]
    core, one, two = da.gen_integrals(24, 1)
    L, na, nb, c2, o2, t2 = da.read_fcidump(str(p))
    assert (L, na, nb, c2) == (24, 8, 8, core)
    assert np.array_equal(o2, one) and np.array_equal(t2, two)
    txt = p.read_text()
    assert not txt.endswith("\n") and txt.splitlines()[-1].split()[1:] == ["0", "0", "0", "0"]


def test_generator_statistics_and_symmetry():
    """h_ii follows the linear ramp, ERIs are PSD (sum of squares of B^P)."""
    L = 12
    core, one, two = da.gen_integrals(L, 5)
    P_one = one.reshape(L, L)
    diag = np.array([P_one[i, i] for i in range(L)])
    ramp = -12 + 11 * np.arange(L) / (L - 1)
    assert np.abs(diag - ramp).max() < 0.3 and core == 10.0
    # unpack the canonical entries to a full (ij|kl) supermatrix and test PSD
    P = np.zeros((L * L, L * L))
    t = two.reshape(L, L, L, L)  # [l,k,j,i]
    for (l_, k, j, i) in zip(*np.nonzero(t)):
        v = t[l_, k, j, i]
        for a, b, c, d in [(i, j, k, l_), (j, i, k, l_), (i, j, l_, k), (j, i, l_, k)]:
            P[a * L + b, c * L + d] = v
            P[c * L + d, a * L + b] = v
    w = np.linalg.eigvalsh(P)
    assert w.min() > -1e-10


def test_gen_vector_normalised():
    v = da.gen_vector(1000, 3)
    assert abs(np.linalg.norm(v) - 1) < 1e-14
    assert np.array_equal(v, da.gen_vector(1000, 3))


def test_inputfile_parser_matches_reference_echo():
    """tests/DOCI_input.in resolves as tests/coHF.out:1-19 echoes it: blank
    DAVIDSON_KEEP / DAVIDSON_MAX_VEC keep the defaults 2 and 30."""
    d = da.read_inputfile(os.path.join(GOLDEN, "DOCI_input.in"))
    assert d == dict(fcidump="co_631g_1.FCIDUMP", solver="D", max_its=1000, tol=1e-8,
                     davidson_keep=2, davidson_max_vec=30, hf_init=1, precision=0, eigvec="")


def test_inputfile_defaults_and_keys(tmp_path):
    (tmp_path / "a.in").write_text("FCIDUMP=x.FCIDUMP\n")
    d = da.read_inputfile(str(tmp_path / "a.in"))
    assert d["solver"] == "D" and d["max_its"] == 200 and d["tol"] == 1e-8
    assert d["davidson_keep"] == 2 and d["davidson_max_vec"] == 30 and d["hf_init"] == 1
    (tmp_path / "b.in").write_text("FCIDUMP = y\nSOLVER=CGP\nTOL = 1e-6\nMAX_ITS=7\n"
                                   "DAVIDSON_KEEP=3\nDAVIDSON_MAX_VEC = 12\nHF_INIT=0\n"
                                   "# SOLVER = CG\nDOCI_PRECISION = MIXED\n")
    d = da.read_inputfile(str(tmp_path / "b.in"))
    assert d == dict(fcidump="y", solver="CGP", max_its=7, tol=1e-6, davidson_keep=3,
                     davidson_max_vec=12, hf_init=0, precision=1, eigvec="")


def test_cli_rejects_bad_solver_and_missing_input(tmp_path):
    """Errors found before any GPU work (executable.c:207-210, :242-245)."""
    (tmp_path / "bad.in").write_text("FCIDUMP=x\nSOLVER=LANCZOS\n")
    r = subprocess.run([da.CLI_PATH, str(tmp_path / "bad.in")], capture_output=True, text=True)
    assert r.returncode == 1 and "no correct solver inputted!" in r.stdout
    r = subprocess.run([da.CLI_PATH, str(tmp_path / "nope.in")], capture_output=True, text=True)
    assert r.returncode == 1 and "Error reading input file" in r.stderr
    r = subprocess.run([da.CLI_PATH, "-m", "1", str(tmp_path / "bad.in")], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "Only DOCI works" in r.stderr


def test_reference_main_links_against_library():
    """The reference's own main() (executable.c + macros.c) links against
    libdoci_amd.so in place of its operator and solver sources: every symbol
    it binds from doci.h, wrapper_solvers.h and read_fcidump.h resolves in the
    library.  Built only where the reference sources exist, and never run
    (tests/c/dropin_caller makes the same calls on the GPU)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "doci_main_on_amd")
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "refmain"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    if not os.path.exists(exe):
        pytest.skip("reference sources absent: nothing to link")
    und = subprocess.run(["nm", "-D", "--undefined-only", exe], capture_output=True,
                         text=True, check=True).stdout.split()
    lib = subprocess.run(["nm", "-D", "--defined-only", da.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout.split()
    for sym in ("fill_doci_data", "doci_get_diagonal", "doci_get_basis_size",
                "doci_cleanup_data", "doci_matvec", "sparse_eigensolve"):
        assert sym in und and sym in lib, sym


def test_eigenvector_file_round_trip_and_key(tmp_path):
    """EIGENVECTOR_FILE = path is read from the input file (the reference's
    parser ignores the key), and doci_write_vector / doci_read_vector round
    trip a vector bit for bit with its (L, N) header; a damaged file is a
    format error."""
    (tmp_path / "in.in").write_text("FCIDUMP = a.FCIDUMP\nEIGENVECTOR_FILE = out.vec\n")
    assert da.read_inputfile(str(tmp_path / "in.in"))["eigvec"] == "out.vec"
    assert da.read_inputfile(os.path.join(GOLDEN, "DOCI_input.in"))["eigvec"] == ""
    x = da.gen_vector(1287, 5)
    p = str(tmp_path / "x.vec")
    da.write_vector(p, 13, 5, x)
    L, N, y = da.read_vector(p)
    assert (L, N) == (13, 5) and np.array_equal(x.view(np.uint64), y.view(np.uint64))
    raw = open(p, "rb").read()
    assert raw.startswith(b"DOCI_VECTOR 13 5 1287\n") and len(raw) == 22 + 8 * 1287
    open(p, "wb").write(raw[:-8])
    with pytest.raises(da.DociError):
        da.read_vector(p)
