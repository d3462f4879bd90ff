"""The GPU solvers (through the C ABI) against runs of the reference's own
solvers (tests/golden/solvers.json, see test_ref_solvers.py for what they
pin and the bars), the reference's captured log of the CO example
(tests/golden/logs/coHF.out), and the mixed-precision forms against the FP64
ones:

* CG / CGP rows, iteration counts and energies; Davidson with deflation every
  few iterations; the synthetic L=24 / N=8 input with D and CGP;
* the drop-in call sequence of the reference's main() from a C program;
* the CLI's trace and result lines against coHF.out;
* mixed-precision Davidson at L=34 / N=10 (131,128,140 determinants) against
  the FP64 solve: |dE| <= 1e-9 Eh, eigenvector overlap >= 1 - 1e-8, FP64
  residual <= 1e-8 (checked with a separate FP64 sigma of the returned x);
* the FP32-storage sigma at L=34 against the FP64 sigma (3e-7 relative, the
  FP32 rounding of the output);
* mixed-precision CG / CGP against the converged energy, and mixed Davidson
  always finishing with an FP64 check.
"""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import doci_amd as da
from conftest import GOLDEN, ROOT
from test_ref_solvers import RUNS, SENSITIVE, check_cg_rows, check_davidson_rows

pytestmark = pytest.mark.gpu

ROW = re.compile(r"^(\d+)\t(-?\d+)\t(\S+)\t(\S+)$")
CALLER = os.path.join(ROOT, "tests", "c", "dropin_caller")


def traced_solve(capfd, op, **kw):
    """op.solve with the solver's trace on; returns the result and the
    printed rows [it, info, residue, energy]."""
    capfd.readouterr()
    r = op.solve(trace=True, **kw)
    C.CDLL(None).fflush(None)
    out = capfd.readouterr().out
    rows = [[int(m.group(1)), int(m.group(2)), float(m.group(3)), float(m.group(4))]
            for m in map(ROW.match, out.splitlines()) if m]
    return r, np.array(rows), out


def path_of(inp, fcidump_dir, syn24):
    return str(syn24) if inp.startswith("syn") else str(fcidump_dir / f"{inp}.FCIDUMP")


def sample_x(x, ref):
    k = int(np.argmax(np.abs(x)))
    return (x * np.sign(x[k]))[ref["x_idx"]]


@pytest.mark.parametrize("tag", [t for t, r in RUNS.items() if r["solver"] != "D"])
def test_gpu_cg_matches_reference_runs(tag, capfd, fcidump_dir, syn24, golden):
    r = RUNS[tag]
    print(f"[case] CG fixture {tag}", flush=True)
    with da.Operator.from_fcidump(path_of(r["input"], fcidump_dir, syn24)) as op:
        q, rows, _ = traced_solve(capfd, op, solver=r["solver"], tol=r["tol"],
                                  max_its=r["max_its"])
    assert q.code == 1
    ref0 = r["trace"][0]
    assert abs(rows[0, 2] - ref0[2]) <= 2e-6 * ref0[2] and abs(rows[0, 3] - ref0[3]) <= 1e-6
    check_cg_rows(tag, r, rows[1:, 2], rows[1:, 3], q.info.iterations)
    if tag in SENSITIVE:
        assert abs(q.energy - golden[r["input"]]["dense_lowest"][0]) <= 1e-9
    else:
        assert abs(q.energy - r["energy"]) <= 1e-9, (tag, q.energy - r["energy"])
    assert abs(np.linalg.norm(q.x) - 1) <= 1e-10


@pytest.mark.parametrize("tag", [t for t, r in RUNS.items() if r["solver"] == "D"])
def test_gpu_davidson_matches_reference_runs(tag, capfd, fcidump_dir, syn24):
    r = RUNS[tag]
    print(f"[case] Davidson fixture {tag}", flush=True)
    with da.Operator.from_fcidump(path_of(r["input"], fcidump_dir, syn24)) as op:
        q, rows, _ = traced_solve(capfd, op, solver="D", tol=r["tol"], max_its=r["max_its"],
                                  keep=r["keep"], max_vec=r["max_vec"])
    assert q.code == r["rc"]
    check_davidson_rows(tag, r, rows[:, 2], rows[:, 3], q.info.iterations)
    assert abs(q.energy - r["energy"]) <= 1e-9
    assert np.abs(sample_x(q.x, r) - np.array(r["x_val"])).max() <= 1e-7


def test_gpu_dropin_caller(fcidump_dir, golden):
    """tests/c/dropin_caller makes the reference main()'s calls
    (src/executable.c:62-81) through the C ABI: the CO example's energy and
    iteration count (tests/coHF.out:24,39,43), and LiF with CGP against the
    reference's CGP run."""
    print("[case] drop-in C caller", flush=True)
    co = golden["logs"]["coHF"]
    r = subprocess.run([CALLER, str(fcidump_dir / "co_631g_1.FCIDUMP"), "D", "1e-8", "1000",
                        "2", "30"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    out = dict(l.split(" ", 1) for l in r.stdout.splitlines() if l.split(" ")[0] in
               ("basis", "code", "energy", "norm"))
    assert int(out["basis"]) == co["dim"] and int(out["code"]) == 0
    assert abs(float(out["energy"]) - co["energy"]) < 1e-9
    assert abs(float(out["norm"]) - 1) < 1e-10
    assert f"** DAVIDSON ITERATIONS : {co['iterations']}" in r.stdout
    ref = RUNS["lif_631g_1_CGP"]
    r = subprocess.run([CALLER, str(fcidump_dir / "lif_631g_1.FCIDUMP"), "CGP", "1e-8", "60",
                        "2", "30"], capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stderr
    out = dict(l.split(" ", 1) for l in r.stdout.splitlines() if l.split(" ")[0] in
               ("basis", "code", "energy"))
    assert int(out["code"]) == 1 and abs(float(out["energy"]) - ref["energy"]) < 1e-9
    assert ref["summary"].split(" with ")[0] in r.stdout  # "CG converged in 60 iterations"


def test_gpu_cli_trace_matches_reference_log(fcidump_dir, tmp_path):
    """The CLI with DOCI_TRACE=1 on the reference's input file prints the
    lines of tests/coHF.out: the echo of the input (the reference's template
    with MAX_ITS 1000), the banner, the trace rows (residues within 1e-4
    relative, energies within 1e-6), the summary line with iterations,
    residue and d_energy, the timing lines and the result lines.  Only the
    thread count of the banner and the timings differ."""
    print("[case] CLI trace vs coHF.out", flush=True)
    ref = open(os.path.join(GOLDEN, "logs", "coHF.out")).read().splitlines()
    (tmp_path / "co_631g_1.FCIDUMP").write_bytes((fcidump_dir / "co_631g_1.FCIDUMP").read_bytes())
    (tmp_path / "DOCI_input.in").write_text(
        "FCIDUMP= co_631g_1.FCIDUMP\nSOLVER = D\nMAX_ITS=1000\nTOL=1e-8\n"
        "DAVIDSON_KEEP=\nDAVIDSON_MAX_VEC=\nHF_INIT= 1\n")
    env = dict(os.environ, DOCI_TRACE="1")
    r = subprocess.run([da.CLI_PATH, "DOCI_input.in"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=180, env=env)
    assert r.returncode == 0, r.stderr
    got = r.stdout.splitlines()
    assert len(got) == len(ref), "\n".join(got)
    num = re.compile(r"[-+]?\d+\.\d+(e[-+]\d+)?")
    for k, (g, w) in enumerate(zip(got, ref)):
        if "executing DOCI" in w:
            assert g.startswith("\t** executing DOCI ("), g
        elif w.startswith(("elapsed time", "average time", "\t\t ** ELAPSED TIME")):
            assert num.sub("#", g) == num.sub("#", w), (k, g, w)
        elif ROW.match(w):
            mg, mw = ROW.match(g), ROW.match(w)
            assert mg and mg.group(1, 2) == mw.group(1, 2), (k, g, w)
            assert abs(float(mg.group(3)) / float(mw.group(3)) - 1) <= 1e-4, (k, g, w)
            assert abs(float(mg.group(4)) - float(mw.group(4))) <= 1e-6, (k, g, w)
        elif w.startswith("Davidson converged"):
            pat = r"Davidson converged in (\d+) iterations with (\S+) residue and d_energy (\S+)\."
            mg, mw = re.match(pat, g), re.match(pat, w)
            assert mg and mg.group(1) == mw.group(1), (g, w)
            assert abs(float(mg.group(2)) / float(mw.group(2)) - 1) <= 1e-3, (g, w)
            assert abs(float(mg.group(3))) <= 1e-12, g
        elif "DOCI ENERGY" in w:
            assert abs(float(g.split(":")[1]) - float(w.split(":")[1])) <= 1e-9, (g, w)
        else:
            assert g == w, (k, g, w)


def test_gpu_mixed_davidson_always_checks_fp64(fcidump_dir):
    """switch_tol below tol: the FP32 phase may reach tol by itself, yet the
    result is still confirmed by an FP64 iteration (FP64 sigma, residual and
    Rayleigh quotient); the returned x has an FP64 residual <= tol."""
    print("[case] mixed Davidson, switch_tol < tol", flush=True)
    with da.Operator.from_fcidump(str(fcidump_dir / "co_631g_1.FCIDUMP")) as op:
        r = op.solve("D", precision="mixed", switch_tol=1e-9)
        assert r.code == 0 and r.info.converged
        assert 0 < r.info.iterations_fp32 < r.info.iterations
        assert r.info.sigma_calls > r.info.sigma_calls_fp32
        x = r.x / np.linalg.norm(r.x)
        hx = op.sigma(x)
        e = float(x @ hx)
        assert np.linalg.norm(hx - e * x) <= 1e-8 * 1.05
        assert abs(e - r.energy) <= 1e-12 * abs(e)


@pytest.mark.parametrize("solver", ["CG", "CGP"])
def test_gpu_mixed_cg(solver, fcidump_dir, golden):
    """Mixed-precision RQ-CG / CGP (FP32 x, p, Ax, Ap, g; FP64 dots, 2x2
    problem, finish): converges on the reference's stopping test with an
    FP64 |g|^2 <= tol, to the FP64 solver's energy."""
    print(f"[case] mixed {solver} H2O_631G", flush=True)
    e0 = golden["H2O_631G"]["dense_lowest"][0]
    with da.Operator.from_fcidump(str(fcidump_dir / "H2O_631G.FCIDUMP")) as op:
        r64 = op.solve(solver, max_its=3000)
        rmx = op.solve(solver, max_its=3000, precision="mixed")
        assert rmx.code == 1 and rmx.info.converged and rmx.info.residual <= 1e-8
        assert rmx.info.iterations_fp32 > 0 and rmx.info.sigma_calls_fp32 > 0
        assert rmx.info.sigma_calls > rmx.info.sigma_calls_fp32  # the FP64 refresh
        # converged to the CG stopping test |2r|^2 <= 1e-8: energy error
        # O(|r|^2 / gap), a few 1e-9 here
        assert abs(rmx.energy - e0) <= 1e-8 and abs(r64.energy - e0) <= 1e-8
        assert abs(rmx.energy - r64.energy) <= 1e-8
        x = rmx.x / np.linalg.norm(rmx.x)
        hx = op.sigma(x)
        assert abs(float(x @ hx) - rmx.energy) <= 1e-12 * abs(rmx.energy)
        assert 4 * np.linalg.norm(hx - rmx.energy * x) ** 2 <= 1e-8 * 1.05


def test_gpu_mixed_cgp_configs0(syn24):
    """Mixed CGP on the synthetic L=24 / N=8 input, run to the reference's
    stopping test (its FP64 run needs more than the 200 iterations of the
    reference's, whose energy is still 2e-5 Eh high there): the FP64-finished
    |g|^2 <= tol, and the energy within 1e-7 Eh of the reference's converged
    Davidson energy, like the FP64 CGP's.  The FP32 iteration is not the FP64
    trajectory: rounding x, p, Ax and Ap to FP32 every step weakens the
    conjugacy of the directions (after 200 iterations its energy lags the
    FP64 run's by about 2e-6 Eh), so it is compared at convergence."""
    print("[case] mixed CGP L24", flush=True)
    e_d = RUNS["syn_L24_N8_s1_D"]["energy"]
    with da.Operator.from_fcidump(str(syn24)) as op:
        r64 = op.solve("CGP", max_its=5000)
        rmx = op.solve("CGP", max_its=5000, precision="mixed")
    print(f"[case]   its fp64 {r64.info.iterations} mixed {rmx.info.iterations} "
          f"(fp32 {rmx.info.iterations_fp32}) dE {rmx.energy - e_d:.2e} {r64.energy - e_d:.2e}",
          flush=True)
    assert r64.info.converged and rmx.info.converged and rmx.info.residual <= 1e-8
    assert rmx.info.iterations_fp32 > 0
    assert abs(rmx.energy - e_d) <= 1e-7 and abs(r64.energy - e_d) <= 1e-7


def test_gpu_fp32_sigma_L34():
    """FP32-storage sigma at L=34 / N=10 (FP32 gathers, FP64 sums) against the
    FP64 sigma of the same FP32-rounded vector: the FP32 rounding of the
    output, 3e-7 relative."""
    from test_gpu_parity import _sigma_fp32
    print("[case] FP32 sigma L34", flush=True)
    with da.Operator.synthetic(34, 10, 2) as op:
        c = da.gen_vector(op.n, 3).astype(np.float32).astype(np.float64)
        s64 = op.sigma(c)
        s32 = _sigma_fp32(op, c)
    err = np.abs(s32.astype(np.float64) - s64).max() / np.abs(s64).max()
    assert err <= 3e-7, err


def test_gpu_mixed_solve_L34():
    """configs: L=34 / N=10, mixed precision against plain FP64 (Davidson,
    TOL 1e-8, HF guess, MAX_VEC 30, KEEP 2)."""
    print("[case] L34 mixed vs FP64 Davidson", flush=True)
    with da.Operator.synthetic(34, 10, 2) as op:
        r64 = op.solve("D")
        rmx = op.solve("D", precision="mixed")
        assert r64.code == 0 and rmx.code == 0 and rmx.info.converged
        assert rmx.info.iterations_fp32 > 0
        assert abs(rmx.energy - r64.energy) <= 1e-9, rmx.energy - r64.energy
        assert abs(abs(rmx.x @ r64.x) - 1) <= 1e-8
        assert rmx.info.residual <= 1e-8
        x = rmx.x / np.linalg.norm(rmx.x)
        hx = op.sigma(x)
        e = float(x @ hx)
        assert abs(e - rmx.energy) <= 1e-9
        assert np.linalg.norm(hx - e * x) <= 1e-8 * 1.05


def test_gpu_cli_writes_eigenvector(fcidump_dir, golden, tmp_path):
    """EIGENVECTOR_FILE: the CLI writes the converged unit eigenvector in
    determinant order; its Rayleigh quotient (FP64 sigma) is the printed
    energy and it equals the library solve's vector up to sign."""
    print("[case] CLI eigenvector file", flush=True)
    path = fcidump_dir / "H2O_631G.FCIDUMP"
    (tmp_path / "in.in").write_text(f"FCIDUMP = {path}\nEIGENVECTOR_FILE = {tmp_path / 'x.vec'}\n")
    r = subprocess.run([da.CLI_PATH, str(tmp_path / "in.in")], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    e_cli = float([l for l in r.stdout.splitlines() if "DOCI ENERGY" in l][0].split(":")[1])
    L, N, x = da.read_vector(str(tmp_path / "x.vec"))
    with da.Operator.from_fcidump(str(path)) as op:
        assert (L, N, x.size) == (op.L, op.N, op.n)
        assert abs(np.linalg.norm(x) - 1) <= 1e-10
        e = float(x @ op.sigma(x))
        ref = op.solve("D")
    assert abs(e - e_cli) <= 1e-9 and abs(e - golden["H2O_631G"]["dense_lowest"][0]) <= 1e-9
    assert abs(abs(x @ ref.x) - 1) <= 1e-10
