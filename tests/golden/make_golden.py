"""Writes the golden fixtures of this directory from the reference itself.

Run in a container that has the reference sources at /root/reference and the
reference operator built by oracle/Makefile (oracle/_ref/libdoci_ref.so):

    python tests/golden/make_golden.py

Fixtures (data only):
  fcidump/*.FCIDUMP.gz   the reference's test inputs (tests/*.FCIDUMP), gzip'd
  DOCI_input.in          the reference's template input file
  golden.json            per input: the reference operator's diagonal and
                         sigma(seeded vector) as SHA-256 of the raw FP64 bytes
                         plus sampled entries (exact, as uint64 bit patterns);
                         the dense spectrum bottom for the small inputs; the
                         Davidson trajectories and energies the reference's
                         captured logs hold (tests/coHF.out, tests/lifHF.out,
                         tests/corandom.out, tests/lifrandom.out)
  small.npz              full diagonal and sigma for inputs with n <= 1287
"""
from __future__ import annotations

import gzip
import hashlib
import json
import os
import re
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
REF = "/root/reference"
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))

import oracle as O  # noqa: E402

import doci_amd as da  # noqa: E402

SMALL = ["H2O_STO3G", "N2_STO3G", "lili_321g", "H2O_631G"]
LARGE = ["lif_631g_1", "co_631g_1"]
SIGMA_SEED = 3
NSAMPLE = 512


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a, dtype=np.float64).tobytes()).hexdigest()


def samples(a: np.ndarray, seed: int):
    rng = np.random.default_rng(seed)
    idx = np.unique(np.concatenate([[0, a.size - 1], rng.integers(0, a.size, NSAMPLE)]))
    return [int(i) for i in idx], [int(v) for v in a[idx].view(np.uint64)]


def parse_log(path: str):
    traj, energy, its = [], None, None
    for line in open(path):
        m = re.match(r"^(\d+)\t(\d+)\t([-+.e\d]+)\t([-+.\d]+)$", line.strip("\n"))
        if m:
            traj.append([int(m.group(1)), float(m.group(3)), float(m.group(4))])
        m = re.search(r"DOCI ENERGY : ([-+.\d]+)", line)
        if m:
            energy = float(m.group(1))
        m = re.search(r"converged in (\d+) iterations", line)
        if m:
            its = int(m.group(1))
        m = re.search(r"basis size : (\d+)", line)
        if m:
            dim = int(m.group(1))
    return dict(energy=energy, iterations=its, trajectory=traj, dim=dim)


def case(path: str, name: str, small: dict, out: dict):
    R = O.RefOperator(path)
    n = R.n
    D = R.diagonal()
    c = da.gen_vector(n, SIGMA_SEED)
    s = R.matvec(c)
    di, dv = samples(D, 11)
    si, sv = samples(s, 12)
    rec = dict(n=n, diag_sha256=sha(D), diag_idx=di, diag_bits=dv, sigma_seed=SIGMA_SEED,
               sigma_sha256=sha(s), sigma_idx=si, sigma_bits=sv)
    if n <= 1287:
        H = np.column_stack([R.matvec(np.eye(n)[:, k]) for k in range(n)])
        rec["dense_lowest"] = [float(x) for x in np.linalg.eigvalsh(0.5 * (H + H.T))[:4]]
        small[name + "_diag"] = D
        small[name + "_sigma"] = s
    R.close()
    out[name] = rec


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libdoci_ref.so missing: run make -C oracle")
    os.makedirs(os.path.join(HERE, "fcidump"), exist_ok=True)
    for f in SMALL + LARGE:
        src = os.path.join(REF, "tests", f + ".FCIDUMP")
        dst = os.path.join(HERE, "fcidump", f + ".FCIDUMP.gz")
        with open(src, "rb") as fi, open(dst, "wb") as raw, \
                gzip.GzipFile(fileobj=raw, mode="wb", compresslevel=9, mtime=0) as fo:
            shutil.copyfileobj(fi, fo)
    dst = os.path.join(HERE, "DOCI_input.in")
    if os.path.exists(dst):
        os.chmod(dst, 0o644)
    shutil.copyfile(os.path.join(REF, "tests", "DOCI_input.in"), dst)
    out, small = {}, {}
    with tempfile.TemporaryDirectory() as td:
        for f in SMALL + LARGE:
            case(os.path.join(REF, "tests", f + ".FCIDUMP"), f, small, out)
        syn = os.path.join(td, "syn_L24_N8_s1.FCIDUMP")
        da.gen_fcidump(syn, 24, 8, 1)
        case(syn, "syn_L24_N8_s1", small, out)
        with open(syn, "rb") as fh:
            out["syn_L24_N8_s1"]["fcidump_sha256"] = hashlib.sha256(fh.read()).hexdigest()
    out["logs"] = {k: parse_log(os.path.join(REF, "tests", k + ".out"))
                   for k in ["coHF", "lifHF", "corandom", "lifrandom"]}
    with open(os.path.join(HERE, "golden.json"), "w") as fh:
        json.dump(out, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "small.npz"), **small)
    print("wrote", os.path.join(HERE, "golden.json"))


if __name__ == "__main__":
    main()
