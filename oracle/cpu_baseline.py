"""CPU baseline of the DOCI Davidson solve: the oracle's OpenMP restatement of
the reference (src/doci.c lower-triangle sigma with thread-private
accumulators, src/davidson.c MGS Davidson) timed on the host.

TEST INFRASTRUCTURE ONLY: run by bench.py's cpu_baseline leg as a child
process, so that OMP_NUM_THREADS / OMP_PROC_BIND / OMP_PLACES take effect
before the OpenMP runtime starts.  Prints one JSON object.

    python oracle/cpu_baseline.py L N SEED ITERATIONS BUDGET_S

The diagonal is computed first and not timed (the GPU's time-to-converge also
starts with the diagonal resident).  One FP64 sigma is timed to project the
cost of the whole solve; if ITERATIONS sigma-dominated iterations fit in
BUDGET_S the full Davidson solve (TOL 1e-8, HF guess, MAX_VEC 30, KEEP 2) runs
to convergence and its wall time is the value; otherwise three iterations run
and the rest are extrapolated from the third iteration's time (flagged).
"""
from __future__ import annotations

import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)


def cpu_info() -> dict:
    """CPU model, sockets x cores per socket (physical cores), from lscpu."""
    info = {"model": "unknown", "physical_cores": os.cpu_count() or 1}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, check=True).stdout
        kv = dict(l.split(":", 1) for l in out.splitlines() if ":" in l)
        kv = {k.strip(): v.strip() for k, v in kv.items()}
        info["model"] = kv.get("Model name", "unknown")
        sockets = int(kv.get("Socket(s)", "1"))
        cps = int(kv.get("Core(s) per socket", "0"))
        if cps:
            info["physical_cores"] = sockets * cps
    except (OSError, subprocess.CalledProcessError, ValueError):
        pass
    return info


def main():
    L, N, seed, its, budget = (int(sys.argv[1]), int(sys.argv[2]), int(sys.argv[3]),
                               int(sys.argv[4]), float(sys.argv[5]))
    import numpy as np

    import doci_amd as da
    import oracle as O
    threads = int(os.environ.get("OMP_NUM_THREADS", "1"))
    with tempfile.TemporaryDirectory() as td:
        p = os.path.join(td, "b.FCIDUMP")
        da.gen_fcidump(p, L, N, seed)
        P = O.Problem(p)
    t = time.perf_counter()
    D = P.diagonal(threads)
    t_diag = time.perf_counter() - t
    c = da.gen_vector(P.n, 3)
    t = time.perf_counter()
    P.sigma(c, D, threads=threads)
    t_sigma = time.perf_counter() - t
    del c
    full = its * t_sigma * 1.3 <= budget
    r = P.davidson(O.hf_guess(D), tol=1e-8, max_its=200 if full else 3, D=D, threads=threads)
    tt = r["t"]
    out = {"threads": threads, "diagonal_s": round(t_diag, 3), "sigma_s": round(t_sigma, 3),
           "iterations_run": int(r["its"]), "iteration_s": [round(float(x), 3) for x in
                                                            np.diff(np.concatenate([[0], tt]))]}
    if full:
        out.update(value=round(float(tt[-1]), 3), extrapolated=False, energy=r["energy"],
                   sample=f"full Davidson solve of the L={L} N={N} problem on the host "
                          f"({r['its']} iterations to TOL 1e-8), diagonal precomputed")
    else:
        per = float(tt[2] - tt[1])
        out.update(value=round(float(tt[2]) + (its - 3) * per, 3), extrapolated=True,
                   sample=f"3 Davidson iterations of the L={L} N={N} problem on the host "
                          f"({tt[2]:.1f} s), the remaining {its - 3} of the GPU FP64 solve's "
                          f"{its} at the third iteration's {per:.2f} s (understates the growth "
                          f"of the subspace work with m)")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
