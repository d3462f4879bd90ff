"""Benchmark: L=34 / N=10 DOCI ground state (131,128,140 determinants) to
TOL 1e-8 on one MI355X, mixed precision vs plain FP64.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

A step is one full Davidson solve (HF guess, TOL 1e-8, MAX_VEC 30, KEEP 2)
of the synthetic L=34, N=10 integrals (generator seed 2) with the operator
tables and diagonal already resident in HBM.  `value` is the mixed-precision
time-to-converge (seconds, lower is better); the plain FP64 solve is timed
over the same number of steps for the ratio.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "L34N10 time-to-1e-8 conv, mixed vs plain FP64; mixed-size solves/s + p99 latency"
L, N, SEED = 34, 10, 2
HBM_PEAK = 8.0e12


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(plain_its: int, threads: int, budget_s: float) -> dict:
    """The oracle's OpenMP restatement of the reference's solve (lower-triangle
    sigma with thread-private accumulators, MGS Davidson) on the host cores,
    run as a child process (oracle/cpu_baseline.py) with OMP_NUM_THREADS = the
    physical cores (lscpu sockets x cores per socket) unless --cpu-threads
    says otherwise, OMP_PROC_BIND=close, OMP_PLACES=cores: the full solve when
    it fits the budget, else three iterations extrapolated (flagged)."""
    import subprocess
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from cpu_baseline import cpu_info
    info = cpu_info()
    t = threads or info["physical_cores"]
    env = dict(os.environ, OMP_NUM_THREADS=str(t), OMP_PROC_BIND="close", OMP_PLACES="cores")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "oracle", "cpu_baseline.py"), str(L),
                        str(N), str(SEED), str(plain_its), str(budget_s)], env=env,
                       capture_output=True, text=True, timeout=budget_s * 3 + 600)
    if r.returncode != 0:
        raise RuntimeError("cpu baseline failed:\n" + r.stderr[-2000:])
    res = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"cpu baseline: {res['value']} s on {t} threads ({info['model']}, "
        f"{info['physical_cores']} physical cores), extrapolated={res['extrapolated']}")
    return {"value": res["value"], "unit": "s", "cores": t, "kind": "port",
            "extrapolated": res["extrapolated"], "sample": res["sample"],
            "cpu_model": info["model"], "physical_cores": info["physical_cores"],
            "omp": "OMP_PROC_BIND=close OMP_PLACES=cores",
            "iteration_s": res["iteration_s"], "sigma_s": res["sigma_s"]}


def measured_traffic(op, prec: str):
    """HBM bytes per sigma call of the kernel in use, from the committed
    rocprof PMC summary (FETCH_SIZE + WRITE_SIZE scaled by a 1 GiB copy), or
    None when no summary covers this kernel and shape."""
    path = os.path.join(ROOT, "profiles", "sigma_star_L34N10_traffic.json")
    if op.kernel != "star" or (op.L, op.N) != (34, 10) or not os.path.exists(path):
        return None
    with open(path) as f:
        return round(json.load(f)["hbm_bytes_per_sigma"][prec] / 1e9, 2)


def batched_leg(da, count: int) -> dict:
    """The batched-solve set (doci_amd.workloads.batched_instances): all
    instances queued at once and run to convergence in one batch, FP64 and
    mixed; baseline = the same instances as plain FP64 single solves one
    after the other.  Integrals are generated beforehand (not timed)."""
    import numpy as np
    from doci_amd.workloads import batched_instances
    inst = batched_instances(count)
    ints = [da.gen_integrals(L, s) for L, N, s in inst]
    out = {"instances": count, "dims": [min(da.binom(N, L) for L, N, _ in inst),
                                        max(da.binom(N, L) for L, N, _ in inst)]}
    # One engine serves every batch, as a long-lived service would: its first
    # batch (untimed, reported as cold_s) allocates the device arena and the
    # pinned buffers and runs the tile-size autotuner once, as the other legs
    # warm their once-per-shape tables.  Each timed batch then covers request
    # intake (the host tables), admission, the lock-step iterations and
    # retirement, with the intake pipelined in chunks of 64 behind the GPU's
    # iterations (Batch.add_pipelined); the median of three batches is
    # reported.
    items = [(L, N) + tuple(ints[k]) for k, (L, N, _) in enumerate(inst)]
    energies = {}
    with da.Batch() as b:
        t = time.perf_counter()
        b.add_pipelined(items)
        b.run()
        out["cold_s"] = round(time.perf_counter() - t, 3)
        for prec in ("fp64", "mixed"):
            runs, it0 = [], b.stats()["iterations"]
            for rep in range(3):
                t = time.perf_counter()
                ids = b.add_pipelined(items, precision=prec)
                b.run()
                runs.append(time.perf_counter() - t)
            res = [b.result(i)[0] for i in ids]
            st = b.stats()
            t = float(np.median(runs))
            energies[prec] = np.array([r.energy for r in res])
            out[prec] = {"seconds": round(t, 4), "solves_per_s": round(count / t, 1),
                         "runs_s": [round(x, 4) for x in runs],
                         "lockstep_iterations": (st["iterations"] - it0) // 3,
                         "tile_rows": st["tile_rows"],
                         "all_converged": all(r.status == 2 for r in res),
                         "max_iterations": max(r.iterations for r in res)}
            log(f"batched {prec}: {count} solves in {t:.4f} s ({count / t:.1f}/s, median of "
                f"{[round(x, 4) for x in runs]}), tile rows {st['tile_rows']}")
    t = time.perf_counter()
    e1 = []
    for k, (L, N, _) in enumerate(inst):
        with da.Operator.from_integrals(L, N, *ints[k]) as op:
            e1.append(op.solve("D").energy)
    t1 = time.perf_counter() - t
    e1 = np.array(e1)
    out["one_at_a_time_fp64"] = {"seconds": round(t1, 3), "solves_per_s": round(count / t1, 1)}
    out["speedup_fp64"] = round(t1 / out["fp64"]["seconds"], 2)
    out["speedup_mixed"] = round(t1 / out["mixed"]["seconds"], 2)
    out["max_abs_dE_vs_single"] = {p: float(np.abs(energies[p] - e1).max()) for p in energies}
    log(f"one-at-a-time fp64: {t1:.3f} s; max |dE| {out['max_abs_dE_vs_single']}")
    return out


def serving_leg(da, count: int, rates, precision: str) -> dict:
    """The serving stream (doci_amd.workloads.serving_requests) at Poisson
    rates lam = r * mu0, latencies arrival -> converged.  mu0 is the
    one-at-a-time plain FP64 throughput over all the requests: each solved as
    a single-instance solve (operator build from the in-memory integrals +
    Davidson), back to back, every service time measured.  At each rate the
    one-at-a-time latencies are those of a FIFO server with those service
    times and the same arrivals, finish_i = max(arrival_i, finish_{i-1}) +
    service_i; at the highest rate a real-time FIFO server is also run and
    its measured latencies reported beside them."""
    import numpy as np
    from doci_amd import serve
    from doci_amd.workloads import poisson_arrivals, serving_requests
    reqs = serving_requests(count)
    ints = [da.gen_integrals(L, s) for L, _, s in reqs]
    ints = [(L, N) + tuple(g) for (L, N, _), g in zip(reqs, ints)]
    # warm: one untimed solve per shape, so that the per-process set-up of a
    # shape (its star tables, built once and shared) is not charged to the
    # one-at-a-time baseline; the serving runs below find them built too
    shapes = {}
    for k, q in enumerate(reqs):
        shapes.setdefault(q[:2], k)
    serve.one_at_a_time_queue([reqs[k] for k in shapes.values()], [0.0] * len(shapes),
                              integrals=[ints[k] for k in shapes.values()])
    b2b = serve.one_at_a_time_queue(reqs, [0.0] * count, integrals=ints)
    service = b2b["service"]
    mu0 = count / float(service.sum())
    log(f"serving: one-at-a-time FP64 {mu0:.2f} solves/s over all {count} requests back to back")
    top = max(rates)
    out = {"requests": count, "precision": precision,
           "one_at_a_time_fp64_solves_per_s": round(mu0, 3), "rates": []}
    for mult in rates:
        lam = mult * mu0
        arr = poisson_arrivals(count, lam)
        res = serve.run_stream(reqs, arr, precision, integrals=list(ints))
        one_lat = serve.fifo_latencies(arr, service)
        one_rate = count / float(np.max(np.asarray(arr) + one_lat))
        dE = float(np.abs(res["energy"] - b2b["energy"]).max())
        rec = {"lambda_over_mu0": mult, "lambda": round(lam, 3),
               "solves_per_s": round(res["solves_per_s"], 3),
               "one_at_a_time_solves_per_s": round(one_rate, 3),
               "speedup_vs_one_at_a_time": round(res["solves_per_s"] / one_rate, 2),
               "p50_s": round(res["p50"], 4), "p99_s": round(res["p99"], 4),
               "one_at_a_time_p50_s": round(float(np.percentile(one_lat, 50)), 4),
               "one_at_a_time_p99_s": round(float(np.percentile(one_lat, 99)), 4),
               "peak_running": res["peak_running"], "all_converged": bool((res["status"] == 2).all()),
               "max_abs_dE_vs_single": dE}
        if mult == top:
            q = serve.one_at_a_time_queue(reqs, arr, integrals=ints)
            rec["one_at_a_time_measured_p50_s"] = round(q["p50"], 4)
            rec["one_at_a_time_measured_p99_s"] = round(q["p99"], 4)
            rec["one_at_a_time_measured_solves_per_s"] = round(count / q["makespan"], 3)
        out["rates"].append(rec)
        log(f"serving lam={mult}x: {res['solves_per_s']:.2f} solves/s vs {one_rate:.2f} one at a "
            f"time ({rec['speedup_vs_one_at_a_time']}x), p50 {res['p50']:.3f} s, "
            f"p99 {res['p99']:.3f} s (one at a time {rec['one_at_a_time_p50_s']:.3f} / "
            f"{rec['one_at_a_time_p99_s']:.3f} s), peak running {res['peak_running']}, "
            f"max |dE| {dE:.1e}")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="threads of the CPU baseline (0: the physical cores)")
    ap.add_argument("--cpu-budget", type=float, default=150.0,
                    help="seconds the CPU baseline may spend on a full solve")
    ap.add_argument("--calib", action="store_true",
                    help="first run a 1 GiB device copy (byte-counter calibration)")
    ap.add_argument("--cgp", type=int, default=1, help="1: also solve the L=34 problem with CGP")
    ap.add_argument("--batch", type=int, default=512,
                    help="instances of the batched-solve leg (0 skips it)")
    ap.add_argument("--serve", type=int, default=2000,
                    help="requests of the serving leg (0 skips it)")
    ap.add_argument("--serve-rates", default="1,5,10",
                    help="comma-separated arrival rates as multiples of mu0")
    ap.add_argument("--serve-precision", default="mixed", choices=["fp64", "mixed"],
                    help="precision flag of every served request (the baseline is plain FP64)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the DOCI solve here is a single-GPU "
              "workload (see DESIGN.md for how it would shard)", file=sys.stderr)
        return 2
    import numpy as np
    import doci_amd as da

    if args.calib:
        import ctypes
        hip = ctypes.CDLL("libamdhip64.so")
        nb = 1 << 30
        a, b = ctypes.c_void_p(), ctypes.c_void_p()
        assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nb)) == 0
        assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nb)) == 0
        assert hip.hipMemset(a, 0, ctypes.c_size_t(nb)) == 0
        da._check(da.lib().doci_dev_copy(b, a, nb // 8), "doci_dev_copy")
        hip.hipFree(a)
        hip.hipFree(b)
        log("calibration copy: 1 GiB read + 1 GiB written by k_convert<double,double>")

    log(f"operator L={L} N={N} seed={SEED}")
    t0 = time.perf_counter()
    op = da.Operator.synthetic(L, N, SEED)
    # one untimed sigma per storage precision: the per-process loading of the
    # kernels' code objects is set-up, not part of a solve (also with --warmup 0)
    op.time_sigma(1, 1)
    op.time_sigma(0, 1)
    log(f"setup {time.perf_counter() - t0:.2f} s, n={op.n}")

    def solve(prec):
        t = time.perf_counter()
        r = op.solve("D", tol=1e-8, max_its=200, keep=2, max_vec=30, precision=prec,
                     want_x=False)
        return r, time.perf_counter() - t

    for w in range(args.warmup):
        r, t = solve("mixed")
        log(f"warmup mixed {t:.3f} s E={r.energy:.12f} its={r.info.iterations}")
    runs = {"fp64": [], "mixed": []}
    for k in range(args.steps):
        for prec in ("mixed", "fp64"):
            r, t = solve(prec)
            runs[prec].append((r, t))
            log(f"step {k} {prec}: {t:.3f} s its={r.info.iterations} "
                f"(fp32 {r.info.iterations_fp32}) E={r.energy:.12f} res={r.info.residual:.2e} "
                f"sigma {r.info.seconds_sigma:.3f} s / {r.info.sigma_calls}")
    t_mixed = float(np.median([t for _, t in runs["mixed"]]))
    t_plain = float(np.median([t for _, t in runs["fp64"]]))
    rm, rp = runs["mixed"][-1][0], runs["fp64"][-1][0]
    dE = abs(rm.energy - rp.energy)

    # dominant kernel: the FP32-storage sigma of the mixed solve (event-timed on its stream)
    n = op.n
    if rm.info.sigma_calls_fp32:
        avg = rm.info.seconds_sigma_fp32 / rm.info.sigma_calls_fp32
        bytes_per = n * (2 * 4 + 8)
        kern = "sigma FP32 storage"
    else:
        avg = rm.info.seconds_sigma / rm.info.sigma_calls
        bytes_per = n * (2 * 8 + 8)
        kern = "sigma FP64"
    achieved = bytes_per / avg
    avg64 = rp.info.seconds_sigma / max(1, rp.info.sigma_calls)
    traffic = measured_traffic(op, "fp32" if "FP32" in kern else "fp64")
    out = {
        "metric": METRIC,
        "value": round(t_mixed, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * t_mixed, 2),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-storage/f64-accumulate, f64 finish",
        "data": "synthetic (seeded generator, L=34 N=10 seed 2)",
        "config": {"workload": "DOCI L=34 N=10 Davidson to TOL 1e-8, HF guess, MAX_VEC 30, KEEP 2",
                   "determinants": n, "solver": "D"},
        "plain_fp64_s": round(t_plain, 4),
        "speedup_mixed_vs_fp64": round(t_plain / t_mixed, 3),
        "energy_mixed": rm.energy,
        "energy_fp64": rp.energy,
        "abs_dE": dE,
        "iterations": {"mixed": rm.info.iterations, "mixed_fp32": rm.info.iterations_fp32,
                       "fp64": rp.info.iterations},
        "sigma_ms": {"fp32": round(1e3 * avg, 3), "fp64": round(1e3 * avg64, 3)},
        "roofline": {"bound": "hbm", "kernel": kern, "achieved": round(achieved / 1e9, 1),
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK, 4),
                     "traffic": traffic,
                     "traffic_unit": "GB per sigma call (rocprof FETCH_SIZE + WRITE_SIZE, calibrated)",
                     # the kernels' own HBM rate: measured traffic over the live sigma time
                     "traffic_GBps": None if traffic is None else round(traffic / avg, 1),
                     "algorithmic_GB": round(bytes_per / 1e9, 3)},
    }
    if args.cgp:
        # the same integrals solved with CGP (src/CG.c), plain FP64, reported beside
        # the Davidson figures (its stopping test is |g|^2 <= TOL, not |r| <= TOL)
        t = time.perf_counter()
        rc = op.solve("CGP", tol=1e-8, max_its=200, want_x=False)
        tc = time.perf_counter() - t
        out["cgp_fp64"] = {"seconds": round(tc, 4), "iterations": rc.info.iterations, "max_its": 200,
                           "energy": rc.energy, "abs_dE_vs_davidson_fp64": abs(rc.energy - rp.energy),
                           "sigma_calls": rc.info.sigma_calls}
        log(f"cgp fp64: {tc:.3f} s its={rc.info.iterations} E={rc.energy:.12f} "
            f"|dE| vs Davidson {abs(rc.energy - rp.energy):.1e}")
    if not args.no_cpu:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(rp.info.iterations, args.cpu_threads, args.cpu_budget)
    op.close()
    if args.batch:
        out["batched"] = batched_leg(da, args.batch)
    if args.serve:
        rates = [float(x) for x in args.serve_rates.split(",")]
        out["serving"] = serving_leg(da, args.serve, rates, args.serve_precision)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
