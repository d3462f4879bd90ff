"""Serving loop: DOCI solve requests arriving over time, continuously batched
on one GPU through the batch engine (doci_batch_* in libdoci_amd.so).

Each request carries its integrals in memory (or an FCIDUMP path; both are
read before the clock starts).  At arrival the loop hands the request to the
batch, which transforms its integrals (intake); the engine
starts it at the next iteration boundary once its Davidson subspace fits
the device-memory budget (FIFO), runs it in lock step with everything else
that is running, and retires it when it converges.  Time-to-converge of a
request = its completion time - its arrival time, on the batch's clock.

`one_at_a_time_queue` times the baseline the stream is compared with: the
same requests solved one after the other as plain FP64 single-instance
solves, as a real-time FIFO server fed by given arrivals (all at 0: back to
back); `fifo_latencies` replays measured service times against other
arrivals.
"""
from __future__ import annotations

import time

import numpy as np

from . import Batch, Operator, binom, gen_integrals, read_fcidump


def _integrals(req):
    if isinstance(req, str):
        L, Na, Nb, core, one, two = read_fcidump(req)
        return L, Na, core, one, two
    L, N, seed = req
    core, one, two = gen_integrals(L, seed)
    return L, N, core, one, two


def run_stream(requests, arrivals, precision="fp64", integrals=None, **batch_opts) -> dict:
    """Serve `requests` ((L, N, seed) tuples or FCIDUMP paths) arriving at
    `arrivals` (s, ascending).  precision: "fp64", "mixed", or a list with
    one entry per request.  integrals: the requests' (L, N, core, one, two)
    if already at hand.  Returns per-request latencies and energies."""
    R = len(requests)
    prec = [precision] * R if isinstance(precision, str) else list(precision)
    # requests carry their integrals: produce them before the clock starts
    ints = list(integrals) if integrals is not None else [_integrals(q) for q in requests]
    batch_opts.setdefault("keep_vectors", False)  # energies only: free vectors at completion
    # a server holds its device memory: reserve the arena before the clock starts
    batch_opts.setdefault("mem_reserve", 1 << 62)
    b = Batch(**batch_opts)
    ids = [None] * R
    t_arr = [0.0] * R
    nxt = 0
    t0 = b.clock()
    peak_running = 0
    steps = 0
    while True:
        now = b.clock() - t0
        while nxt < R and arrivals[nxt] <= now:
            L, N, core, one, two = ints[nxt]
            ids[nxt] = b.add(L, N, core, one, two, prec[nxt])
            ints[nxt] = None
            t_arr[nxt] = t0 + arrivals[nxt]
            nxt += 1
            now = b.clock() - t0
        running, waiting = b.step()
        steps += 1
        peak_running = max(peak_running, running)
        if nxt == R and running == 0 and waiting == 0:
            break
        if running == 0 and waiting == 0 and nxt < R:
            # idle: wait for the next arrival
            delay = arrivals[nxt] - (b.clock() - t0)
            if delay > 0:
                time.sleep(min(delay, 0.01))
    lat, energies, its, status = [], [], [], []
    t_first, t_last = 1e300, 0.0
    for r in range(R):
        res, _ = b.result(ids[r])
        lat.append(res.t_done - t_arr[r])
        energies.append(res.energy)
        its.append(res.iterations)
        status.append(res.status)
        t_first = min(t_first, t_arr[r])
        t_last = max(t_last, res.t_done)
    st = b.stats()
    b.close()
    lat = np.array(lat)
    return dict(latency=lat, energy=np.array(energies), iterations=np.array(its),
                status=np.array(status), makespan=t_last - t_first,
                solves_per_s=R / (t_last - t_first), p50=float(np.percentile(lat, 50)),
                p99=float(np.percentile(lat, 99)), peak_running=peak_running,
                steps=steps, tile_rows=st["tile_rows"])


def one_at_a_time_queue(requests, arrivals, integrals=None) -> dict:
    """A one-at-a-time server measured in real time: requests are solved in
    arrival order as plain FP64 single-instance solves (operator build from
    the in-memory integrals + Davidson), each starting when it has arrived
    and the previous one has finished.  Returns per-request time-to-converge
    (finish - arrival) and service times."""
    ints = list(integrals) if integrals is not None else [_integrals(q) for q in requests]
    t0 = time.perf_counter()
    lat, service, energies = [], [], []
    for k, (L, N, core, one, two) in enumerate(ints):
        wait = arrivals[k] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ts = time.perf_counter()
        with Operator.from_integrals(L, N, core, one, two) as op:
            r = op.solve("D", want_x=False)
        tf = time.perf_counter()
        service.append(tf - ts)
        lat.append(tf - t0 - arrivals[k])
        energies.append(r.energy)
    lat = np.array(lat)
    return dict(latency=lat, service=np.array(service), energy=np.array(energies),
                p50=float(np.percentile(lat, 50)), p99=float(np.percentile(lat, 99)),
                makespan=time.perf_counter() - t0)


def fifo_latencies(arrivals, service) -> np.ndarray:
    """Time-to-converge of a one-at-a-time FIFO queue with these service
    times: finish_i = max(arrival_i, finish_{i-1}) + service_i."""
    fin, out = 0.0, []
    for a, s in zip(arrivals, service):
        fin = max(a, fin) + s
        out.append(fin - a)
    return np.array(out)


def main(argv=None) -> int:
    """python -m doci_amd.serve [--requests R] [--rate X] [--precision P]:
    the serving workload's first R requests (doci_amd.workloads) at Poisson
    rate X times mu0, the one-at-a-time plain FP64 throughput of the same
    requests solved back to back; prints solves/s and p50/p99
    time-to-converge beside those of a one-at-a-time FIFO server with the
    measured service times and the same arrivals."""
    import argparse
    from .workloads import poisson_arrivals, serving_requests
    ap = argparse.ArgumentParser(prog="python -m doci_amd.serve")
    ap.add_argument("--requests", type=int, default=500)
    ap.add_argument("--rate", type=float, default=10.0)
    ap.add_argument("--precision", default="mixed", choices=["fp64", "mixed"])
    ap.add_argument("--by-size", action="store_true",
                    help="also serve the requests from 2^20 determinants on and the smaller "
                         "ones as two separate sets, all arriving at once, and print each "
                         "set's engine time beside its one-at-a-time service time")
    a = ap.parse_args(argv)
    reqs = serving_requests(a.requests)
    ints = [_integrals(q) for q in reqs]
    first = {}
    for k, q in enumerate(reqs):
        first.setdefault(q[:2], k)
    one_at_a_time_queue([reqs[k] for k in first.values()], [0.0] * len(first),
                        integrals=[ints[k] for k in first.values()])  # per-shape tables
    service = one_at_a_time_queue(reqs, [0.0] * len(reqs), integrals=ints)["service"]
    mu0 = len(reqs) / float(service.sum())
    arr = poisson_arrivals(a.requests, a.rate * mu0)
    res = run_stream(reqs, arr, a.precision, integrals=list(ints))
    one = fifo_latencies(arr, service)
    print(f"{a.requests} requests at {a.rate} x {mu0:.1f}/s ({a.precision}): "
          f"{res['solves_per_s']:.1f} solves/s, p50 {res['p50']:.3f} s, p99 {res['p99']:.3f} s, "
          f"peak running {res['peak_running']}, all converged "
          f"{bool((res['status'] == 2).all())}; one at a time: "
          f"{a.requests / float(np.max(np.asarray(arr) + one)):.1f} solves/s, "
          f"p50 {np.percentile(one, 50):.3f} s, p99 {np.percentile(one, 99):.3f} s", flush=True)
    if a.by_size:
        n = np.array([binom(q[1], q[0]) for q in ints])
        for name, sel in (("n >= 2^20", n >= (1 << 20)), ("n < 2^20", n < (1 << 20))):
            idx = np.nonzero(sel)[0]
            if not len(idx):
                continue
            r = run_stream([reqs[i] for i in idx], [0.0] * len(idx), a.precision,
                           integrals=[ints[i] for i in idx])
            print(f"{name}: {len(idx)} requests, one at a time {float(service[idx].sum()):.3f} s "
                  f"of service, engine ({a.precision}, all at once) "
                  f"{len(idx) / r['solves_per_s']:.3f} s", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
