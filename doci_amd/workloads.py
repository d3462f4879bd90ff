"""Seeded synthetic workloads: the batched-solve set and the serving request
stream.

Both draw from splitmix64 streams (the generator of doci_gen_integrals), so
the same instance list and arrival times come out on every machine:

* batched set: instance k (k = 0..count-1) has L uniform in {12..22} and N
  uniform in [ceil(L/4), floor(L/2)], both drawn in that order from the
  stream seeded 999; its integrals use seed 1000 + k.
* serving stream: request r has L uniform in {14..30} from the stream seeded
  99999, N = round(L/3), integrals seed 100000 + r; inter-arrival times are
  exponential (Poisson arrivals) with rate lam from the stream seeded 99998.

A uniform integer in [lo, hi] is lo + (draw mod (hi - lo + 1)); a uniform
real in (0, 1] is ((draw >> 11) + 1) / 2^53.
"""
from __future__ import annotations

import math

_M64 = (1 << 64) - 1


def splitmix64(seed: int):
    s = seed & _M64
    while True:
        s = (s + 0x9E3779B97F4A7C15) & _M64
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        yield z ^ (z >> 31)


def batched_instances(count: int = 512):
    """[(L, N, seed)] of the batched-solve set."""
    g = splitmix64(999)
    out = []
    for k in range(count):
        L = 12 + next(g) % 11
        lo, hi = math.ceil(L / 4), L // 2
        N = lo + next(g) % (hi - lo + 1)
        out.append((L, N, 1000 + k))
    return out


def serving_requests(count: int = 2000):
    """[(L, N, seed)] of the serving stream (N = round(L/3), half-up)."""
    g = splitmix64(99999)
    out = []
    for r in range(count):
        L = 14 + next(g) % 17
        N = int(math.floor(L / 3 + 0.5))
        out.append((L, N, 100000 + r))
    return out


def poisson_arrivals(count: int, lam: float):
    """Arrival times (s) of `count` requests at rate lam (1/s), first at 0."""
    g = splitmix64(99998)
    t, out = 0.0, []
    for r in range(count):
        out.append(t)
        u = ((next(g) >> 11) + 1) / float(1 << 53)
        t += -math.log(u) / lam
    return out
