"""Batched multi-instance Davidson (doci_batch_*): every instance's result
against an independent single-instance solve of the same integrals, the
oracle on small cases, deflation, mixed precision, HIP-graph replay, and
continuous batching (an instance gives the same result alone and in a mix)."""
import numpy as np
import pytest

import doci_amd as da
import oracle as O

pytestmark = pytest.mark.gpu

CASES = [(8, 2, 11), (10, 3, 12), (12, 4, 13), (12, 6, 14), (14, 5, 15), (16, 6, 16),
         (13, 4, 17), (18, 5, 18)]


def _single(L, N, seed, **kw):
    with da.Operator.synthetic(L, N, seed) as op:
        op.set_kernel("star-checked")  # index arithmetic of this shape proven in bounds first
        op.sigma(da.gen_vector(op.n, 3))
        assert op.check_violations() == 0
        op.set_kernel("star")
        r = op.solve("D", want_x=True, **kw)
    return r


@pytest.fixture(scope="module")
def singles():
    return {c: _single(*c) for c in CASES}


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_fp64_matches_single_solves(singles, graphs):
    print(f"[case] batch fp64 graphs={graphs}", flush=True)
    with da.Batch(use_graphs=graphs, tile_rows=1024) as b:
        ids = [b.add_synthetic(L, N, s) for L, N, s in CASES]
        b.run()
        st = b.stats()
        assert (st["graph_launches"] > 0) == graphs
        for i, c in zip(ids, CASES):
            r, x = b.result(i, want_x=True)
            ref = singles[c]
            assert r.status == 2, (c, r.status)
            # same algorithm, different rounding (classical Gram-Schmidt twice vs
            # the single solve's measured-loss loop): a residual that lands next
            # to TOL may take one iteration more or less
            assert abs(r.iterations - ref.info.iterations) <= 1, c
            assert r.residual <= 1e-8
            assert abs(r.energy - ref.energy) <= 1e-9, c
            assert abs(np.linalg.norm(x) - 1) <= 1e-10
            assert abs(abs(x @ ref.x) - 1) <= 1e-8, c


def test_batch_against_oracle(fcidump_dir):
    """Two reference inputs in one batch against the oracle's Davidson."""
    names = ["H2O_631G", "lili_321g"]
    print("[case] batch vs oracle", flush=True)
    with da.Batch(tile_rows=512) as b:
        ids = []
        for nm in names:
            L, Na, Nb, core, one, two = da.read_fcidump(str(fcidump_dir / f"{nm}.FCIDUMP"))
            ids.append(b.add(L, Na, core, one, two))
        b.run()
        for i, nm in zip(ids, names):
            P = O.Problem(str(fcidump_dir / f"{nm}.FCIDUMP"))
            D = P.diagonal()
            orc = P.davidson(O.hf_guess(D), D=D)
            r, x = b.result(i, want_x=True)
            assert abs(r.energy - orc["energy"]) <= 1e-9, nm
            assert abs(abs(x @ orc["x"]) - 1) <= 1e-8, nm


def test_batch_deflation_matches_single():
    """max_vec 4 / keep 2 forces deflation every few iterations."""
    print("[case] batch deflation", flush=True)
    cases = [(14, 5, 15), (16, 6, 16), (12, 4, 13)]
    refs = [_single(*c, keep=2, max_vec=4) for c in cases]
    with da.Batch(keep=2, max_vec=4, tile_rows=512) as b:
        ids = [b.add_synthetic(*c) for c in cases]
        b.run()
        for i, c, ref in zip(ids, cases, refs):
            r, x = b.result(i, want_x=True)
            assert r.status == 2 and r.iterations > 4, c
            assert abs(r.energy - ref.energy) <= 1e-9, c
            assert abs(abs(x @ ref.x) - 1) <= 1e-8, c


def test_batch_mixed_precision(singles):
    print("[case] batch mixed", flush=True)
    with da.Batch(tile_rows=1024) as b:
        ids = [b.add_synthetic(L, N, s, precision="mixed") for L, N, s in CASES]
        b.run()
        for i, c in zip(ids, CASES):
            r, x = b.result(i, want_x=True)
            ref = singles[c]
            assert r.status == 2 and r.iterations_fp32 > 0, c
            assert abs(r.energy - ref.energy) <= 1e-9, c
            assert abs(abs(x @ ref.x) - 1) <= 1e-8, c
            assert r.residual <= 1e-8


def test_batch_alone_equals_in_mix():
    """Continuous batching: an instance admitted into a running mix (other
    instances joining and leaving around it, FP64 and mixed) returns the same
    result as the same instance solved alone: same iteration count, energy
    within 1e-12 Eh, eigenvector within 1e-13 (max abs).  Every reduction is
    per instance in a fixed order; only the launch instantiation (sized by the
    largest basis in the mix) changes."""
    print("[case] batch alone vs in mix", flush=True)
    probe = (16, 6, 16)
    with da.Batch(tile_rows=1024) as b:
        i = b.add_synthetic(*probe)
        b.run()
        alone, xa = b.result(i, want_x=True)
    with da.Batch(tile_rows=1024) as b:
        first = [b.add_synthetic(L, N, s) for L, N, s in CASES[:4]]
        b.step()
        b.step()
        i = b.add_synthetic(*probe)        # joins a running mix
        b.step()
        late = [b.add_synthetic(L, N, s, precision="mixed") for L, N, s in CASES[4:]]
        b.run()
        mixr, xm = b.result(i, want_x=True)
        for j in first + late:
            assert b.result(j)[0].status == 2
    assert mixr.iterations == alone.iterations
    assert abs(mixr.energy - alone.energy) <= 1e-12
    assert np.abs(xa - xm).max() <= 1e-13


def test_batch_memory_budget_admission():
    """A budget that fits two instances at a time: all finish, never more
    than two run at once."""
    print("[case] batch admission by memory", flush=True)
    o = da.Batch(tile_rows=512)
    fp = o.footprint(14, 5)
    o.close()
    with da.Batch(tile_rows=512, mem_budget=2 * fp + fp // 2) as b:
        ids = [b.add_synthetic(14, 5, 100 + k) for k in range(5)]
        peak = 0
        while True:
            run, wait = b.step()
            peak = max(peak, run)
            if not run and not wait:
                break
        assert peak <= 2
        for i in ids:
            assert b.result(i)[0].status == 2


def test_batch_staggered_admission_matches_all_at_once():
    """Serving pattern: instances join one or two per iteration while others
    finish (tile layouts change every few iterations); every result equals the
    all-at-once batch within the solver tolerance."""
    print("[case] batch staggered admission", flush=True)
    cases = [(12 + k % 7, 3 + k % 4, 500 + k) for k in range(24)]
    with da.Batch(tile_rows=512) as b:
        ids = [b.add_synthetic(*c) for c in cases]
        b.run()
        ref = [b.result(i)[0].energy for i in ids]
    with da.Batch(tile_rows=512) as b:
        ids = []
        for k, c in enumerate(cases):
            ids.append(b.add_synthetic(*c, precision="mixed" if k % 3 == 0 else "fp64"))
            if k % 2:
                b.step()
        b.run()
        for i, c, e in zip(ids, cases, ref):
            r = b.result(i)[0]
            assert r.status == 2 and r.residual <= 1e-8, c
            assert abs(r.energy - e) <= 1e-9, (c, r.energy - e)


def test_batch_large_instances_use_star_kernel():
    """Instances from 2^20 determinants on take the star-kernel sigma inside
    the batch (operands read on the device); FP64 and mixed, next to small
    lane-sigma instances, against independent single solves."""
    print("[case] batch large instances", flush=True)
    cases = [(23, 10, 71), (24, 9, 72), (14, 5, 73)]
    refs = [_single(*c) for c in cases]
    assert da.binom(10, 23) >= 1 << 20 and da.binom(9, 24) >= 1 << 20
    # tile rows 512: the two large instances need more rows per tile than that
    # (at most 2048 tiles each), which the engine rounds to a multiple of 512
    for prec, rows in (("fp64", 2048), ("mixed", 2048), ("fp64", 512), ("mixed", 512)):
        with da.Batch(tile_rows=rows) as b:
            ids = [b.add_synthetic(*c, precision=prec) for c in cases]
            b.run()
            for i, c, ref in zip(ids, cases, refs):
                r, x = b.result(i, want_x=True)
                assert r.status == 2 and r.residual <= 1e-8, (prec, rows, c)
                assert abs(r.energy - ref.energy) <= 1e-9, (prec, rows, c, r.energy - ref.energy)
                assert abs(abs(x @ ref.x) - 1) <= 1e-8, (prec, rows, c)


def test_batch_alone_equals_in_mix_star_instance():
    """Continuous batching with a star-kernel instance (2^20 determinants or
    more): alone and admitted into a running mix of lane- and star-sigma
    instances, FP64 and mixed.  The star pass sums a centre's hops with LDS
    atomics, so the order of a few FP64 additions varies from run to run
    (about 1e-16 relative per sigma): same iteration count, energy within
    1e-11 Eh, eigenvector within 1e-9 (max abs) and overlap 1 - 1e-12."""
    print("[case] batch alone vs in mix, star instance", flush=True)
    probe = (23, 10, 71)
    assert da.binom(10, 23) >= 1 << 20
    for prec in ("fp64", "mixed"):
        with da.Batch(tile_rows=1024) as b:
            i = b.add_synthetic(*probe, precision=prec)
            b.run()
            alone, xa = b.result(i, want_x=True)
        with da.Batch(tile_rows=1024) as b:
            first = [b.add_synthetic(L, N, s) for L, N, s in CASES[:3]] + \
                [b.add_synthetic(24, 9, 72)]
            b.step()
            i = b.add_synthetic(*probe, precision=prec)  # joins a running mix
            b.step()
            late = [b.add_synthetic(L, N, s, precision="mixed") for L, N, s in CASES[5:]]
            b.run()
            mixr, xm = b.result(i, want_x=True)
            for j in first + late:
                assert b.result(j)[0].status == 2
        assert alone.status == 2 and mixr.status == 2
        assert mixr.iterations == alone.iterations, prec
        assert abs(mixr.energy - alone.energy) <= 1e-11, (prec, mixr.energy - alone.energy)
        assert np.abs(xa - xm).max() <= 1e-9, prec
        assert abs(abs(xa @ xm) - 1) <= 1e-12, prec


def test_batch_engine_reused_across_batches(singles):
    """One engine runs three batches back to back (FP64, mixed, FP64), as a
    long-lived service does: the arena and the pinned admission / restart
    lists are reused, and every batch gives the results of a fresh engine."""
    print("[case] batch engine reuse", flush=True)
    with da.Batch(tile_rows=1024) as fresh:
        ids = [fresh.add_synthetic(L, N, s, precision="mixed") for L, N, s in CASES]
        fresh.run()
        ref_mixed = [fresh.result(i)[0].energy for i in ids]
    with da.Batch(tile_rows=1024) as b:
        for prec in ("fp64", "mixed", "fp64"):
            ids = [b.add_synthetic(L, N, s, precision=prec) for L, N, s in CASES]
            b.run()
            for k, (i, c) in enumerate(zip(ids, CASES)):
                r, x = b.result(i, want_x=True)
                assert r.status == 2 and r.residual <= 1e-8, (prec, c)
                assert abs(r.energy - singles[c].energy) <= 1e-9, (prec, c)
                assert abs(abs(x @ singles[c].x) - 1) <= 1e-8, (prec, c)
                if prec == "mixed":
                    assert r.iterations_fp32 > 0
                    assert abs(r.energy - ref_mixed[k]) <= 1e-11, c


def test_batch_add_pipelined_matches_all_at_once(singles):
    """Batch.add_pipelined: instances queued in chunks with engine steps in
    between give every instance the result of its single solve."""
    print("[case] batch add_pipelined", flush=True)
    items = [(L, N) + tuple(da.gen_integrals(L, s)) for L, N, s in CASES]
    with da.Batch(tile_rows=1024) as b:
        ids = b.add_pipelined(items, chunk=3)
        b.run()
        for i, c in zip(ids, CASES):
            r, x = b.result(i, want_x=True)
            assert r.status == 2 and r.residual <= 1e-8, c
            assert abs(r.energy - singles[c].energy) <= 1e-9, c
            assert abs(abs(x @ singles[c].x) - 1) <= 1e-8, c


def test_batch_add_many_equals_add():
    """doci_batch_add_many (the integrals of 40 instances transformed on
    three threads) queues the same instances as 40 doci_batch_add calls:
    same iterations and energies; a bad instance in the list queues none."""
    print("[case] batch add_many", flush=True)
    rng = np.random.default_rng(5)
    shapes = [(int(L), int(rng.integers(2, L // 2 + 1))) for L in rng.integers(8, 15, 40)]
    items = [(L, N) + tuple(da.gen_integrals(L, 100 + k)) for k, (L, N) in enumerate(shapes)]
    res = []
    for many in (False, True):
        with da.Batch(tile_rows=512) as b:
            ids = b.add_many(items) if many else [b.add(*it) for it in items]
            b.run()
            res.append([b.result(i)[0] for i in ids])
    for k, (r0, r1) in enumerate(zip(*res)):
        assert r0.status == 2 and r1.status == 2, k
        assert r0.iterations == r1.iterations, k
        assert abs(r0.energy - r1.energy) <= 1e-12, k
    with da.Batch() as b:
        bad = items[:3] + [(4, 5) + tuple(da.gen_integrals(4, 1))]
        with pytest.raises(da.DociError):
            b.add_many(bad)
        assert b.add_many(items[:2]) == [0, 1]
