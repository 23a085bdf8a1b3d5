"""Serving loop (lg_server_*): continuous batching with admission by HBM
footprint.  Every proof that comes out of the mix must equal the proof of
the same instance proved alone; a failing request fails alone."""
import sys
import threading
import time

import numpy as np
import pytest


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _requests():
    from loom_amd import workloads, prover
    reqs = []
    reqs.append([prover.Prepared(*workloads.plonk_instance(8, 500 + s), 12) for s in range(3)])
    reqs.append([prover.Prepared(*workloads.plonk_instance(10, 510), 12)])
    reqs.append([prover.Prepared(*workloads.uneven_instance([6, 8], 520 + s), 12) for s in range(2)])
    reqs.append([prover.Prepared(*workloads.uneven_range_instance([7], 530), 12)])
    reqs.append([prover.Prepared(*workloads.plonk_instance(8, 540 + s), 12) for s in range(2)])
    reqs.append([prover.Prepared(*workloads.wide_instance(9, 550, 2, 2), 12)])
    reqs.append([prover.Prepared(*workloads.plonk_instance(12, 560 + s), 12) for s in range(4)])
    reqs.append([prover.Prepared(*workloads.plonk_instance(8, 570), 12)])
    return reqs


def test_plan_footprint_estimate_scales():
    from loom_amd import workloads, plan, serve
    small = serve.plan_bytes_per_instance(plan.Plan(workloads.plonk_instance(8, 0)[0], 8).words())
    big = serve.plan_bytes_per_instance(plan.Plan(workloads.plonk_instance(12, 0)[0], 8).words())
    assert small > 0 and 12 < big / small < 20


@pytest.mark.gpu
def test_mixed_requests_equal_alone():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, serve
    reqs = _requests()
    log("serving: %d requests, %d instances" % (len(reqs), sum(len(r) for r in reqs)))
    ctx = prover.Context(0)
    alone = [[prover.prove_prepared(ctx, [it])[0] for it in r] for r in reqs]
    ctx.close()
    srv = serve.Server(0, slots=3, max_batch=16)
    tickets = [None] * len(reqs)

    def submitter(idx):
        for i in idx:
            tickets[i] = srv.submit(reqs[i])
            time.sleep(0.001)

    th = [threading.Thread(target=submitter, args=(list(range(k, len(reqs), 2)),)) for k in range(2)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    for i, t in enumerate(tickets):
        got = srv.wait(t)
        assert len(got) == len(reqs[i])
        for a, b in zip(got, alone[i]):
            assert np.array_equal(a, b), "request %d differs from its proofs alone" % i
    st = srv.stats()
    log("serving stats %s" % st)
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_running"] >= 2, "no two requests overlapped"
    srv.close()


@pytest.mark.gpu
def test_budget_serialises_and_failure_is_isolated():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, serve, workloads, _lib
    log("serving: 1-byte budget, one failing request")
    good = [[prover.Prepared(*workloads.plonk_instance(7, 600 + s), 8)] for s in range(3)]
    prog, tr = workloads.uneven_range_instance([5], 3)
    tr = dict(tr)
    tr["L_0"] = tr["L_0"].copy()
    tr["L_0"][2] = 999
    bad = [prover.Prepared(prog, tr, 8)]
    srv = serve.Server(0, slots=3, budget_bytes=1, max_batch=1)
    tk = [srv.submit(r) for r in good[:1]] + [srv.submit(bad)] + [srv.submit(r) for r in good[1:]]
    ctx = prover.Context(0)
    for r, t in zip([good[0], None, good[1], good[2]], tk):
        if r is None:
            with pytest.raises(_lib.LoomGpuError, match="not in the range table"):
                srv.wait(t)
            continue
        got = srv.wait(t)
        assert np.array_equal(got[0], prover.prove_prepared(ctx, r)[0])
    st = srv.stats()
    log("serving stats %s" % st)
    assert st["peak_running"] == 1 and st["failed"] == 1 and st["completed"] == 3
    ctx.close()
    srv.close()


def _stream(seed):
    """A seeded Poisson request stream of the serving shape: per request 1-40
    plonk-like instances (geometric, skewed toward small counts) of one size
    with log2 N geometric on 12..16, arrivals at 400 requests/s; plus one
    request at 2^18 and one at 2^20 (the tail of the size distribution).
    Returns (requests, [(log2 N, instance seeds)] per request, gaps)."""
    from loom_amd import workloads, prover
    rng = np.random.Generator(np.random.PCG64(seed))
    meta = []
    for r in range(22):
        ln = 12 + min(int(rng.geometric(0.55)) - 1, 4)
        cnt = min(int(rng.geometric(0.15)), 40)
        if ln >= 15:
            cnt = min(cnt, 6)
        meta.append((ln, [10_000 * r + k for k in range(cnt)]))
    meta.insert(7, (18, [777]))
    meta.insert(15, (20, [778]))
    reqs = [[prover.Prepared(*workloads.plonk_instance(ln, sd), 200) for sd in seeds] for ln, seeds in meta]
    gaps = rng.exponential(1 / 400.0, size=len(reqs))
    return reqs, meta, gaps


@pytest.mark.gpu
def test_poisson_stream_served_equals_alone():
    """The serving workload: a Poisson stream of requests of 1-40 instances
    and 2^12..2^20 rows over 3 slots with merging on.  Every served proof
    equals the same instance proved alone; one 2^12 instance equals the
    oracle; the 2^16-and-larger proofs pass the proof checker."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, serve, workloads
    from oracle import prover as OP, verifier as V
    reqs, meta, gaps = _stream(4242)
    log("poisson stream: %d requests, %d instances, (log2 N, count): %s"
        % (len(reqs), sum(len(r) for r in reqs), [(ln, len(sd)) for ln, sd in meta]))
    srv = serve.Server(0, slots=3, max_batch=64)
    tickets = []
    t0 = time.time()
    for r, g in zip(reqs, gaps):
        time.sleep(float(g))
        tickets.append(srv.submit(r))
    got = [srv.wait(t) for t in tickets]
    dt = time.time() - t0
    st = srv.stats()
    log("served in %.2f s, stats %s" % (dt, st))
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_running"] >= 2, "no two requests overlapped"
    srv.close()
    ctx = prover.Context(0)
    try:
        for i, r in enumerate(reqs):
            for k in range(len(r)):
                assert np.array_equal(got[i][k], prover.prove_prepared(ctx, [r[k]])[0]), \
                    "request %d instance %d differs from its proof alone" % (i, k)
        log("every served proof equals its proof alone")
    finally:
        ctx.close()
    i = next(i for i, (ln, _) in enumerate(meta) if ln == 12)
    prog, tr = workloads.plonk_instance(12, meta[i][1][0])
    _, enc, _ = OP.prove(prog, tr, 200)
    assert np.array_equal(got[i][0], enc), "2^12 served proof differs from the oracle"
    log("request %d instance 0 equals the oracle" % i)
    for i, (ln, _) in enumerate(meta):
        if ln >= 16:
            V.verify(reqs[i][0].plan.program, 200, got[i][0])
            log("request %d (2^%d rows) passes the proof checker" % (i, ln))


@pytest.mark.gpu
def test_invalid_instance_fails_alone_in_a_merge():
    """Requests of one plan queued behind busy slots are merged into one
    prove; the request carrying an invalid instance (a range-lookup value
    outside the table, or a copy-permutation entry outside the cells) fails
    alone and its merge-mates get their proofs, equal to the proofs alone."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, serve, workloads, _lib
    # three 4-instance requests occupy the three slots (max_batch 4 keeps them
    # apart); the 1-instance burst behind them is merged when a slot frees
    blockers = [[prover.Prepared(*workloads.plonk_instance(16, 900 + 4 * j + k), 200) for k in range(4)]
                for j in range(3)]

    def range_req(seed, bad=False):
        prog, tr = workloads.uneven_range_instance([12], seed)
        if bad:
            tr = dict(tr)
            tr["L_0"] = tr["L_0"].copy()
            tr["L_0"][5] = 300
        return [prover.Prepared(prog, tr, 200)]

    def perm_req(seed, bad=False):
        it = prover.Prepared(*workloads.plonk_instance(12, seed), 200)
        if bad:
            it.perms = [p.copy() for p in it.perms]
            it.perms[0][3] = len(it.perms[0]) + 17
        return [it]

    for kind, make in (("lookup", range_req), ("permutation", perm_req)):
        burst = [make(50), make(51, bad=True), make(52), make(53)]
        srv = serve.Server(0, slots=3, max_batch=4)
        tb = [srv.submit(r) for r in blockers]
        tk = [srv.submit(r) for r in burst]
        ctx = prover.Context(0)
        try:
            for i, (r, t) in enumerate(zip(burst, tk)):
                if i == 1:
                    with pytest.raises(_lib.LoomGpuError, match="instance 0"):
                        srv.wait(t)
                    with pytest.raises(_lib.LoomGpuError):
                        prover.prove_prepared(ctx, r)
                    continue
                got = srv.wait(t)
                assert np.array_equal(got[0], prover.prove_prepared(ctx, r)[0]), (kind, i)
            for t in tb:
                srv.wait(t)
            st = srv.stats()
            log("%s: stats %s" % (kind, st))
            assert st["failed"] == 1 and st["completed"] == len(burst) + len(blockers) - 1
            assert st["merged"] >= len(burst) - 1, "the burst was not merged into one prove"
        finally:
            ctx.close()
            srv.close()


@pytest.mark.gpu
def test_batch_with_invalid_instance():
    """lg_prove on a batch with one invalid instance: -3, the failed instance
    named, the others' proofs equal to their proofs alone."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, workloads, _lib
    items = [prover.Prepared(*workloads.uneven_range_instance([9], 60 + k), 16) for k in range(5)]
    bad = items[2]
    bad.columns = [c.copy() for c in bad.columns]
    li = bad.plan.input_names().index("L_0")
    bad.columns[li][7] = prover.to_mont([1000])[0]
    ctx = prover.Context(0)
    try:
        with pytest.raises(_lib.LoomGpuError, match="instance 2: lookup"):
            prover.prove_prepared(ctx, items)
        res = prover.prove_prepared(ctx, items, allow_failures=True)
        assert res[2] is None
        for k in (0, 1, 3, 4):
            assert np.array_equal(res[k], prover.prove_prepared(ctx, [items[k]])[0]), k
    finally:
        ctx.close()
