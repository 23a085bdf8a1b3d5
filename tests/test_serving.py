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
