"""General lookups on the device (arguments/lookup.go: LookupTuple,
CLookupTuple with selectors, LookupUnion across modules): the multiplicity
hash-table kernels and the logup steps give proofs equal to the oracle's, in
batches; the device verifier accepts them; an instance whose source tuple is
in no target fails alone."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KINDS = ["tuple", "cond", "union", "wide"]


def log(msg):
    print("[gpu-lookups] " + msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover
    c = prover.Context(0)
    yield c
    c.close()


def _oracle(prog, tr, nq):
    from oracle import prover as OP
    return OP.prove(prog, tr, nq)[1]


@pytest.mark.parametrize("kind", KINDS)
def test_general_lookup_proofs_equal_oracle(ctx, kind):
    from loom_amd import workloads, prover
    log("general lookups %s, batch of 3 (2^7 sources, 2^5 targets)" % kind)
    inst = [workloads.lookup_instance(kind, 40 + s, log_n=7, log_t=5) for s in range(3)]
    items = [prover.Prepared(p, t, 12) for p, t in inst]
    got = prover.prove_prepared(ctx, items)
    for (prog, tr), w in zip(inst, got):
        assert np.array_equal(w, _oracle(prog, tr, 12))
    st = prover.verify_prepared(ctx, items, got)
    assert all(s == 0 for s, _ in st), st


@pytest.mark.parametrize("kind,log_n,log_t", [("tuple", 13, 12), ("cond", 12, 10), ("union", 14, 8), ("wide", 12, 10)])
def test_general_lookup_larger(ctx, kind, log_n, log_t):
    """Hash tables of 2^13..2^15 slots, many sources per target tuple."""
    from loom_amd import workloads, prover
    log("general lookups %s at 2^%d sources, 2^%d targets" % (kind, log_n, log_t))
    prog, tr = workloads.lookup_instance(kind, 77, log_n=log_n, log_t=log_t)
    w = prover.prove(ctx, prog, tr, 16)
    assert np.array_equal(w, _oracle(prog, tr, 16))


@pytest.mark.parametrize("kind", KINDS)
def test_missing_tuple_fails_alone(ctx, kind):
    from loom_amd import workloads, prover, _lib
    log("general lookups %s, bad instance in a batch" % kind)
    good = [workloads.lookup_instance(kind, 90 + s, log_n=6, log_t=4) for s in range(2)]
    bad = workloads.lookup_instance(kind, 95, log_n=6, log_t=4, bad=True)
    items = [prover.Prepared(*good[0], 8), prover.Prepared(*bad, 8), prover.Prepared(*good[1], 8)]
    with pytest.raises(_lib.LoomGpuError, match="instance 1.*not in any target"):
        prover.prove_prepared(ctx, items)
    res = prover.prove_prepared(ctx, items, allow_failures=True)
    assert res[1] is None
    for k, (prog, tr) in ((0, good[0]), (2, good[1])):
        assert np.array_equal(res[k], _oracle(prog, tr, 8))


def test_general_lookup_graph_replay(ctx):
    """A tuple lookup within one module (one DEEP class): the device phase
    with the general multiplicity kernels is recorded into a HIP graph on the
    second call and replayed on the third; every call equals the oracle."""
    from loom_amd import workloads, prover
    inst = [workloads.lookup_instance("self", 70 + s, log_n=7) for s in range(2)]
    items = [prover.Prepared(p, t, 12) for p, t in inst]
    want = [_oracle(p, t, 12) for p, t in inst]
    for rep in range(3):
        log("general lookups, one size class, call %d (eager, recorded, replayed)" % rep)
        got = prover.prove_prepared(ctx, items)
        for k in range(2):
            assert np.array_equal(got[k], want[k]), (rep, k)
    st = prover.graph_stats(ctx, items)
    assert st["fallbacks"] == 0 and st["replays"] >= 1, st
