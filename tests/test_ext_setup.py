"""Extension-field setup columns (expr.ExtSetup; setup.Setup commits them in
the setup tree's extension rail, setup/setup.go:68-124): the C++ plan checker
accepts the plan, the oracle's proof passes its checker with the key and
fails with another key; on the device the committed root equals the oracle's
key, proofs of a batch (eager, recorded and replayed graphs) equal the
oracle's, and the device verifier accepts them."""
import numpy as np
import pytest

from loom_amd import workloads


def test_oracle_ext_setup_verifies():
    from oracle import prover as OP, verifier as OV
    prog, tr = workloads.ext_setup_instance(6, 1, 2)
    key = OP.setup_roots(prog, tr)
    _, enc, _ = OP.prove(prog, tr, 8)
    OV.verify(prog, 8, enc, setup_roots=key)
    other = OP.setup_roots(*workloads.ext_setup_instance(6, 9, 2))
    with pytest.raises(Exception):
        OV.verify(prog, 8, enc, setup_roots=other)


def test_ext_setup_plan_passes_the_plan_checker():
    from loom_amd import prover
    prog, tr = workloads.ext_setup_instance(6, 1, 2)
    it = prover.Prepared(prog, tr, 8)
    assert it.plan.setup == ["S1", "SX"]
    assert [c.size for c in it.setup] == [64, 6 * 64]
    assert prover.check_plan(it.words) > 0


@pytest.mark.gpu
def test_gpu_ext_setup_equals_oracle():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover
    from oracle import prover as OP
    nq = 12
    insts = [workloads.ext_setup_instance(8, 4, 50 + k) for k in range(3)]
    key = OP.setup_roots(*insts[0])
    items = [prover.Prepared(p, t, nq) for p, t in insts]
    want = [OP.prove(p, t, nq)[1] for p, t in insts]
    ctx = prover.Context(0)
    try:
        h = ctx.plan_handle(items[0].words)
        root = prover.set_setup(h, items[0].setup)
        assert [int(x) for x in root] == [int(x) for x in key[0]], "device setup root differs from the oracle's key"
        ctx.setups[items[0].words.tobytes()] = None
        for rep in range(3):  # eager, recorded, replayed
            print("[ext-setup] batch of 3, call %d" % rep, flush=True)
            got = prover.prove_prepared(ctx, items)
            for k in range(3):
                assert np.array_equal(got[k], want[k]), (rep, k)
        st = prover.verify_prepared(ctx, items, got, setup_roots=key)
        assert all(s == 0 for s, _ in st), st
    finally:
        ctx.close()
