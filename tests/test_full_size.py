"""Full-size proofs checked by the proof checker (GPU).  The oracle cannot
recompute these sizes in test time, so the checks are the size-independent
ones a verifier makes (oracle/verifier.py): transcript replay, AIR identity
at zeta, logup buses, every opening against its commitment, DEEP against FRI
and FRI fold consistency -- plus determinism of the prover."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd.prover import Context
    c = Context(0)
    yield c
    c.close()


def test_batch_40x2e14_checks(ctx):
    """The bench shape: 40 plonk-like 2^14 instances, 200 queries, batched;
    three of the proofs are checked in full."""
    from loom_amd import workloads, prover
    from oracle import verifier as V
    log("40 x 2^14 batched, checking 3 proofs")
    inst = [workloads.plonk_instance(14, s) for s in range(40)]
    items = [prover.Prepared(p, t, 200) for p, t in inst]
    words = prover.prove_prepared(ctx, items)
    for b in (0, 17, 39):
        V.verify(inst[b][0], 200, words[b])


def test_wide_2e22_checks(ctx):
    """One 2^22-row, 96-column Fibonacci/cubic proof, 200 queries: proved
    twice (identical), then checked."""
    from loom_amd import workloads, prover
    from oracle import verifier as V
    log("wide 2^22 x 96: generating")
    prog, tr = workloads.wide_instance(22, 0)
    item = prover.Prepared(prog, tr, 200)
    del tr
    log("wide 2^22 x 96: proving")
    a = prover.prove_prepared(ctx, [item])[0]
    b = prover.prove_prepared(ctx, [item])[0]
    assert np.array_equal(a, b)
    log("wide 2^22 x 96: checking")
    V.verify(prog, 200, a)


def test_uneven64_with_lookups_checks(ctx):
    """64 plonk-like modules of 2^10..2^20 rows with range lookups, one proof,
    200 queries, checked."""
    from loom_amd import workloads, prover
    from oracle import verifier as V
    sizes = workloads.uneven_sizes(64, 0)
    log("uneven64 %s: generating" % sizes)
    prog, tr = workloads.uneven_range_instance(sizes, 0)
    item = prover.Prepared(prog, tr, 200)
    del tr
    log("uneven64: proving")
    w = prover.prove_prepared(ctx, [item])[0]
    log("uneven64: checking")
    V.verify(prog, 200, w)
