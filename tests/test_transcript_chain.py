"""prover.WithTranscript / verifier.WithTranscript with a running transcript
(prover/prover.go:54-60, 196-205; verifier/verifier.go:60-65, 167-176): the
proof's challenges are registered after the caller's, so the first one's
message carries the caller's last digest after its name, like every later
challenge carries its predecessor's (internal/fiat-shamir/transcript.go:
285-315).  CPU: the oracle's message shape and checker; GPU: chained proofs
equal the oracle's and the batched verifier checks them."""
import numpy as np
import pytest

from oracle import prover as OP
from oracle import transcript as T
from oracle import verifier as OV


def _caller_digest(seed):
    """A running transcript of the caller's with one computed challenge."""
    t = T.Transcript(["caller_%d" % seed])
    t.bind("caller_%d" % seed, [seed, seed + 1, seed + 2])
    return t.compute("caller_%d" % seed)


def test_first_message_carries_the_running_digest():
    prev = _caller_digest(1)
    a = T.Transcript(["x", "y"])
    a.chain_from(prev)
    a.bind("x", [5, 6])
    assert a.message("x") == T.string_to_elements(T.FSID_TAG, "x") + prev + [5, 6]
    # the same as one transcript holding the caller's challenge first
    b = T.Transcript(["caller_1", "x", "y"])
    b.bind("caller_1", [1, 2, 3])
    b.compute("caller_1")
    b.bind("x", [5, 6])
    assert a.compute("x") == b.compute("x")


def test_chained_oracle_proof_is_checked():
    from loom_amd import workloads
    prog, tr = workloads.plonk_instance(5, 6)
    prev = _caller_digest(2)
    _, enc, _ = OP.prove(prog, tr, 6, fs_prev=prev)
    _, plain, _ = OP.prove(prog, tr, 6)
    assert not np.array_equal(enc, plain)
    OV.verify(prog, 6, enc, fs_prev=prev)
    with pytest.raises(OV.ProofError):
        OV.verify(prog, 6, enc)
    with pytest.raises(OV.ProofError):
        OV.verify(prog, 6, enc, fs_prev=_caller_digest(3))


@pytest.mark.gpu
def test_gpu_chained_proofs_equal_oracle():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, workloads
    print("chained transcripts", flush=True)
    ctx = prover.Context(0)
    try:
        inst = [workloads.plonk_instance(6, 30 + s) for s in range(3)]
        items = [prover.Prepared(p, t, 10) for p, t in inst]
        prevs = [_caller_digest(10 + s) for s in range(3)]
        words = prover.prove_prepared(ctx, items, fs_prev=prevs)
        for (p, t), pv, w in zip(inst, prevs, words):
            _, enc, _ = OP.prove(p, t, 10, fs_prev=pv)
            assert np.array_equal(w, enc)
        assert all(r == (0, "") for r in prover.verify_prepared(ctx, items, words, fs_prev=prevs))
        wrong = [prevs[0], prevs[2], prevs[1]]
        res = prover.verify_prepared(ctx, items, words, fs_prev=wrong)
        assert res[0] == (0, "") and res[1][0] != 0 and res[2][0] != 0, res
        assert all(r[0] != 0 for r in prover.verify_prepared(ctx, items, words))
        # device inputs, and a fresh transcript again afterwards
        dev = prover.DeviceInputs(items, 0)
        w2 = prover.prove_prepared(ctx, items, dev_inputs=dev, fs_prev=prevs)
        assert all(np.array_equal(a, b) for a, b in zip(w2, words))
        _, enc0, _ = OP.prove(inst[0][0], inst[0][1], 10)
        assert np.array_equal(prover.prove_prepared(ctx, items[:1])[0], enc0)
    finally:
        ctx.close()
