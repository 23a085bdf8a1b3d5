"""FRI proof-of-work grinding (fri.WithGrinding, internal/fri/fri.go:71-78,
1053-1090; transcript.go:216-365): every fold challenge's message ends with
[FSPW, bits, salt] for the smallest salt whose digest has `bits` zero low
bits.  CPU: the oracle's search against brute force, the checker's accept
and reject cases.  GPU: device-searched proofs equal the oracle's word for
word and the batched verifier checks them."""
import numpy as np
import pytest

from oracle import field as F
from oracle import prover as OP
from oracle import transcript as T
from oracle import verifier as OV
from oracle.poseidon2 import sponge1


def test_first_salt_is_the_smallest_passing_one():
    # closed definition: scan salts 0, 1, ... with the plain sponge
    t = T.Transcript(["a", "b"])
    t.bind("a", [7, 8, 9])
    t.compute("a")
    t.bind("b", list(range(8)))
    msg = t.message("b")
    for bits in (1, 4, 9):
        tt = T.Transcript(["a", "b"])
        tt.bind("a", [7, 8, 9])
        tt.compute("a")
        tt.bind("b", list(range(8)))
        v = tt.compute_grind("b", bits, chunk=64)
        got_bits, salt = tt.pow["b"]
        assert got_bits == bits
        assert v == sponge1(msg + [T.FSPW_TAG, bits, salt])
        assert v[0] % (1 << bits) == 0
        for s in range(salt):
            assert sponge1(msg + [T.FSPW_TAG, bits, s])[0] % (1 << bits) != 0, "a smaller salt passes"


def test_zero_bits_span_elements():
    # 31 bits per element (the low bits of the canonical value), element 0 first
    d = [0] * 8
    assert T.has_zero_grinding_bits(d, 31)
    d[0] = 1 << 30
    assert T.has_zero_grinding_bits(d, 30) and not T.has_zero_grinding_bits(d, 31)
    d = [0, 1] + [0] * 6
    assert T.has_zero_grinding_bits(d, 31) and not T.has_zero_grinding_bits(d, 32)


def _plonk(logn=5, seed=3):
    from loom_amd import workloads
    return workloads.plonk_instance(logn, seed)


def test_oracle_proof_with_grinding_is_checked():
    prog, tr = _plonk()
    pf, enc, _ = OP.prove(prog, tr, 8, grinding=5)
    pw = pf["fri"]["pow"]
    assert sorted(pw) == sorted("fri_fold_%d" % j for j in range(22))
    assert all(b == 5 for b, _ in pw.values())
    OV.verify(prog, 8, enc, grinding=5)
    with pytest.raises(OV.ProofError):
        OV.verify(prog, 8, enc)  # the checker without grinding expects no proofs of work
    with pytest.raises(OV.ProofError):
        OV.verify(prog, 8, enc, grinding=6)  # bit count differs


def _salt_word_index(enc, name):
    """Word index of the salt of `name` in a proof's proof-of-work list."""
    b = name.encode()
    packed = [int.from_bytes(b[i:i + 4].ljust(4, b"\0"), "little") for i in range(0, len(b), 4)]
    w = [int(x) for x in enc]
    needle = [len(b)] + packed
    for i in range(len(w) - len(needle)):
        if w[i:i + len(needle)] == needle:
            return i + len(needle) + 1
    raise AssertionError("name not found")


def test_checker_rejects_a_wrong_salt():
    prog, tr = _plonk()
    pf, enc, _ = OP.prove(prog, tr, 8, grinding=5)
    bits, salt = pf["fri"]["pow"]["fri_fold_7"]
    # another salt: its digest lacks the zero bits or changes the fold challenge
    i = _salt_word_index(enc, "fri_fold_7")
    assert F.from_mont(int(enc[i])) == salt
    bad = enc.copy()
    bad[i] = F.to_mont((salt + 1) % F.P)
    with pytest.raises(OV.ProofError):
        OV.verify(prog, 8, bad, grinding=5)


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover
    c = prover.Context(0)
    yield c
    c.close()


def _gpu_prove(ctx, prog, tr, nq, bits, copies=1):
    from loom_amd import prover
    items = [prover.Prepared(prog, tr, nq) for _ in range(copies)]
    prover.set_grinding(ctx, items, bits)
    try:
        return items, prover.prove_prepared(ctx, items)
    finally:
        pass


@pytest.mark.gpu
@pytest.mark.parametrize("shape", ["plonk", "uneven"])
def test_gpu_grinding_equals_oracle(ctx, shape):
    """One FRI level (the device derives every fold challenge) and several
    levels (the zero-bound rounds chain from the host's level gammas)."""
    from loom_amd import prover, workloads
    print("grinding %s" % shape, flush=True)
    if shape == "plonk":
        prog, tr = workloads.plonk_instance(6, 21)
    else:
        prog, tr = workloads.uneven_instance([5, 7], 22)
    _, enc, _ = OP.prove(prog, tr, 10, grinding=6)
    items, words = _gpu_prove(ctx, prog, tr, 10, 6, copies=3)
    for w in words:
        assert np.array_equal(w, enc)
    res = prover.verify_prepared(ctx, items, words)
    assert all(r == (0, "") for r in res), res
    bad = words[1].copy()
    i = _salt_word_index(bad, "fri_fold_3")
    bad[i] = F.to_mont((F.from_mont(int(bad[i])) + 1) % F.P)
    res = prover.verify_prepared(ctx, items, [words[0], bad, words[2]])
    assert res[0] == (0, "") and res[2] == (0, "") and res[1][0] != 0, res
    prover.set_grinding(ctx, items, 0)
    plain = prover.prove_prepared(ctx, items[:1])[0]
    _, enc0, _ = OP.prove(prog, tr, 10)
    assert np.array_equal(plain, enc0)


@pytest.mark.gpu
def test_gpu_grinding_16_bits(ctx):
    """16 bits (the device search spans many workgroups per instance): the
    proof is accepted by the oracle's checker and the batched verifier."""
    from loom_amd import prover, workloads
    print("grinding 16 bits", flush=True)
    prog, tr = workloads.plonk_instance(8, 23)
    items, words = _gpu_prove(ctx, prog, tr, 12, 16, copies=2)
    assert np.array_equal(words[0], words[1])
    OV.verify(prog, 12, words[0], grinding=16)
    assert all(r == (0, "") for r in prover.verify_prepared(ctx, items, words))
    prover.set_grinding(ctx, items, 0)
