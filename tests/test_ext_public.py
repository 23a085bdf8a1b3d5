"""Extension-field public inputs (public.Entry.SetExt with
expr.PublicInputExt, public/public_inputs.go:30-101, prover/prover.go:236-270):
entries of six limbs bound to the first challenge as (row, field 1, limbs),
base entries of an extension column lifted, the column materialised with all
six limbs.

CPU: the oracle's proof passes its checker with the statement's values and
fails with a changed limb or a changed field (a base entry given as the same
value lifted to six limbs binds different transcript words); a relation that
fails at a public row is refused; an extension entry of a base-declared
input makes its column an extension column; the plans pass the C++ plan
checker.
GPU: proofs equal the oracle's (batch of two statements), and the device
verifier accepts them and refuses a changed limb."""
import numpy as np
import pytest

from loom_amd import workloads

P = 2**31 - 2**24 + 1


def test_oracle_ext_public_inputs_verify():
    from oracle import prover as OP, verifier as OV
    prog, tr, pi = workloads.ext_public_instance(6, 3)
    _, enc, _ = OP.prove(prog, tr, 8, public_inputs=pi)
    OV.verify(prog, 8, enc, public_inputs=pi)
    mod, ents = pi["PX"]
    r0, v0 = ents[0]
    wrong = dict(pi)
    wrong["PX"] = (mod, [(r0, [(v0[0] + 1) % P] + v0[1:])] + ents[1:])
    with pytest.raises(Exception):
        OV.verify(prog, 8, enc, public_inputs=wrong)
    # the last PX entry is a base entry; the same value as six limbs (lifted)
    # is the same column but different transcript words (field 1)
    rl, vl = ents[-1]
    lifted = dict(pi)
    lifted["PX"] = (mod, ents[:-1] + [(rl, [vl, 0, 0, 0, 0, 0])])
    with pytest.raises(Exception):
        OV.verify(prog, 8, enc, public_inputs=lifted)


def test_failing_relation_at_a_public_row_is_refused():
    from oracle import prover as OP, verifier as OV
    prog, tr, pi = workloads.ext_public_instance(6, 3, bad=True)
    _, enc, _ = OP.prove(prog, tr, 8, public_inputs=pi)
    with pytest.raises(Exception, match="AIR identity"):
        OV.verify(prog, 8, enc, public_inputs=pi)


def _pb_ext(pi):
    """The statement with PB's first entry (an input declared with
    expr.PublicInput) an extension value."""
    out = dict(pi)
    mod, ents = pi["PB"]
    out["PB"] = (mod, [(ents[0][0], [1, 2, 3, 4, 5, 6])] + ents[1:])
    return out


def test_extension_entry_of_a_base_input_makes_an_extension_column():
    """prover.go:241-268 materialises a public column in the field of its
    entries, whatever the declaration: PB with one E6 entry is an extension
    column whose other entry is lifted; the proof passes the checker with
    that statement and fails with the entry's limbs changed, or with the
    original base statement."""
    from loom_amd import prover
    from oracle import prover as OP, verifier as OV
    prog, tr, pi = workloads.ext_public_instance(6, 3)
    pe = _pb_ext(pi)
    it = prover.Prepared(prog, tr, 8, public_inputs=pe)
    assert prover.check_plan(it.words) > 0
    assert [c.size for c in it.columns[-2:]] == [6 + 1, 6 * 3 + 1]
    _, enc, _ = OP.prove(prog, tr, 8, public_inputs=pe)
    OV.verify(prog, 8, enc, public_inputs=pe)
    mod, ents = pe["PB"]
    wrong = dict(pe)
    wrong["PB"] = (mod, [(ents[0][0], [1, 2, 3, 4, 5, 7])] + ents[1:])
    for bad in (wrong, pi):
        with pytest.raises(Exception):
            OV.verify(prog, 8, enc, public_inputs=bad)


def test_ext_public_plan_passes_the_plan_checker():
    from loom_amd import prover
    prog, tr, pi = workloads.ext_public_instance(6, 3)
    it = prover.Prepared(prog, tr, 8, public_inputs=pi)
    assert it.plan.public == ["PB", "PX"]
    assert prover.check_plan(it.words) > 0
    # value words: 1 per base entry, 6 per extension entry, in row order
    assert [c.size for c in it.columns[-2:]] == [2, 6 * 3 + 1]


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover
    c = prover.Context(0)
    yield c
    c.close()


@pytest.mark.gpu
def test_gpu_ext_public_inputs_equal_oracle_and_verify(ctx):
    from loom_amd import prover
    from oracle import prover as OP
    print("[ext-public] 2 statements of 2^7 rows", flush=True)
    insts = [workloads.ext_public_instance(7, 40 + k) for k in range(2)]
    items = [prover.Prepared(p, t, 12, public_inputs=pi) for p, t, pi in insts]
    got = prover.prove_prepared(ctx, items)
    for (prog, tr, pi), w in zip(insts, got):
        assert np.array_equal(w, OP.prove(prog, tr, 12, public_inputs=pi)[1])
    st = prover.verify_prepared(ctx, items, got)
    assert all(s == 0 for s, _ in st), st
    # a changed limb of an extension entry in the verifier's values
    pubs = [[c.copy() for c in it.columns[-2:]] for it in items]
    pubs[1][1][3] ^= 1
    st = prover.verify(ctx, items[0].words, got, publics=pubs)
    assert st[0][0] == 0 and st[1][0] != 0, st


@pytest.mark.gpu
def test_gpu_extension_entry_of_a_base_input(ctx):
    """A base-declared input with an E6 entry: the proof equals the oracle's
    and the device verifier accepts it."""
    from loom_amd import prover
    from oracle import prover as OP
    print("[ext-public] base-declared input with an E6 entry, 2^7 rows", flush=True)
    prog, tr, pi = workloads.ext_public_instance(7, 44)
    pe = _pb_ext(pi)
    items = [prover.Prepared(prog, tr, 12, public_inputs=pe)]
    got = prover.prove_prepared(ctx, items)
    assert np.array_equal(got[0], OP.prove(prog, tr, 12, public_inputs=pe)[1])
    assert prover.verify_prepared(ctx, items, got) == [(0, "")]
