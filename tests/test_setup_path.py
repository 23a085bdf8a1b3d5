"""The setup path and public inputs.

A plonk-like module with its selectors as setup columns
(PrepareIthPlonkWithSetup, integration_test/gnark_plonk/plonk_gnark_utils.go:
217-232) and public-input rows:

* setup.Setup commits the selectors into one setup tree (setup/setup.go:
  68-124); its root is the verification key, bound to the first challenge
  after the module sizes (prover/prover.go:224-229) and left out of the
  proof's commitments, while the tree is opened like the others;
* public.Inputs become a verifier column the prover materialises
  (prover.go:236-273) and the verifier evaluates at zeta
  (verifier/hookFns.go:51-70); their transcript elements follow the setup
  roots (public/public_inputs.go:50-74).

CPU: the oracle's proof passes the checker with the key and the public
inputs and fails with a wrong key or a wrong public value; the first
challenge's message is checked word by word.  GPU: lg_plan_set_setup's root
equals the oracle's key, and the proofs of a batch sharing one key are the
oracle's, bit for bit (eager, recorded and replayed graphs, device inputs).
"""
import sys

import numpy as np
import pytest

P = 2**31 - 2**24 + 1
FSID, HASH, SIZE, PUBI = 0x46534944, 0x48415348, 0x53495A45, 0x50554249


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _chunks3(s):
    b = s.encode()
    return [int.from_bytes(b[i:i + 3], "little") for i in range(0, len(b), 3)]


def test_oracle_setup_and_public_inputs_verify():
    from loom_amd import workloads
    from oracle import prover as OP, verifier as V
    prog, tr, pi = workloads.plonk_setup_instance(5, 3, 11, n_public=3)
    key = OP.setup_roots(prog, tr)
    assert len(key) == 1
    _, enc, _ = OP.prove(prog, tr, 8, pi)
    pf = V.verify(prog, 8, enc, setup_roots=key, public_inputs=pi)
    assert len(pf["commitments"]) == 3  # round 0, round 2, AIR; the setup root is the key's
    assert len(pf["claimed"]) == 4     # the setup tree is opened too
    wrong_pi = {"PI": ("plonk", [(i, (v + (i == 1)) % P) for i, v in pi["PI"][1]])}
    with pytest.raises(Exception):
        V.verify(prog, 8, enc, setup_roots=key, public_inputs=wrong_pi)
    with pytest.raises(Exception):
        V.verify(prog, 8, enc, setup_roots=[[(x + 1) % P for x in key[0]]], public_inputs=pi)
    with pytest.raises(Exception):
        V.verify(prog, 8, enc, setup_roots=[], public_inputs=pi)
    # another witness of the same circuit: the same key
    prog2, tr2, _ = workloads.plonk_setup_instance(5, 3, 12, n_public=3)
    assert OP.setup_roots(prog2, tr2) == key


def test_first_challenge_binds_setup_root_then_public_inputs():
    """prover.go:214-234: HASH id, module sizes, the setup root, then the
    public inputs' transcript elements (public_inputs.go:50-74)."""
    from loom_amd import workloads
    from oracle import prover as OP
    prog, tr, pi = workloads.plonk_setup_instance(4, 5, 6, n_public=2)
    pv = OP.Prover(prog, tr, 4, pi)
    pv.prove()
    key = [int(x) for x in OP.setup_roots(prog, tr)[0]]
    (r0, v0), (r1, v1) = pi["PI"][1]
    want = [FSID, 16] + _chunks3("challenge@loom_0")
    want += [HASH, 9] + _chunks3("poseidon2") + [SIZE, 1, 16] + key
    want += [PUBI, 1, PUBI, 2] + _chunks3("PI") + [PUBI, 5] + _chunks3("plonk") + [2, r0, 0, v0, r1, 0, v1]
    msg = pv.ts.message("challenge@loom_0")
    assert msg[:len(want)] == want
    assert len(msg) == len(want) + 8  # then the round-0 trace root


def test_setup_plan_compiles_without_device():
    from loom_amd import workloads, prover
    prog, tr, pi = workloads.plonk_setup_instance(5, 3, 11, n_public=3)
    it = prover.Prepared(prog, tr, 8, pi)
    assert it.plan.setup == ["QK", "QL", "QM", "QO", "QR"]
    assert it.plan.public == ["PI"]
    assert len(it.columns) == 3 + 1 and len(it.columns[-1]) == 3
    assert prover.check_plan(it.words) > 0


@pytest.mark.gpu
def test_setup_path_equals_oracle():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, workloads
    from oracle import prover as OP, verifier as V
    nq = 16
    insts = [workloads.plonk_setup_instance(9, 21, 100 + k, n_public=5) for k in range(5)]
    key = OP.setup_roots(insts[0][0], insts[0][1])
    items = [prover.Prepared(p, t, nq, pi) for p, t, pi in insts]
    want = [OP.prove(p, t, nq, pi)[1] for p, t, pi in insts]
    ctx = prover.Context(0)
    try:
        h = ctx.plan_handle(items[0].words)
        root = prover.set_setup(h, items[0].setup)
        assert [int(x) for x in root] == [int(x) for x in key[0]], "device setup root differs from the oracle's key"
        ctx.setups[items[0].words.tobytes()] = None  # plan_handle recommits on the next call
        for rep in range(3):  # eager, recorded, replayed
            log("setup path: batch of 5, call %d" % rep)
            got = prover.prove_prepared(ctx, items)
            for k in range(5):
                assert np.array_equal(got[k], want[k]), (rep, k)
        one = prover.prove_prepared(ctx, items[2:3])[0]
        assert np.array_equal(one, want[2])
        d = prover.prove_prepared(ctx, items, dev_inputs=prover.DeviceInputs(items))
        for k in range(5):
            assert np.array_equal(d[k], want[k]), ("device inputs", k)
        V.verify(insts[0][0], nq, got[0], setup_roots=key, public_inputs=insts[0][2])
    finally:
        ctx.close()


@pytest.mark.gpu
def test_setup_plan_refuses_to_prove_without_key():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, workloads, _lib
    p, t, pi = workloads.plonk_setup_instance(6, 1, 2, n_public=1)
    it = prover.Prepared(p, t, 8, pi)
    it.setup = []  # no key given
    ctx = prover.Context(0)
    try:
        with pytest.raises(_lib.LoomGpuError, match="setup"):
            prover.prove_prepared(ctx, [it])
    finally:
        ctx.close()
