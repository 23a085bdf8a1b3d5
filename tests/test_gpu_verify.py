"""Batched proof checking on the device (lg_verify), against the oracle's
checker.

verifier.Verify (verifier/verifier.go:479-533) for many proofs of one plan in
one call: valid batches (a plonk-like module; uneven modules with range
lookups and level injections; setup columns with public inputs) pass, and the
oracle's checker (oracle/verifier.py) accepts the same proofs.  Planted
mistakes are rejected by the check the reference runs first for them
(fri/verify.go:127-155 orders the FRI proof, the point samplings and the
bridge):
  a claimed value of a quotient chunk -> AIR identity at zeta;
  a logup total -> logup bus;
  a sibling of an opening path -> opening;
  an injection leaf of an opening path -> opening;
  an opened row value -> opening (its leaf no longer reaches the root);
  a FRI layer value -> FRI;
  a truncated proof -> format;
and every mutated proof is rejected by the oracle's checker too, while the
untouched proofs of the same batch stay valid.  prover.SkipFRI proofs carry
the full proof's words up to the claimed values and an empty opening; the
SkipFRI verifier accepts them.
"""
import sys

import numpy as np
import pytest

P = 2**31 - 2**24 + 1
R = (1 << 32) % P
OK, FORMAT, BUS, AIR, OPENING, DEEP, FRI = range(7)


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


class Offsets:
    """Walks the proof words by include/loomgpu.h's grammar and records where
    the planted mistakes go."""

    def __init__(self, words):
        self.w = [int(x) for x in words]
        self.i = 0
        self.exposed = []
        self.commitments = []
        self.claimed = []  # [tree][group][poly][shift] -> offset of the E6
        self.layers = []   # [query][layer] -> (offset of P, path)
        self.samp = []     # [query][tree] -> dict(lo_base, path)
        self._walk()

    def u(self):
        v = self.w[self.i]
        self.i += 1
        return v

    def skip(self, n):
        at = self.i
        self.i += n
        return at

    def string(self):
        n = self.u()
        self.skip((n + 3) // 4)

    def path(self):
        idx = self.u()
        nsib = self.u()
        sib = self.skip(8 * nsib)
        ninj = self.u()
        inj = self.skip(8 * ninj)
        return dict(idx=idx, nsib=nsib, sib=sib, ninj=ninj, inj=inj)

    def row(self):
        nb = self.u()
        base = self.skip(nb)
        ne = self.u()
        ext = self.skip(6 * ne)
        return dict(nb=nb, base=base, ne=ne, ext=ext)

    def opening(self):
        assert self.u() == 1
        self.u()
        p = self.skip(6)
        self.skip(6)
        return p, self.path()

    def _walk(self):
        self.string()
        for _ in range(self.u()):
            self.string()
            assert self.u() == 1
            self.u()
            self.exposed.append(self.skip(6))
        for _ in range(self.u()):
            self.commitments.append(self.skip(8))
        for _ in range(self.u()):
            groups = []
            for _ in range(self.u()):
                polys = []
                for _ in range(2):
                    for _ in range(self.u()):
                        polys.append([self.skip(6) for _ in range(self.u())])
                groups.append(polys)
            self.claimed.append(groups)
        self.head_end = self.i  # the words up to the claimed values
        self.skip(8)
        for _ in range(self.u()):
            for _ in range(self.u()):
                self.opening()
        self.skip(8 * self.u())
        assert self.u() == 1
        self.skip(6 * self.u())
        for _ in range(self.u()):
            self.layers.append([self.opening() for _ in range(self.u())])
        assert self.u() == 0
        for _ in range(self.u()):
            trees = []
            for _ in range(self.u()):
                lo = self.row()
                self.row()
                pth = self.path()
                for _ in range(self.u()):
                    self.row()
                    self.row()
                trees.append(dict(lo=lo, path=pth))
            self.samp.append(trees)
        assert self.i == len(self.w)


def _bump(words, at):
    w = np.array(words, dtype=np.uint32).copy()
    w[at] = (int(w[at]) + R) % P  # +1 in Montgomery form: still a reduced field word
    return w


@pytest.fixture(scope="module")
def ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover
    c = prover.Context(0)
    yield c
    c.close()


def _mutants(pf, air_tree, lookups):
    o = Offsets(pf)
    out = {}
    # the last polynomial of the AIR tree's first group is a quotient chunk
    out["quotient chunk claimed value"] = (_bump(pf, o.claimed[air_tree][0][-1][0]), AIR)
    if lookups:
        out["logup total"] = (_bump(pf, o.exposed[0]), BUS)
    s = o.samp[5][0]
    out["opening sibling"] = (_bump(pf, s["path"]["sib"] + 8 * 3), OPENING)
    out["opened row value"] = (_bump(pf, s["lo"]["base"] if s["lo"]["nb"] else s["lo"]["ext"]), OPENING)
    inj = [(k, t) for k in range(len(o.samp)) for t in range(len(o.samp[k])) if o.samp[k][t]["path"]["ninj"]]
    if inj:
        k, t = inj[0]
        out["injection leaf"] = (_bump(pf, o.samp[k][t]["path"]["inj"]), OPENING)
    p_off, _ = o.layers[3][2]
    out["FRI layer value"] = (_bump(pf, p_off), FRI)
    out["truncated"] = (np.array(pf[:-5], dtype=np.uint32), FORMAT)
    return out


def _check_batch(ctx, items, proofs, air_tree, lookups, oracle_args, setup_roots=()):
    from loom_amd import prover
    from oracle import verifier as V
    log("verify: %d valid proofs" % len(proofs))
    res = prover.verify_prepared(ctx, items, proofs, setup_roots)
    assert res == [(OK, "")] * len(proofs), res
    for it, pf, a in zip(items, proofs, oracle_args):
        V.verify(*a, pf, **({"setup_roots": setup_roots} if setup_roots else {}),
                 **({"public_inputs": it.public_inputs} if it.public_inputs else {}))
    muts = _mutants(proofs[0], air_tree, lookups)
    names = list(muts)
    # every mutant beside an untouched proof of the batch
    batch = [muts[n][0] for n in names] + [proofs[1]]
    log("verify: %d planted mistakes" % len(names))
    res = prover.verify_prepared(ctx, [items[0]] * len(names) + [items[1]], batch, setup_roots)
    for n, (st, msg) in zip(names, res):
        log("  %-30s -> %d %s" % (n, st, msg))
        assert st == muts[n][1], (n, st, msg)
        assert msg
    assert res[-1] == (OK, "")
    for n in names:
        with pytest.raises(Exception):
            V.verify(*oracle_args[0], muts[n][0], **({"setup_roots": setup_roots} if setup_roots else {}),
                     **({"public_inputs": items[0].public_inputs} if items[0].public_inputs else {}))


@pytest.mark.gpu
def test_verify_plonk_batch(ctx):
    from loom_amd import prover, workloads
    nq = 24
    insts = [workloads.plonk_instance(9, 700 + k) for k in range(6)]
    items = [prover.Prepared(p, t, nq) for p, t in insts]
    proofs = prover.prove_prepared(ctx, items)
    air = [i for i, t in enumerate(items[0].plan.layout.trees) if t.kind == "air"][0]
    _check_batch(ctx, items, proofs, air, False, [(p, nq) for p, _ in insts])
    # verifier.SkipFRI: a broken FRI layer passes, a broken AIR claim does not
    muts = _mutants(proofs[0], air, False)
    res = prover.verify_prepared(ctx, items[:1] * 2, [muts["FRI layer value"][0],
                                                      muts["quotient chunk claimed value"][0]], skip_fri=True)
    assert res[0] == (OK, "") and res[1][0] == AIR


@pytest.mark.gpu
def test_verify_uneven_modules_with_lookups(ctx):
    """Several DEEP classes and FRI levels, groups of different sizes as level
    injections, logup buses."""
    from loom_amd import prover, workloads
    nq = 16
    insts = [workloads.uneven_range_instance([7, 9, 8, 9], 40 + k) for k in range(3)]
    items = [prover.Prepared(p, t, nq) for p, t in insts]
    proofs = prover.prove_prepared(ctx, items)
    air = [i for i, t in enumerate(items[0].plan.layout.trees) if t.kind == "air"][0]
    _check_batch(ctx, items, proofs, air, True, [(p, nq) for p, _ in insts])


@pytest.mark.gpu
def test_verify_setup_and_public_inputs(ctx):
    from loom_amd import prover, workloads
    from oracle import prover as OP
    nq = 16
    insts = [workloads.plonk_setup_instance(8, 5, 60 + k, n_public=4) for k in range(3)]
    items = [prover.Prepared(p, t, nq, pi) for p, t, pi in insts]
    proofs = prover.prove_prepared(ctx, items)
    key = [[int(x) for x in r] for r in OP.setup_roots(insts[0][0], insts[0][1])]
    air = [i for i, t in enumerate(items[0].plan.layout.trees) if t.kind == "air"][0]
    _check_batch(ctx, items, proofs, air, False, [(p, nq) for p, _, _ in insts], setup_roots=key)
    # a wrong key or a wrong public value fails
    bad_key = [[(key[0][0] + 1) % P] + key[0][1:]]
    assert prover.verify_prepared(ctx, items[:1], proofs[:1], bad_key)[0][0] != OK
    it = prover.Prepared(*insts[0][:2], nq, insts[0][2])
    it.columns = list(it.columns)
    it.columns[-1] = it.columns[-1].copy()
    it.columns[-1][0] = (int(it.columns[-1][0]) + R) % P
    assert prover.verify_prepared(ctx, [it], proofs[:1], key)[0][0] != OK


@pytest.mark.gpu
def test_verify_full_size_batch(ctx):
    """The 40 x 2^14 batch of the bench: every proof valid."""
    from loom_amd import prover, workloads
    items = [prover.Prepared(p, t, 200) for p, t in (workloads.plonk_instance(14, s) for s in range(40))]
    proofs = prover.prove_prepared(ctx, items)
    log("verify: 40 x 2^14")
    res = prover.verify_prepared(ctx, items, proofs)
    assert res == [(OK, "")] * 40


@pytest.mark.gpu
def test_prove_skip_fri(ctx):
    """prover.SkipFRI (prover/prover.go:626-653, 691-700): the same words as
    the full proof up to the claimed values (the transcript up to zeta is the
    same), then the empty opening; verifier.SkipFRI accepts it, the full
    verifier does not, and proofs are full again once the option is cleared."""
    from loom_amd import prover, workloads
    nq = 16
    insts = [workloads.plonk_instance(8, 900 + k) for k in range(3)]
    items = [prover.Prepared(p, t, nq) for p, t in insts]
    full = [f.copy() for f in prover.prove_prepared(ctx, items)]
    prover.set_options(ctx, items, skip_fri=True)
    try:
        short = [f.copy() for f in prover.prove_prepared(ctx, items)]
    finally:
        prover.set_options(ctx, items, skip_fri=False)
    empty = [0] * 8 + [0, 0, 1, 0, 0, 0, 0]
    for f, sh in zip(full, short):
        h = Offsets(f).head_end
        assert np.array_equal(sh[:h], f[:h])
        assert [int(x) for x in sh[h:]] == empty
    assert prover.verify_prepared(ctx, items, short, skip_fri=True) == [(OK, "")] * 3
    assert all(st == FORMAT for st, _ in prover.verify_prepared(ctx, items, short))
    again = prover.prove_prepared(ctx, items)
    for f, g in zip(full, again):
        assert np.array_equal(f, g)


@pytest.mark.gpu
def test_phase_timings(ctx):
    """prover.WithPhaseCallback (prover/prover.go:99-108, 676-716): an eager
    prove of modules of different sizes (no recorded graph) reports three
    positive phase durations, execute-steps, compute-air-quotients and
    pcs-open, whose sum stays within the call's wall time; under SkipFRI the
    third slot is pcs-claimed-values-only and the proof is still the same."""
    import time
    from loom_amd import prover, workloads
    log("phase timings: uneven (9, 11, 8), 16 queries")
    prog, tr = workloads.uneven_instance([9, 11, 8], 77)
    items = [prover.Prepared(prog, tr, 16)]
    for skip in (False, True):
        prover.set_options(ctx, items, skip_fri=skip)
        try:
            ph = []
            t0 = time.perf_counter()
            w = prover.prove_prepared(ctx, items, phases=ph)[0]
            wall = time.perf_counter() - t0
        finally:
            prover.set_options(ctx, items, skip_fri=False)
        assert w is not None and len(ph) == 3
        assert all(x > 0 for x in ph), ph
        assert sum(ph) <= wall, (ph, wall)
