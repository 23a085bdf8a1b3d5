"""The proof word stream, decoded with the grammar include/loomgpu.h states
and nothing else.

The decoder below is written from the header comment alone (it does not use
oracle.verifier.decode_proof); it is run on a proof with exposed values (the
uneven-module shape with range lookups: a source and a target logup total
per lookup) and on a single plonk module, and it must consume every
word.  The structural facts it checks follow from the reference's types
(proof/proof.go:31-51, proof/public_inputs.go:23-49, internal/fri/fri.go:
232-248): exposed names in sorted order with one entry at index 0 for a logup
total, one root per commitment, a path with as many siblings as the tree is
deep, queries of the requested count.  A last check holds the header text to
the grammar decoded here.
"""
import os
import re

P = 2**31 - 2**24 + 1
HEADER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "..", "include", "loomgpu.h")


class HeaderDecoder:
    """Reads the word stream exactly as include/loomgpu.h's grammar says."""

    def __init__(self, words):
        self.w = [int(x) for x in words]
        self.i = 0

    def u(self):
        v = self.w[self.i]
        self.i += 1
        return v

    def field(self):
        v = self.u()
        assert v < P, "field word not reduced"
        return v

    def e6(self):
        return tuple(self.field() for _ in range(6))

    def digest(self):
        return tuple(self.field() for _ in range(8))

    def bytes_string(self):
        n = self.u()
        out = b""
        for _ in range((n + 3) // 4):
            out += self.u().to_bytes(4, "little")
        assert all(b == 0 for b in out[n:]), "padding bytes must be zero"
        return out[:n]

    def path(self):
        leaf = self.u()
        sib = [self.digest() for _ in range(self.u())]
        inj = [self.digest() for _ in range(self.u())]
        return dict(leaf=leaf, siblings=sib, inj=inj)

    def row(self):
        base = [self.field() for _ in range(self.u())]
        ext = [self.e6() for _ in range(self.u())]
        return dict(base=base, ext=ext)

    def opening(self):
        assert self.u() == 1, "opening: extension-field flag"
        return dict(row=self.u(), p=self.e6(), q=self.e6(), path=self.path())

    def proof(self):
        pf = {"hash_id": self.bytes_string()}
        ex = []
        for _ in range(self.u()):
            name = self.bytes_string()
            entries = [(self.u(), self.e6()) for _ in range(self.u())]
            ex.append((name, entries))
        pf["exposed"] = ex
        pf["commitments"] = [self.digest() for _ in range(self.u())]
        claimed = []
        for _ in range(self.u()):
            groups = []
            for _ in range(self.u()):
                rails = []
                for _ in range(2):
                    rails.append([[self.e6() for _ in range(self.u())] for _ in range(self.u())])
                groups.append(rails)
            claimed.append(groups)
        pf["claimed"] = claimed
        fri = {"levels_root": self.digest()}
        fri["level_queries"] = [[self.opening() for _ in range(self.u())] for _ in range(self.u())]
        fri["roots"] = [self.digest() for _ in range(self.u())]
        assert self.u() == 1, "final polynomial: extension-field flag"
        fri["final"] = [self.e6() for _ in range(self.u())]
        fri["queries"] = [[self.opening() for _ in range(self.u())] for _ in range(self.u())]
        fri["pow"] = [(self.bytes_string(), self.u(), self.field()) for _ in range(self.u())]
        pf["fri"] = fri
        samp = []
        for _ in range(self.u()):
            trees = []
            for _ in range(self.u()):
                lo, hi = self.row(), self.row()
                path = self.path()
                inj = [(self.row(), self.row()) for _ in range(self.u())]
                trees.append(dict(lo=lo, hi=hi, path=path, inj=inj))
            samp.append(trees)
        pf["samplings"] = samp
        assert self.i == len(self.w), "trailing words: %d of %d read" % (self.i, len(self.w))
        return pf


def _oracle_proof(prog, tr, nq):
    from oracle import prover as OP
    _, enc, _ = OP.prove(prog, tr, nq)
    return enc


def _check_common(pf, prog, nq):
    from loom_amd import protocol as PR
    lay = PR.build_layout(prog)
    assert pf["hash_id"] == b"poseidon2"
    names = [n for n, _ in pf["exposed"]]
    assert names == sorted(names, key=lambda s: s)
    assert len(pf["commitments"]) == len(lay.trees)
    assert len(pf["claimed"]) == len(lay.trees)
    for t, groups in zip(lay.trees, pf["claimed"]):
        assert len(groups) == len(t.groups)
        for g, (base, ext) in zip(t.groups, groups):
            assert [len(v) for v in base] == [len(s) for s in g.base_shifts]
            assert [len(v) for v in ext] == [len(s) for s in g.ext_shifts]
    fri = pf["fri"]
    assert fri["pow"] == []
    assert len(fri["queries"]) == nq and len(pf["samplings"]) == nq
    sizes = sorted({prog.modules[m].N for m in prog.modules}, reverse=True)
    assert len(fri["level_queries"]) == len(sizes) - 1
    for lq in fri["level_queries"]:
        assert len(lq) == nq
        for op in lq:
            assert op["path"] == dict(leaf=0, siblings=[], inj=[])
    # multi-degree FRI over N = 2^24 with D = 2^22 (fri_params.go:29-42): the
    # largest level enters at round 22 - log2 N_max and every round after it
    # folds one layer; the last 21 - r0 rounds have a tree each
    r0 = 22 - (sizes[0].bit_length() - 1)
    assert len(fri["roots"]) == 21 - r0
    assert len(fri["final"]) == 4
    for q in fri["queries"]:
        assert len(q) == 22 - r0
        for j, op in enumerate(q):
            depth = 24 - (r0 + j) - 1
            assert len(op["path"]["siblings"]) == depth
            assert op["path"]["inj"] == []
    for q in pf["samplings"]:
        assert len(q) == len(lay.trees)
        for t, rec in zip(lay.trees, q):
            big = max(g.N for g in t.groups)
            assert len(rec["path"]["siblings"]) == (4 * big // 2).bit_length() - 1
            assert len(rec["inj"]) == len(t.groups) - 1 == len(rec["path"]["inj"])


def test_decode_proof_with_exposed_values():
    from loom_amd import workloads
    prog, tr = workloads.uneven_range_instance([4, 5], 7)
    nq = 6
    pf = HeaderDecoder(_oracle_proof(prog, tr, nq)).proof()
    _check_common(pf, prog, nq)
    names = [n for n, _ in pf["exposed"]]
    # each arguments.Range lookup has a source logup column in its module and
    # a target column in the range module: two totals per plonk module
    assert len(names) == 4, names
    for _, entries in pf["exposed"]:
        assert len(entries) == 1 and entries[0][0] == 0
    # the bus: the exposed logup totals of a range lookup sum to zero
    tot = [0] * 6
    for _, entries in pf["exposed"]:
        tot = [(a + b) % P for a, b in zip(tot, entries[0][1])]
    assert tot == [0] * 6


def test_decode_single_module_proof():
    from loom_amd import workloads
    prog, tr = workloads.plonk_instance(5, 3)
    nq = 5
    pf = HeaderDecoder(_oracle_proof(prog, tr, nq)).proof()
    _check_common(pf, prog, nq)
    assert pf["exposed"] == []


def test_proofs_of_work_by_the_header_grammar():
    """With grinding the list holds one (name, bits, salt) per fold challenge
    in ascending byte order of the names, and the decoder consumes every word."""
    from loom_amd import workloads
    from oracle import prover as OP
    prog, tr = workloads.plonk_instance(5, 4)
    pf, enc, _ = OP.prove(prog, tr, 4, grinding=3)
    d = HeaderDecoder(enc)
    got = d.proof()["fri"]["pow"]
    assert d.i == len(d.w)
    names = [n.decode() for n, _, _ in got]
    assert names == sorted(("fri_fold_%d" % j for j in range(22)), key=lambda x: x.encode())
    R_inv = pow(2**32, P - 2, P)
    for n, bits, salt in got:
        assert bits == 3 and salt * R_inv % P == pf["fri"]["pow"][n.decode()][1]


def test_decoders_agree():
    """The header decoder and the oracle verifier's decoder read the same
    values (two readers of one grammar)."""
    from loom_amd import workloads
    from oracle import verifier as V
    prog, tr = workloads.uneven_range_instance([4], 2)
    enc = _oracle_proof(prog, tr, 4)
    a = HeaderDecoder(enc).proof()
    b = V.decode_proof(enc)
    R_inv = pow(2**32, P - 2, P)
    canon = lambda v: tuple(int(x) * R_inv % P for x in v)
    assert [canon(c) for c in a["commitments"]] == [tuple(int(x) for x in d) for d in b["commitments"]]
    assert sorted(n.decode() for n, _ in a["exposed"]) == sorted(b["exposed"])
    for n, ents in a["exposed"]:
        assert canon(ents[0][1]) == tuple(int(x) for x in b["exposed"][n.decode()])


def test_header_states_this_grammar():
    text = open(HEADER).read()
    block = text[text.index("Proof word stream"):text.index("#ifndef LOOMGPU_H")]
    flat = re.sub(r"\s+", " ", block.replace("*", " "))
    for phrase in ["#exposed, then per exposed value in ascending byte order of its name",
                   "#entries, then per entry: Idx, ExtValue() as E6",
                   "#commitments, roots (8 words each; trace rounds then AIR)",
                   "#levels - 1, then per level l >= 1: #queries { opening }",
                   "#proofs of work (0 without grinding), then per fold challenge in ascending byte order of its "
                   "name (\"fri_fold_0\", \"fri_fold_1\", \"fri_fold_10\", ...): name (byte length, then bytes "
                   "packed 4 per word), NbBits, Salt (Montgomery)",
                   "opening := 1 (extension-field layer), row index, P (E6), Q (E6), path",
                   "path := leaf index, #siblings, siblings (8 words each, leaf to root), "
                   "#injection leaves, leaves (8 words each)",
                   "row := #base, base values, #ext, ext values (6 words each)"]:
        assert phrase in flat, phrase
