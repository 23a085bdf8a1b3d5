"""Pins the oracle (CPU): golden vectors the reference tree holds, closed
forms, brute force on tiny inputs, and the verifier's own identities.

What pins what:
* Poseidon2 constants: the reference's in-tree width-16 table
  (tests/golden/poseidon2_w16_zkc.json, from constants.zkc) -- the round-key
  derivation and the diagonal formula must reproduce it exactly; the width-16
  permutation must equal an independent straight-line restatement of the
  zkc program (poseidon2.zkc:33-176).
* Keccak: hashlib's SHA3-256 (same permutation, different padding).
* E6: brute-force polynomial product reduced mod (u^2 - 3, v^3 - (u + 1)).
* NTT / RS encode: naive O(n^2) DFT and Horner evaluation at the
  bit-reversed points (rs_eval_test.go pattern).
* Merkle with injections: node-by-node recursion written independently
  (merkle/injection_test.go ``referenceRoot`` pattern) and path verification
  (tree.go:392-436).
* Sponge: partial last block overwrites only its lanes (hash_test.go:394-414).
* Whole proof: the verifier's AIR identity at zeta (verifier.go:406-444),
  FRI fold consistency of every query (fri.go:1430-1534), Merkle paths of
  every opening against the committed roots, and the DEEP codeword against a
  slow per-row restatement (deep_test.go pattern).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from oracle import field as F
from oracle import ntt
from oracle import poseidon2 as p2
from oracle.merkle import Tree
from oracle.transcript import Transcript, string_to_elements

P = F.P
HERE = os.path.dirname(os.path.abspath(__file__))


def zkc():
    with open(os.path.join(HERE, "golden", "poseidon2_w16_zkc.json")) as f:
        return json.load(f)


# ------------------------------------------------------------------ keccak / constants
def test_keccak_matches_sha3_with_sha3_padding():
    for msg in [b"", b"abc", b"x" * 135, b"y" * 136, b"z" * 300]:
        assert p2.keccak256(msg, pad=0x06) == hashlib.sha3_256(msg).digest()


def test_keccak_legacy_empty():
    assert p2.keccak256(b"").hex() == "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"


def test_round_keys_width16_match_reference_table():
    g = zkc()
    rows = p2.round_keys(16)
    flat = []
    for r in rows:
        flat += r + [0] * (16 - len(r))
    assert flat == g["round_keys_27x16"]


def test_internal_diag_width16_matches_reference_table():
    assert p2.p2_diag(16) == zkc()["internal_diag"]


def test_product_constants_header_matches_oracle():
    import re
    hdr = open(os.path.join(HERE, "..", "loom_amd", "csrc", "poseidon2_constants.h")).read()

    def arr(name):
        blk = hdr.split(name + "[")[1].split("};")[0]
        return [int(x, 16) for x in re.findall(r"0x([0-9a-f]{8})u", blk)]

    for w in (16, 24):
        rows = p2.round_keys(w)
        full = [v for r in rows if len(r) == w for v in r]
        part = [r[0] for r in rows if len(r) == 1]
        assert arr("P2_W%d_FULL_RK_CANON" % w) == full
        assert arr("P2_W%d_PARTIAL_RK_CANON" % w) == part
        assert arr("P2_W%d_DIAG_CANON" % w) == p2.p2_diag(w)
        assert arr("P2_W%d_FULL_RK" % w) == [F.to_mont(v) for v in full]


def _zkc_permutation16(state):
    """Straight-line restatement of poseidon2.zkc:33-176 on Python ints."""
    g = zkc()
    rk, dg = g["round_keys_27x16"], g["internal_diag"]
    s = [x % P for x in state]

    def m4():
        for c in range(4):
            x0, x1, x2, x3 = s[4 * c:4 * c + 4]
            t01, t23 = (x0 + x1) % P, (x2 + x3) % P
            t0123 = (t01 + t23) % P
            t01123, t01233 = (t0123 + x1) % P, (t0123 + x3) % P
            s[4 * c + 3] = (2 * x0 + t01233) % P
            s[4 * c + 1] = (2 * x2 + t01123) % P
            s[4 * c] = (t01 + t01123) % P
            s[4 * c + 2] = (t23 + t01233) % P

    def ext():
        m4()
        sums = [sum(s[j::4]) % P for j in range(4)]
        for i in range(16):
            s[i] = (s[i] + sums[i % 4]) % P

    def full(r):
        for i in range(16):
            s[i] = pow((s[i] + rk[16 * r + i]) % P, 3, P)
        ext()

    def partial(r):
        s[0] = pow((s[0] + rk[16 * r]) % P, 3, P)
        tot = sum(s) % P
        for i in range(16):
            s[i] = (tot + dg[i] * s[i]) % P

    ext()
    for r in range(3):
        full(r)
    for r in range(3, 24):
        partial(r)
    for r in range(24, 27):
        full(r)
    return s


def test_permutation16_matches_zkc_restatement():
    perm = p2.Permutation(16)
    rng = np.random.default_rng(5)
    for _ in range(4):
        st = [int(x) for x in rng.integers(0, P, size=16)]
        got = [int(x) for x in perm(np.array(st, dtype=np.uint64))]
        assert got == _zkc_permutation16(st)


# ------------------------------------------------------------------ field
def test_generator_and_root_properties():
    # MULT_GEN generates F_p^* (p - 1 = 2^24 * 127)
    assert all(pow(F.MULT_GEN, (P - 1) // q, P) != 1 for q in (2, 127))
    w = F.OMEGA_MAX
    assert pow(w, 1 << 24, P) == 1 and pow(w, 1 << 23, P) == P - 1
    assert F.omega(4) ** 2 % P == P - 1


def test_montgomery_roundtrip():
    rng = np.random.default_rng(1)
    a = rng.integers(0, P, size=100, dtype=np.uint64)
    assert np.array_equal(F.from_mont(F.to_mont(a)), a)
    x, y = 123456789, 987654321
    assert F.to_mont(x) * F.to_mont(y) * pow(2**32, -1, P) % P == F.to_mont(x * y % P)


def _e6_brute(x, y):
    # polynomials in u, v: element = sum c[2i + j] u^j v^i; reduce v^3 = u + 1, u^2 = 3
    prod = {}
    for i in range(3):
        for j in range(2):
            for k in range(3):
                for l in range(2):
                    key = (i + k, j + l)
                    prod[key] = (prod.get(key, 0) + x[2 * i + j] * y[2 * k + l]) % P
    changed = True
    while changed:
        changed = False
        for (vi, uj), c in list(prod.items()):
            if c == 0:
                continue
            if uj >= 2:
                prod[(vi, uj)] = 0
                prod[(vi, uj - 2)] = (prod.get((vi, uj - 2), 0) + 3 * c) % P
                changed = True
            elif vi >= 3:
                prod[(vi, uj)] = 0
                prod[(vi - 3, uj + 1)] = (prod.get((vi - 3, uj + 1), 0) + c) % P
                prod[(vi - 3, uj)] = (prod.get((vi - 3, uj), 0) + c) % P
                changed = True
    return [prod.get((i, j), 0) % P for i in range(3) for j in range(2)]


def test_e6_mul_inverse():
    rng = np.random.default_rng(2)
    for _ in range(20):
        x = [int(v) for v in rng.integers(0, P, size=6)]
        y = [int(v) for v in rng.integers(0, P, size=6)]
        got = [int(v) for v in F.e6_mul(np.array(x, dtype=np.uint64), np.array(y, dtype=np.uint64))]
        assert got == _e6_brute(x, y)
        xi = F.e6_inv(np.array(x, dtype=np.uint64))
        assert [int(v) for v in F.e6_mul(np.array(x, dtype=np.uint64), xi)] == [1, 0, 0, 0, 0, 0]


# ------------------------------------------------------------------ NTT
@pytest.mark.parametrize("logn", [0, 1, 3, 5])
def test_ntt_matches_naive_dft(logn):
    n = 1 << logn
    rng = np.random.default_rng(logn)
    c = rng.integers(0, P, size=n, dtype=np.uint64)
    w = F.omega(n)
    want = [sum(int(c[k]) * pow(w, j * k, P) for k in range(n)) % P for j in range(n)]
    assert [int(v) for v in ntt.evaluate(c)] == want
    assert np.array_equal(ntt.interpolate(ntt.evaluate(c)), c)


@pytest.mark.parametrize("logn", [1, 3, 6])
def test_encode_is_horner_at_bitreversed_points(logn):
    n = 1 << logn
    rng = np.random.default_rng(logn + 10)
    vals = rng.integers(0, P, size=n, dtype=np.uint64)
    coeffs = ntt.interpolate(vals)
    enc = ntt.encode(vals, 4)
    big = 4 * n
    w = F.omega(big)
    br = F.bitrev_indices(big)
    for r in range(big):
        assert int(enc[r]) == ntt.horner(coeffs, pow(w, int(br[r]), P))
    # the first n rows are the original values on H, in bit-reversed order
    assert np.array_equal(enc[:n], vals[F.bitrev_indices(n)])


# ------------------------------------------------------------------ sponge / merkle / transcript
def test_sponge_partial_block_overwrites_only_its_lanes():
    msg = list(range(1, 18))  # 17 elements: one full block, then 1 element
    perm = p2.perm24()
    st = np.zeros(24, dtype=np.uint64)
    st[:16] = msg[:16]
    st = perm(st)
    st[0] = msg[16]          # only lane 0 is overwritten
    st = perm(st)
    assert p2.sponge1(msg) == [int(v) for v in st[:8]]
    assert p2.sponge1([]) == [0] * 8


def _ref_root(leaves, injections):
    level = [tuple(leaves[:, i]) for i in range(leaves.shape[1])]
    inj = list(injections)

    def h(a, b):
        return tuple(int(x) for x in p2.node_compress(np.array(a, dtype=np.uint64)[:, None],
                                                      np.array(b, dtype=np.uint64)[:, None])[:, 0])

    def fold(level):
        while inj and inj[0].shape[1] == len(level):
            x = inj.pop(0)
            level = [h(level[j], tuple(x[:, j])) for j in range(len(level))]
        return level

    level = fold(level)
    while len(level) > 1:
        level = [h(level[2 * i], level[2 * i + 1]) for i in range(len(level) // 2)]
        level = fold(level)
    return [int(v) for v in level[0]]


@pytest.mark.parametrize("logn,widths", [(0, []), (3, []), (4, [16, 4, 4, 1]), (5, [8])])
def test_merkle_matches_node_by_node(logn, widths):
    n = 1 << logn
    rng = np.random.default_rng(logn)
    leaves = rng.integers(0, P, size=(8, n), dtype=np.uint64)
    inj = [rng.integers(0, P, size=(8, w), dtype=np.uint64) for w in widths]
    t = Tree(leaves, inj)
    assert t.root() == _ref_root(leaves, inj)
    # every opening verifies (tree.go:392-436)
    for idx in range(n):
        li, sibs, injl = t.open(idx)
        hcur = [int(v) for v in leaves[:, idx]]
        nxt = 0

        def hh(a, b):
            return [int(x) for x in p2.node_compress(np.array(a, dtype=np.uint64)[:, None],
                                                     np.array(b, dtype=np.uint64)[:, None])[:, 0]]

        def foldp(width):
            nonlocal hcur, nxt
            while nxt < len(widths) and widths[nxt] == width:
                hcur = hh(hcur, injl[nxt])
                nxt += 1

        foldp(n)
        pos = idx
        for k, s in enumerate(sibs):
            hcur = hh(hcur, s) if pos % 2 == 0 else hh(s, hcur)
            pos >>= 1
            foldp(n >> (k + 1))
        assert hcur == t.root()


def test_transcript_message_sequence():
    ts = Transcript(["a", "bb"])
    ts.bind("a", [5, 6])
    assert ts.message("a") == [0x46534944, 1, ord("a"), 5, 6]
    va = ts.compute("a")
    ts.bind("bb", [7])
    assert ts.message("bb") == [0x46534944, 2, ord("b") | ord("b") << 8] + va + [7]
    assert string_to_elements(0x48415348, "poseidon2") == [
        0x48415348, 9, ord("p") | ord("o") << 8 | ord("s") << 16,
        ord("e") | ord("i") << 8 | ord("d") << 16, ord("o") | ord("n") << 8 | ord("2") << 16]


# ------------------------------------------------------------------ whole proof
@pytest.fixture(scope="module")
def small_proof():
    from loom_amd import workloads
    from oracle import prover as OP
    prog, tr = workloads.plonk_instance(5, 21)
    pv = OP.Prover(prog, tr, 6)
    pf = pv.prove()
    return prog, pv, pf


def test_air_identity_at_zeta(small_proof):
    from loom_amd import expr as E
    from oracle import prover as OP
    prog, pv, pf = small_proof
    zeta = pv.zeta
    m = prog.modules["plonk_0"]
    n = m.N
    vals = {}
    for ti, t in enumerate(pv.layout.trees):
        for gi, g in enumerate(t.groups):
            gb, ge = pf["claimed"][ti][gi]
            for nm, ss, vs in list(zip(g.base, g.base_shifts, gb)) + list(zip(g.ext, g.ext_shifts, ge)):
                for s, v in zip(ss, vs):
                    vals[(nm, s)] = np.asarray(v, dtype=np.uint64)

    def leaf(x):
        if x.type == E.CHALLENGE:
            return pv.ext[x.name][0][None, :]
        if x.type == E.VERIFIER:
            zn = F.e6_pow(zeta, n)
            num = F.e6_sub(zn, F.e6_one())
            den = F.e6_scale(F.e6_sub(zeta, F.e6_one()), n)
            return F.e6_mul(num, F.e6_inv(den))[None, :]
        return vals[(x.name, x.shift % n)][None, :]

    ev = OP.eval_expr(m.relation, leaf, 1)[0]
    acc = F.e6_zero()
    for i in range(4):
        q = np.asarray(pf["claimed"][2][0][1][i][0], dtype=np.uint64)
        acc = F.e6_add(acc, F.e6_mul(F.e6_pow(zeta, i * n), q))
    rhs = F.e6_mul(F.e6_sub(F.e6_pow(zeta, n), F.e6_one()), acc)
    assert np.array_equal(ev, rhs)
    # a planted mistake in the oracle (wrong claimed value) breaks the identity
    bad = dict(vals)
    k = next(iter(bad))
    bad[k] = F.e6_add(bad[k], F.e6_one())
    vals, saved = bad, vals
    assert not np.array_equal(OP.eval_expr(m.relation, leaf, 1)[0], rhs)


def test_fri_queries_fold_consistently(small_proof):
    from oracle import prover as OP
    prog, pv, pf = small_proof
    fr = pf["fri"]
    n0 = 4 * prog.modules["plonk_0"].N
    first = OP.FRI_ROUNDS - (n0.bit_length() - 1 - 2)
    # replay the fold challenges from the transcript log
    alphas = {int(name.split("_")[-1]): np.array(v[:6], dtype=np.uint64)
              for name, v in pv.log if name.startswith("fri_fold_")}
    for q in fr["queries"]:
        for i in range(len(q) - 1):
            lay, nxt = q[i], q[i + 1]
            j = first + i
            n = OP.FRI_N >> j
            half = n // 2
            parent = lay["row"] >> 1
            ginv = F.inv(F.omega(n))
            xinv = pow(ginv, int(F.bitrev_indices(half)[parent]), P)
            p, qq = np.asarray(lay["p"], dtype=np.uint64), np.asarray(lay["q"], dtype=np.uint64)
            s = F.e6_scale(F.e6_add(p, qq), OP.INV2)
            d = F.e6_mul(F.e6_scale(F.e6_sub(p, qq), xinv * OP.INV2 % P), alphas[j])
            want = F.e6_add(s, d)
            got = nxt["p"] if parent % 2 == 0 else nxt["q"]
            assert np.array_equal(want, np.asarray(got, dtype=np.uint64))


def test_openings_verify_against_roots(small_proof):
    prog, pv, pf = small_proof
    for q, row in enumerate(pf["samplings"]):
        for t, wp in enumerate(row):
            lo, hi = wp["top"]
            nb, ne = len(lo[0]), len(lo[1])
            msg = [p2.LEAF_TAG, 2 * nb, 2 * ne] + lo[0] + hi[0]
            for v in lo[1]:
                msg += v
            for v in hi[1]:
                msg += v
            h = p2.sponge1(msg)
            idx, sibs, injl = wp["path"]
            for s in sibs:
                a, b = (h, s) if idx % 2 == 0 else (s, h)
                h = [int(x) for x in p2.node_compress(np.array(a, dtype=np.uint64)[:, None],
                                                      np.array(b, dtype=np.uint64)[:, None])[:, 0]]
                idx >>= 1
            assert h == pf["commitments"][t]


def test_deep_codeword_matches_slow_restatement(small_proof):
    """DEEP values on H recomputed row by row with scalar arithmetic, as the
    slow reference of deep_test.go states them."""
    prog, pv, pf = small_proof
    n = prog.modules["plonk_0"].N
    alpha = next(np.array(v[:6], dtype=np.uint64) for nm, v in pv.log if nm == "alpha_DEEP")
    w = F.omega(n)
    for x in [0, 1, n // 2, n - 1]:
        acc = F.e6_zero()
        a = F.e6_one()
        xv = pow(w, x, P)
        for ti, t in enumerate(pv.layout.trees):
            g = t.groups[0]
            bases, exts = pv.tree_polys(ti, g)
            gb, ge = pf["claimed"][ti][0]
            for col, ss, vs in list(zip(bases, g.base_shifts, gb)) + list(zip(exts, g.ext_shifts, ge)):
                px = F.e6_from_base(np.array(col[x])) if col.ndim == 1 else np.asarray(col[x], dtype=np.uint64)
                for s, v in zip(ss, vs):
                    z = F.e6_scale(pv.zeta, pow(w, s, P))
                    num = F.e6_mul(a, F.e6_sub(np.asarray(v, dtype=np.uint64), px))
                    den = F.e6_sub(z, F.e6_from_base(np.array(xv)))
                    acc = F.e6_add(acc, F.e6_mul(num, F.e6_inv(den)))
                a = F.e6_mul(a, alpha)
        assert np.array_equal(acc, pv.deep_on_h[0][x])


def test_claimed_values_equal_horner(small_proof):
    """Claimed values (barycentric, values.go) equal the interpolant evaluated
    at zeta * w^s by Horner in E6."""
    prog, pv, pf = small_proof
    n = prog.modules["plonk_0"].N
    w = F.omega(n)
    t = pv.layout.trees[1]
    g = t.groups[0]
    bases, exts = pv.tree_polys(1, g)
    gb, ge = pf["claimed"][1][0]
    for col, ss, vs in list(zip(bases, g.base_shifts, gb)) + list(zip(exts, g.ext_shifts, ge)):
        cv = F.e6_from_base(col) if col.ndim == 1 else col
        coeffs = ntt.interpolate(cv.T).T
        for s, v in zip(ss, vs):
            z = F.e6_scale(pv.zeta, pow(w, s, P))
            r = F.e6_zero()
            for c in coeffs[::-1]:
                r = F.e6_add(F.e6_mul(r, z), c)
            assert np.array_equal(r, np.asarray(v, dtype=np.uint64))


# ---------------------------------------------------------------- lookups
def _claimed_by_name(pv, pf):
    vals = {}
    for ti, t in enumerate(pv.layout.trees):
        for gi, g in enumerate(t.groups):
            gb, ge = pf["claimed"][ti][gi]
            for nm, ss, vs in list(zip(g.base, g.base_shifts, gb)) + list(zip(g.ext, g.ext_shifts, ge)):
                for s, v in zip(ss, vs):
                    vals[(nm, s)] = np.asarray(v, dtype=np.uint64)
    return vals


def _air_residual(prog, pv, vals, mn):
    """V(zeta) - (zeta^N - 1) * sum_i zeta^{iN} Q_i(zeta) for module mn, with
    every verifier column evaluated from its definition at zeta."""
    from loom_amd import expr as E, protocol as PR
    from oracle import prover as OP
    zeta = pv.zeta
    m = prog.modules[mn]
    n = m.N
    w = F.omega(n)

    def leaf(x):
        if x.type == E.CHALLENGE:
            return pv.ext[x.name][0][None, :]
        if x.type == E.VERIFIER and x.hook == E.EXPOSED_AVERAGE_HOOK:
            return pv.exposed_avg[x.name][None, :]
        if x.type == E.VERIFIER:
            k = int(x.name.rsplit("_", 1)[1])
            i = k % n
            wi = pow(w, i, P)
            num = F.e6_scale(F.e6_sub(F.e6_pow(zeta, n), F.e6_one()), wi)
            den = F.e6_scale(F.e6_sub(zeta, F.e6_from_base(np.array([wi], dtype=np.uint64))[0]), n)
            return F.e6_mul(num, F.e6_inv(den))[None, :]
        return vals[(x.name, x.shift % n)][None, :]

    ev = OP.eval_expr(m.relation, leaf, 1)[0]
    acc = F.e6_zero()
    for i in range(PR.air_chunk_count(m)):
        q = vals[(PR.quotient_chunk_name(mn, i), 0)]
        acc = F.e6_add(acc, F.e6_mul(F.e6_pow(zeta, i * n), q))
    rhs = F.e6_mul(F.e6_sub(F.e6_pow(zeta, n), F.e6_one()), acc)
    return F.e6_sub(ev, rhs)


@pytest.fixture(scope="module")
def range_proof():
    from loom_amd import workloads
    from oracle import prover as OP
    prog, tr = workloads.uneven_range_instance([4, 6, 3], 9, 256)
    pv = OP.Prover(prog, tr, 6)
    pf = pv.prove()
    return prog, tr, pv, pf


def test_range_lookup_multiplicities_brute_force(range_proof):
    """Each Mult column of the range module counts its source column's values
    (the range table holds 0..255 once each)."""
    prog, tr, pv, pf = range_proof
    mult_steps = [st for st in prog.rounds[0].steps if st.kind == 2]
    assert len(mult_steps) == 3
    for st, i in zip(mult_steps, range(3)):
        want = np.bincount(np.asarray(tr["L_%d" % i], dtype=np.int64), minlength=256)
        assert np.array_equal(pv.base[st.outs[0]], want.astype(np.uint64))


def test_logup_bus_balances_and_air_holds(range_proof):
    """Every bus of logup totals sums to zero and every module's relation
    (plonk gates, copy constraint, logup columns with the exposed average)
    matches its quotient at zeta."""
    prog, tr, pv, pf = range_proof
    for bus in prog.logup_bus:
        s = F.e6_zero()
        for nm in bus:
            s = F.e6_add(s, pf["exposed"][nm])
        assert not np.any(s)
    vals = _claimed_by_name(pv, pf)
    for mn in sorted(prog.modules):
        assert not np.any(_air_residual(prog, pv, vals, mn)), mn


def test_logup_planted_wrong_multiplicity_breaks_bus():
    """A multiplicity off by one in the oracle unbalances the bus."""
    from loom_amd import workloads
    from oracle import prover as OP
    prog, tr = workloads.uneven_range_instance([4], 2, 256)
    pv = OP.Prover(prog, tr, 4)
    orig = pv.lookup_multiplicities

    def bad(st):
        orig(st)
        c = pv.base[st.outs[0]]
        c[0] = (c[0] + 1) % P

    pv.lookup_multiplicities = bad
    pf = pv.prove()
    s = F.e6_zero()
    for nm in prog.logup_bus[0]:
        s = F.e6_add(s, pf["exposed"][nm])
    assert np.any(s)


def test_lookup_rejects_value_outside_table():
    from loom_amd import workloads
    from oracle import prover as OP
    prog, tr = workloads.uneven_range_instance([3], 4, 256)
    tr = dict(tr)
    tr["L_0"] = tr["L_0"].copy()
    tr["L_0"][3] = 300
    with pytest.raises(ValueError, match="not found"):
        OP.prove(prog, tr, 4)
