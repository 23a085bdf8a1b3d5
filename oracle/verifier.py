"""Proof checker on the CPU (test infrastructure, like the rest of oracle/).

Decodes the canonical proof word stream (the layout ``prover.encode_proof``
writes and include/loomgpu.h documents) and checks it the way the
reference's verifier does, so that a proof too large for the oracle to
recompute can still be checked at full size:

* Fiat-Shamir replay -- every challenge recomputed from the proof's
  commitments, exposed values, claimed values, FRI roots and final
  polynomial (prover/prover.go:199-234, 360-412, 528-558; fri/open.go:151-167,
  223-285; fri/fri.go:427-437, 1031-1051, 1216-1222);
* logup buses -- the exposed totals of each bus sum to zero
  (recursion/verifier.go:1936-1950 checkLogupBus);
* AIR identity at zeta -- each module's folded relation from the claimed
  values equals (zeta^N - 1) * sum_i zeta^{iN} Q_i(zeta), with the verifier
  columns (Lagrange, exposed averages) computed here (verifier.go:406-444);
* openings -- every opened row pair hashes to a leaf whose path (with the
  injected groups' rows as level injections) reaches its commitment
  (merkle/tree.go:392-436, fri/open.go:290-416);
* DEEP -- at each queried row the DEEP quotient recomputed from the opened
  rows and claimed values equals the FRI level value (fri/deep.go:71-266);
* FRI -- each opened layer folds into the next (plus the levels entering
  there), the last fold lands on the final polynomial, and every layer pair
  opens against its root (fri/fri.go:1321-1534).

``verify`` raises ``ProofError`` with the first failed check.
"""
import numpy as np

from . import field as F
from .field import P, U64
from . import poseidon2 as p2
from .transcript import Transcript, string_to_elements
from .prover import FRI_N, FRI_D, FRI_ROUNDS, INV2, HASH_TAG, EXTP_TAG, eval_expr, log2

from loom_amd import expr as E
from loom_amd import board as B
from loom_amd import protocol as PR


class ProofError(AssertionError):
    pass


def check(cond, what):
    if not cond:
        raise ProofError(what)


# ------------------------------------------------------------------ decode
class _Reader:
    def __init__(self, words):
        self.w = [int(x) for x in np.asarray(words, dtype=np.uint32)]
        self.i = 0

    def u(self):
        check(self.i < len(self.w), "proof words truncated")
        v = self.w[self.i]
        self.i += 1
        return v

    def f(self):
        return F.from_mont(self.u())

    def e6(self):
        return [self.f() for _ in range(6)]

    def digest(self):
        return [self.f() for _ in range(8)]

    def string(self):
        n = self.u()
        b = b"".join(self.u().to_bytes(4, "little") for _ in range((n + 3) // 4))
        return b[:n].decode()

    def path(self):
        idx = self.u()
        sibs = [self.digest() for _ in range(self.u())]
        injl = [self.digest() for _ in range(self.u())]
        return idx, sibs, injl

    def row(self):
        bases = [self.f() for _ in range(self.u())]
        exts = [self.e6() for _ in range(self.u())]
        return bases, exts


def decode_proof(words):
    """Inverse of prover.encode_proof; values canonical."""
    r = _Reader(words)
    pf = {"hash_id": r.string()}
    ex = {}
    for _ in range(r.u()):
        name = r.string()
        ents = [(r.u(), r.e6()) for _ in range(r.u())]
        check(len(ents) == 1 and ents[0][0] == 0, "exposed value %s: one entry at index 0 expected" % name)
        ex[name] = np.array(ents[0][1], dtype=np.uint64)
    pf["exposed"] = ex
    pf["commitments"] = [r.digest() for _ in range(r.u())]
    claimed = []
    for _ in range(r.u()):
        groups = []
        for _ in range(r.u()):
            rails = []
            for _ in range(2):
                rails.append([[r.e6() for _ in range(r.u())] for _ in range(r.u())])
            groups.append(tuple(rails))
        claimed.append(groups)
    pf["claimed"] = claimed
    fr = {"levels_root": r.digest()}
    lq = []
    for _ in range(r.u()):
        qs = []
        for _ in range(r.u()):
            check(r.u() == 1, "level query: extension flag")
            row, p, q = r.u(), r.e6(), r.e6()
            check(r.u() == 0 and r.u() == 0 and r.u() == 0, "level query: no path expected")
            qs.append(dict(row=row, p=p, q=q))
        lq.append(qs)
    fr["level_queries"] = lq
    fr["roots"] = [r.digest() for _ in range(r.u())]
    check(r.u() == 1, "final polynomial: extension flag")
    fr["final"] = [r.e6() for _ in range(r.u())]
    queries = []
    for _ in range(r.u()):
        layers = []
        for _ in range(r.u()):
            check(r.u() == 1, "query layer: extension flag")
            layers.append(dict(row=r.u(), p=r.e6(), q=r.e6(), path=r.path()))
        queries.append(layers)
    fr["queries"] = queries
    pw = {}
    for _ in range(r.u()):  # name, NbBits, Salt (fri.Proof.PoW, sorted by name)
        name = r.string()
        check(name not in pw, "proof of work listed twice")
        pw[name] = (r.u(), r.f())
    fr["pow"] = pw
    pf["fri"] = fr
    samplings = []
    for _ in range(r.u()):
        row = []
        for _ in range(r.u()):
            lo, hi = r.row(), r.row()
            path = r.path()
            inj = [(r.row(), r.row()) for _ in range(r.u())]
            row.append(dict(top=(lo, hi), path=path, inj=inj))
        samplings.append(row)
    pf["samplings"] = samplings
    check(r.i == len(r.w), "trailing proof words")
    return pf


# ------------------------------------------------------------------ helpers
def _leaves(lb, le, hb, he):
    """Leaf digests (8, Q) of Q row pairs: lb/hb (Q, nb) base values, le/he
    (Q, ne, 6) extension values."""
    Q, nb = lb.shape
    ne = le.shape[1]
    msg = np.concatenate([np.full((Q, 1), p2.LEAF_TAG, np.uint64), np.full((Q, 1), 2 * nb, np.uint64),
                          np.full((Q, 1), 2 * ne, np.uint64), lb, hb, le.reshape(Q, 6 * ne),
                          he.reshape(Q, 6 * ne)], axis=1)
    return p2.sponge(msg.T)


def _paths_root(leaves, paths, n_leaves, inj_widths):
    """Roots (8, Q) reached from leaf digests (8, Q) through Q paths, folding
    injection leaves at the levels of their widths (merkle/tree.go:392-436)."""
    Q = leaves.shape[1]
    depth = log2(n_leaves)
    check(all(len(pth[1]) == depth and len(pth[2]) == len(inj_widths) for pth in paths), "path shape")
    idx = np.array([pth[0] for pth in paths], dtype=np.int64)
    h = leaves.copy()
    nxt = 0

    def fold(width):
        nonlocal h, nxt
        while nxt < len(inj_widths) and inj_widths[nxt] == width:
            injv = np.array([pth[2][nxt] for pth in paths], dtype=np.uint64).reshape(Q, 8).T
            h = p2.node_compress(h, injv)
            nxt += 1

    fold(n_leaves)
    for k in range(depth):
        sib = np.array([pth[1][k] for pth in paths], dtype=np.uint64).reshape(Q, 8).T
        even = (idx & 1) == 0
        left = np.where(even[None, :], h, sib)
        right = np.where(even[None, :], sib, h)
        h = p2.node_compress(left, right)
        idx >>= 1
        fold(n_leaves >> (k + 1))
    return h


def _pow_each(base, exps):
    """base^e for every e of exps (scalar canonical base, uint64 exponents)."""
    exps = np.asarray(exps, dtype=np.uint64).copy()
    r = np.ones(exps.shape, dtype=np.uint64)
    b = U64(base % P)
    while np.any(exps):
        r = np.where(exps & U64(1), r * b % U64(P), r)
        b = b * b % U64(P)
        exps >>= U64(1)
    return r


def _lde_points(n_lde, rows):
    b = log2(n_lde)
    rb = np.array([int(format(int(r), "0%db" % b)[::-1], 2) if b else 0 for r in rows], dtype=np.uint64)
    return _pow_each(F.omega(n_lde), rb)


def _e(v):
    return np.asarray(v, dtype=np.uint64)


# ------------------------------------------------------------------ verify
def verify(program, num_queries, words, check_queries=None, setup_roots=(), public_inputs=None, grinding=0,
           fs_prev=None):
    """Checks proof words against the program; returns the decoded proof.
    check_queries bounds how many queries have their openings checked (all by
    default); the transcript replay always covers every query.  setup_roots:
    the verification key (setup.VerificationKey.Roots); public_inputs: name ->
    (module, [(index, value), ...]); grinding: the FRI parameters' proof-of-work
    bits (every fold challenge then carries a proof of work, fri.go:1060-1073);
    fs_prev: the running transcript's last digest (verifier.WithTranscript,
    verifier.go:60-65), None for a fresh one."""
    pf = decode_proof(words)
    lay = PR.build_layout(program)
    prog = program
    pi = dict(public_inputs or {})
    check(pf["hash_id"] == "poseidon2", "hash id")
    nt = len(lay.trees)
    ns = PR.setup_tree_count(lay)
    check(len(setup_roots) == ns, "verification key: setup root count")
    check(len(pf["commitments"]) == nt - ns, "commitment count")
    # roots by canonical tree index: setup roots from the key, then the proof's
    roots_by_tree = [list(map(int, r)) for r in setup_roots] + list(pf["commitments"])
    check(len(pf["claimed"]) == nt, "claimed values: tree count")
    exposed_names = sorted(o for rd in prog.rounds for st in rd.steps for o in st.exposed_outputs())
    check(sorted(pf["exposed"]) == exposed_names, "exposed value names")

    # ---- Fiat-Shamir replay
    nr = len(prog.rounds)
    names = [B.challenge_name(r) for r in range(nr)] + ["__zeta"]
    ts = Transcript(names)
    if fs_prev is not None:
        ts.chain_from(fs_prev)
    ts.bind(names[0], string_to_elements(HASH_TAG, "poseidon2"))
    ts.bind(names[0], PR.module_sizes_transcript(prog))
    for r in setup_roots:
        ts.bind(names[0], list(map(int, r)))
    if pi:
        ts.bind(names[0], PR.public_input_elements(pi))
    chal = {}
    for r, rd in enumerate(prog.rounds):
        cn = B.challenge_name(r)
        if r in lay.trace_tree_of_round:
            ts.bind(cn, roots_by_tree[lay.trace_tree_of_round[r]])
        if rd.fshook == B.BIND_EXPOSED_VALUES:
            for st in rd.steps:
                for name in st.exposed_outputs():
                    ts.bind(cn, [0] + [int(x) for x in pf["exposed"][name]])
        chal[cn] = _e(ts.compute(cn)[:6])
    air_trees = [i for i, t in enumerate(lay.trees) if t.kind == "air"]
    for t in air_trees:
        ts.bind("__zeta", roots_by_tree[t])
    zeta = _e(ts.compute("__zeta")[:6])

    # claimed values by (tree, group) and by (name, shift)
    vals = {}
    for ti, t in enumerate(lay.trees):
        check(len(pf["claimed"][ti]) == len(t.groups), "claimed values: group count")
        for gi, g in enumerate(t.groups):
            gb, ge = pf["claimed"][ti][gi]
            check(len(gb) == len(g.base) and len(ge) == len(g.ext), "claimed values: polynomial count")
            for nm, ss, vs in list(zip(g.base, g.base_shifts, gb)) + list(zip(g.ext, g.ext_shifts, ge)):
                check(len(vs) == len(ss), "claimed values: shift count of " + nm)
                for s, v in zip(ss, vs):
                    vals[(nm, s)] = _e(v)

    # ---- logup buses
    for bus in prog.logup_bus:
        s = F.e6_zero()
        for nm in bus:
            s = F.e6_add(s, pf["exposed"][nm])
        check(not np.any(s), "logup bus does not balance")

    # ---- AIR identity at zeta
    for mn in sorted(prog.modules):
        m = prog.modules[mn]
        if PR.air_chunk_count(m) == 0:
            continue
        n = m.N
        w = F.omega(n)
        zn1 = F.e6_sub(F.e6_pow(zeta, n), F.e6_one())

        def leaf(x):
            if x.type == E.CHALLENGE:
                return chal[x.name][None, :]
            if x.type == E.VERIFIER and x.hook == E.EXPOSED_AVERAGE_HOOK:
                return F.e6_scale(pf["exposed"][x.name][None, :], F.inv(n))
            if x.type == E.VERIFIER and x.hook == E.LAGRANGE_HOOK:
                i = int(x.name.rsplit("_", 1)[1]) % n
                wi = pow(w, i, P)
                num = F.e6_scale(zn1, wi * F.inv(n) % P)
                den = F.e6_sub(zeta, F.e6_from_base(np.array([wi], dtype=np.uint64))[0])
                return F.e6_mul(num, F.e6_inv(den))[None, :]
            if x.type == E.VERIFIER and x.hook == E.PUBLIC_INPUT_HOOK:
                # sum_i v_i L_i(zeta) over the entries (verifier/hookFns.go:51-70)
                mod, entries = pi[x.name]
                check(mod == mn, "public input %s used from module %s" % (x.name, mn))
                acc = F.e6_zero()
                for i, v in entries:
                    wi = pow(w, int(i) % n, P)
                    num = F.e6_scale(zn1, wi * F.inv(n) % P)
                    val = np.zeros(6, dtype=np.uint64)
                    ws = PR.public_entry_words(v)
                    val[:len(ws)] = ws
                    den = F.e6_sub(zeta, F.e6_from_base(np.array([wi], dtype=np.uint64))[0])
                    acc = F.e6_add(acc, F.e6_mul(F.e6_mul(num, val), F.e6_inv(den)))
                return acc[None, :]
            check(x.type in (E.COMMITTED, E.SETUP), "unsupported leaf in a relation")
            return vals[(x.name, x.shift % n)][None, :]

        ev = eval_expr(m.relation, leaf, 1)[0]
        acc = F.e6_zero()
        for i in range(PR.air_chunk_count(m)):
            acc = F.e6_add(acc, F.e6_mul(F.e6_pow(zeta, i * n), vals[(PR.quotient_chunk_name(mn, i), 0)]))
        check(np.array_equal(ev, F.e6_mul(zn1, acc)), "AIR identity fails at zeta for module " + mn)

    # ---- alpha_DEEP
    class_of = {mm: i for i, mm in enumerate(lay.modules)}
    sizes = {class_of[mm]: prog.modules[mm].N for mm in lay.modules}
    ts.new_challenge("alpha_DEEP")
    present = sorted(set(class_of[g.module] for t in lay.trees for g in t.groups))
    for cl in present:
        for ti, t in enumerate(lay.trees):
            for gi, g in enumerate(t.groups):
                if class_of[g.module] != cl:
                    continue
                gb, ge = pf["claimed"][ti][gi]
                for vs in gb + ge:
                    for v in vs:
                        ts.bind("alpha_DEEP", v)
    alpha = _e(ts.compute("alpha_DEEP")[:6])

    # ---- FRI transcript
    fr = pf["fri"]
    levels = sorted(range(len(present)), key=lambda i: -sizes[present[i]])  # stable, by size desc
    lv_cls = [present[i] for i in levels]
    nl = len(lv_cls)
    for l in range(1, nl):
        ts.new_challenge("fri_level_%d_gamma" % l)
    for j in range(FRI_ROUNDS):
        ts.new_challenge("fri_fold_%d" % j)
    for k in range(num_queries):
        ts.new_challenge("fri_query_%d" % k)
    first = log2(FRI_D // sizes[lv_cls[0]])
    enter = [log2(FRI_D // sizes[c]) for c in lv_cls]
    gammas = [None] * nl
    for l in range(1, nl):
        nm = "fri_level_%d_gamma" % l
        if l == 1:
            ts.bind(nm, fr["levels_root"])
        gammas[l] = _e(ts.compute(nm)[:6])
    pw = fr["pow"]
    if grinding:
        check(sorted(pw) == sorted("fri_fold_%d" % j for j in range(FRI_ROUNDS)), "proofs of work: one per fold")
    else:
        check(not pw, "proofs of work: none expected (grinding off)")

    def fold_challenge(name):
        if not grinding:
            return ts.compute(name)
        bits, salt = pw[name]
        check(bits == grinding, "proof of work: bit count")
        try:
            return ts.compute_with_pow(name, bits, salt)
        except ValueError:
            check(False, "invalid proof of work for " + name)

    for j in range(first):
        ts.bind("fri_fold_%d" % j, [0] * 8)
        fold_challenge("fri_fold_%d" % j)
    check(len(fr["roots"]) == FRI_ROUNDS - first - 1, "FRI root count")
    alphas = {}
    for j in range(first, FRI_ROUNDS):
        root = fr["levels_root"] if j == first else fr["roots"][j - first - 1]
        ts.bind("fri_fold_%d" % j, root)
        alphas[j] = _e(fold_challenge("fri_fold_%d" % j)[:6])
    final = [_e(v) for v in fr["final"]]
    check(len(final) == FRI_N >> FRI_ROUNDS, "final polynomial length")
    ts.bind("fri_query_0", [EXTP_TAG, len(final)] + [int(x) for v in final for x in v])
    positions = []
    for k in range(num_queries):
        ch = ts.compute("fri_query_%d" % k)
        positions.append(((ch[0] << 31) ^ ch[1]) % FRI_N)
        if k < num_queries - 1:
            ts.bind("fri_query_%d" % (k + 1), ch)
    check(len(fr["queries"]) == num_queries and len(pf["samplings"]) == num_queries, "query count")
    check(len(fr["level_queries"]) == nl - 1, "level query count")

    # ---- openings, DEEP and FRI, vectorised over the checked queries
    Q = num_queries if check_queries is None else min(check_queries, num_queries)
    pos = np.array(positions[:Q], dtype=np.int64)
    deep_polys = {cl: [] for cl in present}  # (tree, group) per class, canonical order
    for cl in present:
        for ti, t in enumerate(lay.trees):
            for gi, g in enumerate(t.groups):
                if class_of[g.module] == cl:
                    deep_polys[cl].append((ti, gi))
    rows = {}  # (tree, group) -> (lo base (Q,nb), lo ext (Q,ne,6), hi base, hi ext, lo row index (Q,))
    for ti, t in enumerate(lay.trees):
        order = sorted(range(len(t.groups)), key=lambda gi: -t.groups[gi].N)
        n_top = PR.RATE * t.groups[order[0]].N
        lo_idx = (pos >> (log2(FRI_N) - log2(n_top))) & ~1
        wps = [pf["samplings"][k][ti] for k in range(Q)]
        check(all(wp["path"][0] == lo_idx[k] // 2 for k, wp in enumerate(wps)), "opened leaf index")
        check(all(len(wp["inj"]) == len(order) - 1 for wp in wps), "injected row count")
        groups_rows = [(order[0], [wp["top"] for wp in wps], lo_idx)]
        for gpos, gi in enumerate(order[1:]):
            n_lde = PR.RATE * t.groups[gi].N
            r2 = (pos >> (log2(FRI_N) - log2(n_lde))) & ~1
            groups_rows.append((gi, [wp["inj"][gpos] for wp in wps], r2))
        leaves = []
        for gi, pairs, ridx in groups_rows:
            g = t.groups[gi]
            lb = np.array([pr[0][0] for pr in pairs], dtype=np.uint64).reshape(Q, len(g.base))
            le = np.array([pr[0][1] for pr in pairs], dtype=np.uint64).reshape(Q, len(g.ext), 6)
            hb = np.array([pr[1][0] for pr in pairs], dtype=np.uint64).reshape(Q, len(g.base))
            he = np.array([pr[1][1] for pr in pairs], dtype=np.uint64).reshape(Q, len(g.ext), 6)
            rows[(ti, gi)] = (lb, le, hb, he, ridx)
            leaves.append(_leaves(lb, le, hb, he))
        inj_w = [PR.RATE * t.groups[gi].N // 2 for gi in order[1:]]
        for j in range(1, len(leaves)):
            got = np.array([wp["path"][2][j - 1] for wp in wps], dtype=np.uint64).reshape(Q, 8).T
            check(np.array_equal(got, leaves[j]), "injection leaves do not hash the injected rows of tree %d" % ti)
        roots = _paths_root(leaves[0], [wp["path"] for wp in wps], n_top // 2, inj_w)
        check(np.all(roots == np.array(roots_by_tree[ti], dtype=np.uint64)[:, None]),
              "opening of tree %d does not reach its commitment" % ti)
    # DEEP quotient at the two rows of each pair, per class
    deep_at = {}
    for cl in present:
        n = sizes[cl]
        w = F.omega(n)
        out = []
        for half in range(2):
            a = F.e6_one()
            acc = np.zeros((Q, 6), dtype=np.uint64)
            for ti, gi in deep_polys[cl]:
                g = lay.trees[ti].groups[gi]
                lb, le, hb, he, r0 = rows[(ti, gi)]
                xb, xe = (lb, le) if half == 0 else (hb, he)
                x = F.e6_from_base(_lde_points(PR.RATE * n, r0 + half))
                gb, ge = pf["claimed"][ti][gi]
                polys = [(F.e6_from_base(xb[:, c]), g.base_shifts[c], gb[c]) for c in range(len(g.base))]
                polys += [(xe[:, c, :], g.ext_shifts[c], ge[c]) for c in range(len(g.ext))]
                for px, ss, vs in polys:
                    for sh, v in zip(ss, vs):
                        z = F.e6_scale(zeta, pow(w, sh % n, P))
                        num = F.e6_mul(np.broadcast_to(a, (Q, 6)), F.e6_sub(np.broadcast_to(_e(v), (Q, 6)), px))
                        den = F.e6_sub(np.broadcast_to(z, (Q, 6)), x)
                        acc = F.e6_add(acc, F.e6_mul(num, F.e6_inv(den)))
                    a = F.e6_mul(a, alpha)
            out.append(acc)
        deep_at[cl] = out
    # FRI levels: level 0 is layer `first`, the others are the level queries
    qs = [fr["queries"][k] for k in range(Q)]
    check(all(len(q) == FRI_ROUNDS - first for q in qs), "query layer count")
    lv = []
    for l in range(nl):
        rowl = pos >> enter[l]
        if l == 0:
            rr = np.array([q[0]["row"] for q in qs])
            pl = np.array([q[0]["p"] for q in qs], dtype=np.uint64)
            ql = np.array([q[0]["q"] for q in qs], dtype=np.uint64)
        else:
            lqs = fr["level_queries"][l - 1][:Q]
            rr = np.array([x["row"] for x in lqs])
            pl = np.array([x["p"] for x in lqs], dtype=np.uint64)
            ql = np.array([x["q"] for x in lqs], dtype=np.uint64)
        check(np.array_equal(rr, rowl), "level %d row" % l)
        check(np.array_equal(pl, deep_at[lv_cls[l]][0]) and np.array_equal(ql, deep_at[lv_cls[l]][1]),
              "DEEP quotient differs from FRI level %d" % l)
        lv.append((pl, ql))
    inj = [_leaves(np.zeros((Q, 0), np.uint64), lv[l][0][:, None, :], np.zeros((Q, 0), np.uint64),
                   lv[l][1][:, None, :]) for l in range(1, nl)]
    check(all(q[0]["path"][2] == [] for q in qs), "layer-%d path carries no injection leaves" % first)
    paths0 = [(q[0]["path"][0], q[0]["path"][1], [[int(v) for v in inj[l][:, k]] for l in range(nl - 1)])
              for k, q in enumerate(qs)]
    roots = _paths_root(_leaves(np.zeros((Q, 0), np.uint64), lv[0][0][:, None, :], np.zeros((Q, 0), np.uint64),
                                lv[0][1][:, None, :]), paths0, 2 * sizes[lv_cls[0]],
                        [2 * sizes[lv_cls[l]] for l in range(1, nl)])
    check(np.all(roots == np.array(fr["levels_root"], dtype=np.uint64)[:, None]), "levels tree opening")
    for i in range(FRI_ROUNDS - first):
        j = first + i
        row = pos >> j
        check(all(q[i]["row"] == row[k] for k, q in enumerate(qs)), "layer row")
        pp = np.array([q[i]["p"] for q in qs], dtype=np.uint64)
        qq = np.array([q[i]["q"] for q in qs], dtype=np.uint64)
        if i > 0:
            check(all(q[i]["path"][0] == (row[k] & ~1) // 2 for k, q in enumerate(qs)), "layer leaf index")
            roots = _paths_root(_leaves(np.zeros((Q, 0), np.uint64), pp[:, None, :], np.zeros((Q, 0), np.uint64),
                                        qq[:, None, :]), [q[i]["path"] for q in qs], FRI_N >> (j + 1), [])
            check(np.all(roots == np.array(fr["roots"][i - 1], dtype=np.uint64)[:, None]), "FRI layer %d opening" % j)
        else:
            # the opened layer is level 0 alone; levels of its size join the
            # running codeword before the first fold
            for l in range(1, nl):
                if enter[l] == first:
                    pp = F.e6_add(pp, F.e6_mul(np.broadcast_to(gammas[l], (Q, 6)), lv[l][0]))
                    qq = F.e6_add(qq, F.e6_mul(np.broadcast_to(gammas[l], (Q, 6)), lv[l][1]))
        n_j = FRI_N >> j
        half = n_j // 2
        parent = row >> 1
        xinv = _pow_each(F.inv(F.omega(n_j)),
                         [int(format(int(pa), "0%db" % log2(half))[::-1], 2) for pa in parent])
        ssum = F.e6_scale(F.e6_add(pp, qq), np.full(Q, INV2, dtype=np.uint64))
        diff = F.e6_mul(F.e6_scale(F.e6_sub(pp, qq), xinv * U64(INV2) % U64(P)), np.broadcast_to(alphas[j], (Q, 6)))
        nxt = F.e6_add(ssum, diff)
        for l in range(1, nl):
            if enter[l] == j + 1:
                sel = np.where((parent & 1)[:, None] == 0, lv[l][0], lv[l][1])
                nxt = F.e6_add(nxt, F.e6_mul(np.broadcast_to(gammas[l], (Q, 6)), sel))
        if i + 1 < FRI_ROUNDS - first:
            nq_ = np.array([q[i + 1]["p"] for q in qs], dtype=np.uint64)
            nqq = np.array([q[i + 1]["q"] for q in qs], dtype=np.uint64)
            tgt = np.where((parent & 1)[:, None] == 0, nq_, nqq)
        else:
            tgt = np.array([final[int(pa)] for pa in parent], dtype=np.uint64)
        check(np.array_equal(nxt, tgt), "FRI fold inconsistent at round %d" % j)
    return pf
