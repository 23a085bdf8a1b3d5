"""The layout model and the transcript message layout, pinned by literal
expectations that do not come from the code under test.

loom_amd/protocol.py is the one source of tree, group and column order and of
opening shifts for both the product (through the plan words) and the oracle,
so a mistake there would be common to both sides of every parity test.  These
tests hold it to the reference's own literal expectations instead:

* internal/protocol/layout_test.go -- rail-relative PolyIdx (:171-211),
  mixed-size trace round in one tree (:213-249), canonical schedule shifts of
  mixed groups (:251-309), groups follow module order (:385-413), the module
  sizes transcript (:415-436), base-only slot stability with setup columns
  (:31-125) and the mixed-size setup tree (:127-169).
* the config-0 commitment layout of the plonk-like module (T0, T2, AIR).
* internal/fiat-shamir/transcript_test.go:13-78 -- the exact hasher inputs
  of a two-challenge transcript, for the oracle transcript and the host C++
  transcript (tests/host/transcript_record.cpp).
* the registration and binding order of a whole proof's transcript
  (prover/prover.go:208-234, open.go:151-167, fri.go:427-437): every message
  of an oracle proof, checked word by word for its literal parts.

A last group of tests plants the mistakes these checks exist for (two groups
swapped, two columns of a rail swapped, two shift entries swapped) and
requires each literal check to fail.
"""
import copy
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

from loom_amd import board as B
from loom_amd import expr as E
from loom_amd import protocol as PR

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "loom_amd", "csrc")
P = 2**31 - 2**24 + 1
FSID, HASH, SIZE, EXTP = 0x46534944, 0x48415348, 0x53495A45, 0x45585450


# ------------------------------------------------------------------ helpers
def program(modules, rounds, fields=None, setup=()):
    """A compiled program given directly (as the Go tests build
    board.Program): modules {name: (N, relation or None)}, rounds as lists of
    (column, module, field) staged in that order, setup columns as
    (column, module, field)."""
    fields = dict(fields or {})
    mods = {}
    for name, (n, rel) in modules.items():
        r = rel if rel is not None else E.Const(0)
        mods[name] = B.CompiledModule(name, n, r, E.to_dag(r, fields), E.degree(r) if rel is not None else 0, [])
    rds = [B.ProverRound([], [B.ColumnRef(c, m, f) for c, m, f in cols]) for cols in rounds]
    for rd in rds:
        for c in rd.staged:
            fields.setdefault(c.name, c.field)
    return B.Program(mods, rds, fields, [], [B.ColumnRef(c, m, f) for c, m, f in setup])


def col_slots(lay):
    """{column: (tree, group, rail-relative index, field)} -- layout.ColSlot
    and AIRChunkSlot together."""
    out = {}
    for t, tree in enumerate(lay.trees):
        for g, grp in enumerate(tree.groups):
            for i, c in enumerate(grp.base):
                out[c] = (t, g, i, E.BASE)
            for i, c in enumerate(grp.ext):
                out[c] = (t, g, i, E.EXT)
    return out


def tree_groups(lay):
    return [[(g.module, g.N) for g in t.groups] for t in lay.trees]


# ------------------------------------------------------------------ literal checks
def check_rail_relative(lay):
    """layout_test.go:171-211."""
    s = col_slots(lay)
    want = {"base_0": (0, 0, 0, E.BASE), "base_1": (0, 0, 1, E.BASE),
            "ext_0": (0, 1, 0, E.EXT), "ext_1": (0, 1, 1, E.EXT),
            "base.0": (1, 0, 0, E.BASE), "ext.0": (1, 1, 0, E.EXT)}
    for k, v in want.items():
        assert s[k] == v, (k, s[k], v)
    assert tree_groups(lay) == [[("base", 8), ("ext", 8)], [("base", 8), ("ext", 8)]]


def rail_relative_program():
    return program({"base": (8, E.Col("base_air")), "ext": (8, E.ExtCol("ext_air"))},
                   [[("base_0", "base", E.BASE), ("ext_0", "ext", E.EXT),
                     ("base_1", "base", E.BASE), ("ext_1", "ext", E.EXT)]],
                   {"ext_air": E.EXT})


def check_mixed_size_single_tree(lay):
    """layout_test.go:213-249."""
    assert len(lay.trees) == 1
    assert tree_groups(lay) == [[("big", 8), ("small", 4)]]
    s = col_slots(lay)
    want = {"big_0": (0, 0, 0, E.BASE), "big_1": (0, 0, 1, E.BASE),
            "small_0": (0, 1, 0, E.BASE), "small_ext": (0, 1, 0, E.EXT)}
    for k, v in want.items():
        assert s[k] == v, (k, s[k], v)


MIXED_ROUND = [[("small_0", "small", E.BASE), ("big_0", "big", E.BASE),
                ("small_ext", "small", E.EXT), ("big_1", "big", E.BASE)]]


def check_mixed_schedule(lay):
    """layout_test.go:251-309: names per group and rail, normalised shifts."""
    t = lay.trace_tree_of_round[0]
    g0, g1 = lay.trees[t].groups
    assert (g0.module, g0.N, g1.module, g1.N) == ("big", 8, "small", 4)
    assert g0.base == ["big_0", "big_1"]
    assert g1.base == ["small_0"] and g1.ext == ["small_ext"]
    assert g0.base_shifts == [[0], [1]]
    assert g1.base_shifts == [[3]]
    assert g1.ext_shifts == [[0]]


def mixed_schedule_program():
    return program({"big": (8, E.Col("big_0") + E.Col("big_1", 1)),
                    "small": (4, E.Col("small_0", -1) + E.ExtCol("small_ext"))},
                   MIXED_ROUND, {"small_ext": E.EXT})


def check_config0_layout(lay):
    """The plonk-like module (a, b, c = L_0, R_0, O_0): T0 holds the wires
    and the permutation support and sigma columns in staging order, round 1
    commits nothing, T2 the five selectors in sorted-name order on the base
    rail and the grand product on the ext rail (opened at 0 and 1), the AIR
    tree the four extension chunks (degree 4 folded relation)."""
    assert lay.trace_tree_of_round == {0: 0, 2: 1}
    assert [t.kind for t in lay.trees] == ["trace", "trace", "air"]
    (t0,), (t2,), (air,) = [t.groups for t in lay.trees]
    assert t0.base == ["L_0", "plonk_0.ID_0_16", "R_0", "plonk_0.ID_1_16", "O_0", "plonk_0.ID_2_16",
                       "plonk.S_0_0", "plonk.S_0_1", "plonk.S_0_2"] and t0.ext == []
    assert t0.base_shifts == [[0]] * 9
    assert t2.base == ["QK_0", "QL_0", "QM_0", "QO_0", "QR_0"] and t2.ext == ["gp_1"]
    assert t2.base_shifts == [[0]] * 5 and t2.ext_shifts == [[0, 1]]
    assert air.base == [] and air.ext == ["plonk_0.0", "plonk_0.1", "plonk_0.2", "plonk_0.3"]
    assert air.ext_shifts == [[0]] * 4
    n_claims = sum(len(s) for t in lay.trees for g in t.groups for s in g.base_shifts + g.ext_shifts)
    assert n_claims == 20


def config0_program():
    from loom_amd import workloads
    return workloads.plonk_instance(4, 0)[0]


# ------------------------------------------------------------------ layout tests
def test_rail_relative_poly_idx():
    check_rail_relative(PR.build_layout(rail_relative_program()))


def test_mixed_size_trace_round_uses_single_tree():
    check_mixed_size_single_tree(PR.build_layout(program({"big": (8, None), "small": (4, None)}, MIXED_ROUND,
                                                         {"small_ext": E.EXT})))


def test_canonical_schedule_mixed_groups():
    check_mixed_schedule(PR.build_layout(mixed_schedule_program()))


@pytest.mark.parametrize("sizes", [(4, 8, 16), (16, 8, 4), (8, 8, 8)])
def test_groups_follow_module_order(sizes):
    """layout_test.go:385-413."""
    prog = program({"a": (sizes[0], None), "m": (sizes[1], None), "z": (sizes[2], None)},
                   [[("z_0", "z", E.BASE), ("a_0", "a", E.BASE), ("m_0", "m", E.BASE)]])
    lay = PR.build_layout(prog)
    assert tree_groups(lay) == [[("a", sizes[0]), ("m", sizes[1]), ("z", sizes[2])]]
    s = col_slots(lay)
    assert [s[n][1] for n in ("a_0", "m_0", "z_0")] == [0, 1, 2]


def test_base_only_slot_stability():
    """layout_test.go:31-125: a module with setup columns and a lookup
    multiplicity step compiled by the builder: setup columns sorted by name in
    tree 0, the unstaged trace columns swept in name order into tree 1, the
    AIR chunk in tree 2."""
    b = B.Builder()
    m = B.Module("m", 8)
    mul = B.chunk_name("mul", 0)
    m.assert_zero(E.Col("z") + E.Col(mul) + E.Col("s") + E.Col("t") + E.Setup("pz") + E.Setup("pa") + E.Setup("pm"))
    b.add_module(m)
    b.add_lookup_multiplicity_step(0, None, None, [[E.Col("s")]], [[E.Col("t")]], "mul", ["m"], ["m"])
    lay = PR.build_layout(B.compile_program(b))
    want = {"pa": (0, 0, 0, E.BASE), "pm": (0, 0, 1, E.BASE), "pz": (0, 0, 2, E.BASE),
            mul: (1, 0, 0, E.BASE), "s": (1, 0, 1, E.BASE), "t": (1, 0, 2, E.BASE), "z": (1, 0, 3, E.BASE),
            "m.0": (2, 0, 0, E.BASE)}
    assert col_slots(lay) == want
    assert len(lay.trees) == 3 and [t.kind for t in lay.trees] == ["setup", "trace", "air"]
    assert tree_groups(lay) == [[("m", 8)], [("m", 8)], [("m", 8)]]


def test_mixed_size_setup_uses_single_tree():
    """layout_test.go:127-169."""
    prog = program({"big": (8, None), "small": (4, None)}, [],
                   setup=[("pb", "big", E.BASE), ("ps", "small", E.BASE)])
    lay = PR.build_layout(prog)
    assert len(lay.trees) == 1 and lay.trees[0].kind == "setup"
    assert tree_groups(lay) == [[("big", 8), ("small", 4)]]
    s = col_slots(lay)
    assert s["pb"] == (0, 0, 0, E.BASE) and s["ps"] == (0, 1, 0, E.BASE)


def test_module_sizes_transcript():
    """layout_test.go:415-436: [SIZE tag, #modules, N of a, N of b]."""
    prog = lambda a, b: program({"b": (b, None), "a": (a, None)}, [[]])
    assert PR.module_sizes_transcript(prog(8, 16)) == [0x53495A45, 2, 8, 16]
    assert PR.module_sizes_transcript(prog(8, 16)) != PR.module_sizes_transcript(prog(16, 8))


def test_config0_layout():
    check_config0_layout(PR.build_layout(config0_program()))


# ------------------------------------------------------------------ transcript messages
ZETA_MSG = [FSID, 4, 0x74657A, 0x61, 11]                       # "zeta": "zet", "a"
ALPHA_PREFIX = [FSID, 10, 0x706C61, 0x5F6168, 0x454544, 0x50]  # "alpha_DEEP": "alp", "ha_", "DEE", "P"
ALPHA_MSG = ALPHA_PREFIX + [101, 102, 103, 104, 105, 106, 107, 108] + [22]


def test_oracle_transcript_records_challenge_id():
    """transcript_test.go:43-78 on the oracle transcript."""
    from oracle.transcript import Transcript
    inputs = []

    def rec(msg):
        inputs.append([int(x) for x in msg])
        return [100 * len(inputs) + i + 1 for i in range(8)]

    t = Transcript(["zeta", "alpha_DEEP"], hasher=rec)
    t.bind("zeta", [11])
    t.compute("zeta")
    t.bind("alpha_DEEP", [22])
    t.compute("alpha_DEEP")
    assert inputs == [ZETA_MSG, ALPHA_MSG]


def test_host_transcript_records_challenge_id():
    """transcript_test.go:43-78 on transcript.h (by name and by position)."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    try:
        exe = os.path.join(d, "transcript_record")
        subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(HERE, "host", "transcript_record.cpp"),
                        "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    finally:
        shutil.rmtree(d)
    lines = [[int(x) for x in ln.split()] for ln in out.strip().splitlines()]
    assert lines == [ZETA_MSG, ALPHA_MSG, ZETA_MSG, ALPHA_MSG]


def _chunks3(s):
    b = s.encode()
    return [int.from_bytes(b[i:i + 3], "little") for i in range(0, len(b), 3)]


def _prefix(name):
    return [FSID, len(name)] + _chunks3(name)


def test_proof_transcript_binding_order():
    """Every message hashed by an oracle proof of the plonk-like module
    (N = 16, 6 queries), checked against the registration and binding order
    of prover.go:208-234, 360-412, 545-554, open.go:151-167 and
    fri.go:427-437, 606-736: the literal parts word by word, the data parts
    against the proof that carries them."""
    from loom_amd import workloads
    from oracle import prover as OP
    from oracle import verifier as V
    prog, tr = workloads.plonk_instance(4, 0)
    nq = 6
    pv = OP.Prover(prog, tr, nq)
    pf = V.decode_proof(OP.encode_proof(pv.prove()))
    ts = pv.ts
    names = ["challenge@loom_0", "challenge@loom_1", "challenge@loom_2", "__zeta", "alpha_DEEP"]
    names += ["fri_fold_%d" % j for j in range(22)] + ["fri_query_%d" % k for k in range(nq)]
    assert ts.order == names  # one module: no level gammas
    vals = [ts.values[n] for n in names]
    roots = [list(map(int, r)) for r in pf["commitments"]]
    to_m = lambda v: [int(x) for x in v]
    msgs = [ts.message(n) for n in names]
    # challenge@loom_0: HASH id elements, module sizes, then the T0 root
    assert msgs[0] == _prefix(names[0]) + [HASH, 9] + _chunks3("poseidon2") + [SIZE, 1, 16] + roots[0]
    # challenge@loom_1: nothing committed in round 1
    assert msgs[1] == _prefix(names[1]) + vals[0]
    # challenge@loom_2: the T2 root
    assert msgs[2] == _prefix(names[2]) + vals[1] + roots[1]
    # __zeta: the AIR root
    assert msgs[3] == _prefix(names[3]) + vals[2] + roots[2]
    # alpha_DEEP: the 20 claimed values, tree -> group -> base polys -> ext polys -> shifts
    cv = []
    for groups in pf["claimed"]:
        for base, ext in groups:
            for rail in (base, ext):
                for shifts in rail:
                    for v in shifts:
                        cv += to_m(v)
    assert len(cv) == 20 * 6
    assert msgs[4] == _prefix(names[4]) + vals[3] + cv
    # FRI: N = 16 enters at round r0 = 22 - 4 = 18; rounds before it bind the
    # zero digest, round r0 the levels root, later rounds their layer roots
    r0 = 18
    fr = pf["fri"]
    for j in range(22):
        m = msgs[5 + j]
        head = _prefix(names[5 + j]) + vals[4 + j]
        if j < r0:
            assert m == head + [0] * 8, j
        elif j == r0:
            assert m == head + to_m(fr["levels_root"]), j
        else:
            assert m == head + to_m(fr["roots"][j - r0 - 1]), j
    # fri_query_0: [EXTP, #final, final values]; query k+1: query k's value
    fin = []
    for v in fr["final"]:
        fin += to_m(v)
    q0 = 5 + 22
    assert msgs[q0] == _prefix(names[q0]) + vals[q0 - 1] + [EXTP, 4] + fin
    for k in range(1, nq):
        assert msgs[q0 + k] == _prefix(names[q0 + k]) + vals[q0 + k - 1] + vals[q0 + k - 1]
    # the oracle's messages are canonical values; the host transcript's the
    # same in Montgomery form (test_host_code.py), the device's equal digests


# ------------------------------------------------------------------ planted mistakes
def _swap_groups(lay):
    for t in lay.trees:
        if len(t.groups) >= 2:
            t.groups[0], t.groups[1] = t.groups[1], t.groups[0]
            return lay
    raise AssertionError("no tree with two groups")


def _swap_rail_columns(lay):
    for t in lay.trees:
        for g in t.groups:
            for rail, sh in ((g.base, g.base_shifts), (g.ext, g.ext_shifts)):
                if len(rail) >= 2:
                    rail[0], rail[1] = rail[1], rail[0]
                    sh[0], sh[1] = sh[1], sh[0]
                    return lay
    raise AssertionError("no rail with two columns")


def _swap_shift_entries(lay):
    for t in lay.trees:
        for g in t.groups:
            for sh in g.base_shifts + g.ext_shifts:
                if len(sh) >= 2:
                    sh[0], sh[1] = sh[1], sh[0]
                    return lay
    # no polynomial with two shifts: swap the shift lists of two polynomials
    for t in lay.trees:
        for g in t.groups:
            for sh in (g.base_shifts, g.ext_shifts):
                for i in range(len(sh) - 1):
                    if sh[i] != sh[i + 1]:
                        sh[i], sh[i + 1] = sh[i + 1], sh[i]
                        return lay
    raise AssertionError("no shifts to swap")


@pytest.mark.parametrize("check,make,plant", [
    (check_rail_relative, rail_relative_program, _swap_groups),
    (check_rail_relative, rail_relative_program, _swap_rail_columns),
    (check_mixed_schedule, mixed_schedule_program, _swap_groups),
    (check_mixed_schedule, mixed_schedule_program, _swap_rail_columns),
    (check_mixed_schedule, mixed_schedule_program, _swap_shift_entries),
    (check_config0_layout, config0_program, _swap_rail_columns),
    (check_config0_layout, config0_program, _swap_shift_entries),
])
def test_planted_layout_mistake_fails(check, make, plant):
    lay = PR.build_layout(make())
    check(lay)
    bad = plant(copy.deepcopy(lay))
    with pytest.raises(AssertionError):
        check(bad)
