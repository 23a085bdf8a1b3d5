"""Canonical commitment layout and opening schedule (host-side mirror of
internal/protocol/layout.go:165-241 and schedule.go:92-235).

Tree order: [setup] [trace round 0] ... [trace round r-1] [AIR]; each tree has
one group per module in sorted module order, skipping modules with nothing
to commit; inside a group, columns keep their staging order and get a
rail-relative index.  AIR groups hold the quotient chunks of the module.  The
schedule gives every committed polynomial its sorted, normalised list of
opening shifts (AIR chunks: [0]).
"""
from dataclasses import dataclass, field

import numpy as np

from . import expr as E
from .board import challenge_name

RATE = 4
MAX_LOG_MODULE_SIZE = 22
NUM_QUERIES_DEFAULT = 32


def next_pow2(n):
    p = 1
    while p < n:
        p <<= 1
    return p


def quotient_domain_size(deg, n):
    """poly.QuotientDomainSize (compute_quotient.go:508-510)."""
    return next_pow2(max(deg - 1, 1) * n)


def air_chunk_count(m):
    if m.degree <= 0:
        return 0
    return quotient_domain_size(m.degree, m.N) // m.N


def quotient_chunk_name(module, i):
    return "%s.%d" % (module, i)


@dataclass
class Group:
    module: str
    N: int
    base: list = field(default_factory=list)   # column names, rail order
    ext: list = field(default_factory=list)
    base_shifts: list = field(default_factory=list)
    ext_shifts: list = field(default_factory=list)


@dataclass
class Tree:
    kind: str          # "setup", "trace" or "air"
    round: int         # trace round index (-1 for AIR)
    groups: list


@dataclass
class Layout:
    modules: list
    trees: list
    trace_tree_of_round: dict


def build_layout(program):
    order = sorted(program.modules)
    trees = []
    tree_of_round = {}

    def add_tree(kind, rnd, by_module):
        groups = []
        for m in order:
            cols = by_module.get(m, [])
            if not cols:
                continue
            g = Group(m, program.modules[m].N)
            for name, f in cols:
                (g.ext if f == E.EXT else g.base).append(name)
            groups.append(g)
        if groups:
            trees.append(Tree(kind, rnd, groups))
            return len(trees) - 1
        return None

    # setup section: one tree, one group per module, columns sorted by name
    # (protocol.SetupGroups, layout.go:143-166)
    by = {}
    for c in getattr(program, "setup_columns", []):
        if c.module in program.modules:
            by.setdefault(c.module, []).append((c.name, c.field))
    for m in by:
        by[m].sort()
    add_tree("setup", -1, by)
    for r, rd in enumerate(program.rounds):
        by = {}
        for c in rd.staged:
            if c.module in program.modules:
                by.setdefault(c.module, []).append((c.name, c.field))
        t = add_tree("trace", r, by)
        if t is not None:
            tree_of_round[r] = t
    by = {}
    for m in order:
        cm = program.modules[m]
        f = cm.dag.nodes[cm.dag.root].field
        for i in range(air_chunk_count(cm)):
            by.setdefault(m, []).append((quotient_chunk_name(m, i), f))
    add_tree("air", -1, by)
    lay = Layout(order, trees, tree_of_round)
    _fill_schedule(program, lay)
    return lay


def _fill_schedule(program, lay):
    cfg = E.LeafConfig(wo_challenges=True, wo_verifier=True)
    shifts = {}
    for mn in sorted(program.modules):
        m = program.modules[mn]
        for leaf in E.leaves_full(m.relation, cfg):
            s = ((leaf.shift % m.N) + m.N) % m.N if leaf.shift else 0
            shifts.setdefault(leaf.name, set()).add(s)
    for t in lay.trees:
        for g in t.groups:
            if t.kind == "air":
                g.base_shifts = [[0] for _ in g.base]
                g.ext_shifts = [[0] for _ in g.ext]
            else:
                g.base_shifts = [sorted(shifts.get(n, ())) for n in g.base]
                g.ext_shifts = [sorted(shifts.get(n, ())) for n in g.ext]


def setup_tree_count(lay):
    return sum(1 for t in lay.trees if t.kind == "setup")


PUBI_TAG = 0x50554249


def _string_elements(tag, s):
    b = s.encode()
    return [tag, len(b)] + [int.from_bytes(b[i:i + 3], "little") for i in range(0, len(b), 3)]


def public_entry_field(v):
    """field.Base (0) for an int value, field.Ext (1) for 6 limbs."""
    return 0 if np.isscalar(v) or np.ndim(v) == 0 else 1


def public_entry_words(v):
    """An entry's value elements (appendEntryValueElements,
    public/public_inputs.go:96-101): the base value, or the 6 E6 limbs."""
    if public_entry_field(v) == 0:
        return [int(v) % E.P]
    vals = [int(x) % E.P for x in v]
    if len(vals) != 6:
        raise ValueError("an extension public-input entry has 6 limbs")
    return vals


def public_input_elements(public_inputs):
    """public.Inputs.TranscriptElements (public/public_inputs.go:50-74):
    [PUBI, #inputs], then per input in name order the tagged name and module,
    #entries and (index, field, value elements) per entry in index order
    (field 0: one base value; field 1: six E6 limbs).  public_inputs: name ->
    (module, [(index, value), ...]), value an int or 6 limbs."""
    out = [PUBI_TAG, len(public_inputs)]
    for name in sorted(public_inputs):
        module, entries = public_inputs[name]
        out += _string_elements(PUBI_TAG, name) + _string_elements(PUBI_TAG, module)
        ents = sorted(((int(i), v) for i, v in entries), key=lambda e: e[0])
        out.append(len(ents))
        for i, v in ents:
            out += [i, public_entry_field(v)] + public_entry_words(v)
    return out


def module_sizes_transcript(program):
    """protocol.ModuleSizesTranscript (layout.go:121-129): [SIZE, #mod, N...]."""
    order = sorted(program.modules)
    return [0x53495A45, len(order)] + [program.modules[m].N for m in order]


def challenge_names(program):
    return [challenge_name(r) for r in range(len(program.rounds))]
