"""Serialise a compiled program + layout into the plan word stream the C-ABI
(``lg_plan_create``) consumes.

The plan is everything the prover needs that the reference computes on its
Go side before the hot path runs: module sizes, the column table (where each
column's values come from), the rounds with their steps and trace trees, the
AIR tree, the opening shifts of every committed polynomial
(protocol.BuildLayout / BuildCanonicalSchedule) and every module's folded
vanishing relation as a topologically ordered DAG (board.Compile).  A Go
caller would emit the same words from its compiled ``board.Program``.

Word stream (u32, little endian):
  MAGIC, VERSION, num_queries
  n_modules, then per module (sorted names): N, degree, n_chunks
  n_cols, then per column: field, module, src, a0, a1
  n_inputs, then per input: column id      (caller passes inputs in this order)
  n_perms, then per permutation: module, len   (the permutation itself is
            per-instance data, passed beside the trace columns)
  n_rounds, then per round: challenge col, trace tree (or NONE), fs_hook,
            n_steps, per step: kind, module, n_outs, out cols...,
            n_params, params..., n_ins, dag...
            (params: lookup multiplicity -> n_sources, n_targets, width,
            has_sel_s, has_sel_t, module per source, module per target;
            cyclic logup -> n_terms; grand product -> none)
  n_exposed, then per exposed value (sorted by name): column, name
            (byte length, bytes packed 4 per word, little endian)
  n_public, then per public-input column (sorted by name): column, module,
            name, module name (strings as above), n_entries, entry rows...,
            entry fields... (0 base, 1 extension; the caller passes each
            instance's entry values after its input columns: one word per
            base entry, six per extension entry, in row order)
  n_trees, then per tree: kind (0 trace, 1 AIR, 2 setup), n_groups, per group: module, N,
            nb, base cols..., ne, ext cols..., shift lists (base polys then
            ext polys; each: count, shifts...)
  per module: relation dag
  n_buses, then per logup bus: n_members, exposed-value indices (into the
            sorted exposed list; the totals that must sum to zero)
  dag := n_nodes, root, then 4 words per node: op, field, a, b
"""
import numpy as np

from . import expr as E
from . import board as B
from . import protocol as PR

MAGIC = 0x31504C4C
VERSION = 5
NONE = 0xFFFFFFFF

SRC_INPUT, SRC_LAGRANGE, SRC_ID, SRC_SIGMA, SRC_STEP, SRC_CHALLENGE, SRC_AIR, SRC_RANGE, SRC_SETUP, SRC_PUBLIC = range(10)
OP_COL, OP_CONST, OP_CHAL, OP_ADD, OP_SUB, OP_MUL, OP_POW = range(7)


class Plan:
    def __init__(self, program, num_queries=PR.NUM_QUERIES_DEFAULT, public_rows=None):
        """public_rows: public-input column name -> its entry rows (the
        statement's public.Inputs fix which rows are public; their values are
        per-instance inputs)."""
        self.program = program
        # entry rows of each public input, each a row (field.Base entry) or a
        # (row, field) pair (field 1: an extension entry)
        self.public_rows = {k: sorted((int(e[0]), int(e[1])) if isinstance(e, (tuple, list)) else (int(e), 0)
                                      for e in v) for k, v in (public_rows or {}).items()}
        self.layout = PR.build_layout(program)
        self.num_queries = num_queries
        self.modules = sorted(program.modules)
        self.mod_idx = {m: i for i, m in enumerate(self.modules)}
        self.cols = []          # (name, field, module, src, a0, a1)
        self.col_idx = {}
        self.perms = []
        self.inputs = []
        self.setup = []   # setup column names, in lg_plan_set_setup order (the setup tree's)
        self.public = []  # public-input column names, by name
        self._collect()

    def _add(self, name, field, module, src, a0=0, a1=0):
        if name in self.col_idx:
            return self.col_idx[name]
        self.col_idx[name] = len(self.cols)
        self.cols.append((name, field, module, src, a0, a1))
        return self.col_idx[name]

    def _collect(self):
        prog = self.program
        for t in self.layout.trees:
            if t.kind != "setup":
                continue
            for g in t.groups:
                for name in g.base + g.ext:
                    f = prog.column_fields.get(name, E.BASE)
                    self._add(name, f, self.mod_idx[g.module], SRC_SETUP, len(self.setup))
                    self.setup.append(name)
        for name, mn in getattr(prog, "public_columns", []):
            if name not in self.public_rows:
                raise KeyError("public input %r has no entry rows" % name)
            # the column is materialised in the extension field when it is
            # declared there or any entry is an extension value (the join of
            # the entries' fields, prover.go:241-268); base entries are lifted
            ext = name in getattr(prog, "public_ext", []) or any(fe for _, fe in self.public_rows[name])
            f = E.EXT if ext else E.BASE
            self._add(name, f, self.mod_idx[mn], SRC_PUBLIC, len(self.public))
            self.public.append(name)
        gen = {}
        for mn in self.modules:
            m = prog.modules[mn]
            mi = self.mod_idx[mn]
            for g in m.gencol:
                if isinstance(g, B.LagrangeGen):
                    gen[B.lagrange_name(mn, g.i)] = (mi, SRC_LAGRANGE, g.i, 0)
                elif isinstance(g, B.LagrangeRelativeGen):
                    gen[B.lagrange_name_relative(mn, g.i)] = (mi, SRC_LAGRANGE, m.N - 1 - g.i, 0)
                elif isinstance(g, B.RangeColumnGen):
                    gen[B.range_col_name(g.bound)] = (mi, SRC_RANGE, g.bound, 0)
                elif isinstance(g, B.PermutationGen):
                    pi = len(self.perms)
                    self.perms.append((mi, np.asarray(g.S, dtype=np.int64)))
                    k = len(g.S) // m.N
                    for i in range(k):
                        gen[m_id_name(mn, i, m.N)] = (mi, SRC_ID, i, 0)
                        gen[g.chunk_name(i)] = (mi, SRC_SIGMA, i, pi)
                else:
                    raise NotImplementedError(type(g))
        step_out = {}
        for r, rd in enumerate(prog.rounds):
            self._add(B.challenge_name(r), E.EXT, 0, SRC_CHALLENGE, r)
            for s, st in enumerate(rd.steps):
                # a multiplicity column lives in its target's module
                mods = (st.ctx.get("target_modules") if st.kind == B.LOOKUP_MULTIPLICITY else None) or \
                    [st.module] * len(st.outs)
                for o, om in zip(st.outs, mods):
                    step_out[o] = (self.mod_idx[om], r, s)
        fields = prog.column_fields

        def reg(name):
            if name in self.col_idx:
                return self.col_idx[name]
            f = fields.get(name, E.BASE)
            if name in gen:
                mi, src, a0, a1 = gen[name]
                return self._add(name, E.BASE, mi, src, a0, a1)
            if name in step_out:
                mi, r, s = step_out[name]
                return self._add(name, f, mi, SRC_STEP, r, s)
            raise KeyError(name)

        names = set()
        for t in self.layout.trees:
            if t.kind == "air":
                continue
            for g in t.groups:
                names.update(g.base)
                names.update(g.ext)
        for mn in self.modules:
            for leaf in E.leaves_full(prog.modules[mn].relation):
                if leaf.type in (E.COMMITTED, E.VERIFIER):
                    names.add(leaf.name)
        for rd in prog.rounds:
            for st in rd.steps:
                for e in st.ins:
                    for leaf in E.leaves_full(e):
                        if leaf.type in (E.COMMITTED, E.VERIFIER):
                            names.add(leaf.name)
        for name in sorted(names):
            if name in gen or name in step_out:
                reg(name)
        owner = {}
        for mn in self.modules:
            for leaf in E.leaves_full(prog.modules[mn].relation, E.COMMITTED_ONLY):
                owner.setdefault(leaf.name, mn)
        for name in sorted(names):
            if name not in self.col_idx:
                if name not in owner:
                    raise KeyError("column %r has no owning module" % name)
                ci = self._add(name, fields.get(name, E.BASE), self.mod_idx[owner[name]], SRC_INPUT,
                               len(self.inputs))
                self.inputs.append(name)
        for mn in self.modules:
            m = prog.modules[mn]
            for i in range(PR.air_chunk_count(m)):
                self._add(PR.quotient_chunk_name(mn, i), m.dag.nodes[m.dag.root].field,
                          self.mod_idx[mn], SRC_AIR, i)

    def input_names(self):
        return list(self.inputs)

    def perm_arrays(self):
        """The permutations (plan order) as uint32 arrays."""
        return [np.ascontiguousarray(S, dtype=np.uint32) for _, S in self.perms]

    def _dag_words(self, e, n, out):
        dag = E.to_dag(e, self.program.column_fields)
        out += [len(dag.nodes), dag.root]
        for nd in dag.nodes:
            if nd.op == E.OP_LEAF:
                lf = nd.leaf
                if lf.type == E.CONSTANT:
                    out += [OP_CONST, nd.field, lf.value % E.P, 0]
                elif lf.type == E.CHALLENGE:
                    out += [OP_CHAL, nd.field, self.col_idx[lf.name], 0]
                elif lf.type == E.VERIFIER and lf.hook == E.EXPOSED_AVERAGE_HOOK:
                    # the constant column total/N: a per-instance E6 value, like a challenge
                    out += [OP_CHAL, nd.field, self.col_idx[lf.name], 0]
                else:
                    s = ((lf.shift % n) + n) % n
                    out += [OP_COL, nd.field, self.col_idx[lf.name], s]
            elif nd.op == E.OP_POW:
                out += [OP_POW, nd.field, nd.children[0], nd.exp]
            else:
                op = {E.OP_ADD: OP_ADD, E.OP_SUB: OP_SUB, E.OP_MUL: OP_MUL}[nd.op]
                out += [op, nd.field, nd.children[0], nd.children[1]]

    def words(self):
        prog, lay = self.program, self.layout
        w = [MAGIC, VERSION, self.num_queries]
        w.append(len(self.modules))
        for mn in self.modules:
            m = prog.modules[mn]
            w += [m.N, m.degree, PR.air_chunk_count(m)]
        w.append(len(self.cols))
        for name, f, mi, src, a0, a1 in self.cols:
            w += [f, mi, src, a0, a1]
        w.append(len(self.inputs))
        w += [self.col_idx[n] for n in self.inputs]
        w.append(len(self.perms))
        for mi, S in self.perms:
            w += [mi, len(S)]
        w.append(len(prog.rounds))
        for r, rd in enumerate(prog.rounds):
            w += [self.col_idx[B.challenge_name(r)], lay.trace_tree_of_round.get(r, NONE), rd.fshook,
                  len(rd.steps)]
            for st in rd.steps:
                n = prog.modules[st.module].N
                w += [st.kind, self.mod_idx[st.module], len(st.outs)] + [self.col_idx[o] for o in st.outs]
                sizes = [n] * len(st.ins)
                if st.kind == B.LOOKUP_MULTIPLICITY:
                    c = st.ctx
                    tmods = c.get("target_modules") or [st.module] * c["n_targets"]
                    params = [c["n_sources"], c["n_targets"], c["width"], int(c["has_sel_s"]), int(c["has_sel_t"])]
                    params += [self.mod_idx[x] for x in c["source_modules"]]
                    params += [self.mod_idx[x] for x in tmods]
                    ns = [prog.modules[x].N for x in c["source_modules"]]
                    nt = [prog.modules[x].N for x in tmods]
                    sizes = (ns if c["has_sel_s"] else []) + (nt if c["has_sel_t"] else [])
                    sizes += [x for x in ns for _ in range(c["width"])] + [x for x in nt for _ in range(c["width"])]
                elif st.kind == B.CYCLIC_LOGUP:
                    params = [st.ctx["n_terms"]]
                else:
                    params = []
                w += [len(params)] + params + [len(st.ins)]
                for e, sz in zip(st.ins, sizes):
                    self._dag_words(e, sz, w)
        exposed = sorted(o for rd in prog.rounds for st in rd.steps for o in st.exposed_outputs())

        def string(x):
            b = x.encode()
            w.append(len(b))
            for i in range(0, len(b), 4):
                w.append(int.from_bytes(b[i:i + 4].ljust(4, b"\0"), "little"))

        w.append(len(exposed))
        for name in exposed:
            w.append(self.col_idx[name])
            string(name)
        w.append(len(self.public))
        for name in self.public:
            mn = dict(prog.public_columns)[name]
            w += [self.col_idx[name], self.mod_idx[mn]]
            string(name)
            string(mn)
            ents = self.public_rows[name]
            w.append(len(ents))
            w += [r for r, _ in ents] + [f for _, f in ents]
        w.append(len(lay.trees))
        for t in lay.trees:
            w += [{"trace": 0, "air": 1, "setup": 2}[t.kind], len(t.groups)]
            for g in t.groups:
                w += [self.mod_idx[g.module], g.N, len(g.base)] + [self.col_idx[c] for c in g.base]
                w += [len(g.ext)] + [self.col_idx[c] for c in g.ext]
                for ss in g.base_shifts + g.ext_shifts:
                    w += [len(ss)] + list(ss)
        for mn in self.modules:
            m = prog.modules[mn]
            self._dag_words(m.relation, m.N, w)
        w.append(len(prog.logup_bus))
        for bus in prog.logup_bus:
            w.append(len(bus))
            w += [exposed.index(nm) for nm in bus]
        return np.array(w, dtype=np.uint32)


def m_id_name(module, i, n):
    return "%s.ID_%d_%d" % (module, i, n)
