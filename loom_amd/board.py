"""Program construction and compilation (host-side mirror of loom's board/ and
arguments/ for the pieces the prove path needs).

Mirrors:
* ``Builder``/``Module`` relation helpers -- board/board.go:52-267,
  board/module.go:60-293;
* rounds: staging, sweep of unstaged committed columns into the last round
  in sorted-name order -- board/rounds.go:72-303;
* ``compile_program`` -- board/program.go:45-136 (fold each module's relations
  with Coin(last round), field inference);
* ``copy_constraint`` -- arguments/permutation.go:37-127 (tuple fold with
  Coin(0), grand product with Coin(1) at round 2);
* ``range_check`` / ``lookup`` -- arguments/lookup.go:34-236 (range module,
  multiplicity step at round 0, cyclic logup columns at round 2 with their
  exposed totals, logup bus), board/logup.go:30-186.

The compiled ``Program`` is what ``protocol.build_layout`` and the plan
serialiser consume.
"""
from dataclasses import dataclass, field

from . import expr as E

ROUND_FOLD = 0
ROUND_LOG_DERIVATIVE = 1
ROUND_RUNNING_SUMS = 2

NO_FS_HOOK = 0
BIND_EXPOSED_VALUES = 1


def challenge_name(r):
    """constants.CanonicalChallengeName (internal/constants/const.go:99-101)."""
    return "challenge@loom_%d" % r


def Coin(r):
    return E.Challenge(challenge_name(r))


def lagrange_name(module, i):
    return "%s.Lagrange_%d" % (module, i)


def lagrange_name_relative(module, i):
    return "%s.Lagrange_%d" % (module, -(i + 1))


# ------------------------------------------------------------------ GenCol
@dataclass
class LagrangeGen:
    i: int


@dataclass
class LagrangeRelativeGen:
    i: int


@dataclass
class PermutationGen:
    S: list
    name: str

    def chunk_name(self, i):
        return "%s_%d" % (self.name, i)


@dataclass
class RangeColumnGen:
    bound: int


# ------------------------------------------------------------------ steps
GRAND_PRODUCT = 1
LOOKUP_MULTIPLICITY = 2
CYCLIC_LOGUP = 3

# board/logup.go:37, arguments/lookup.go:152
MAX_LOGUP_DEGREE = 5
MAX_LOGUP_BATCH = 4


def range_module_name(bound):
    """constants.RangeModuleName (internal/constants/const.go:43-45)."""
    return "range_%d" % bound


def range_col_name(bound):
    """constants.RangeColName (const.go:47-49)."""
    return "range.bound_%d" % bound


def chunk_name(name, i):
    """constants.LogupChunkName / MultiplicityChunkName (const.go:51-57)."""
    return "%s_%d" % (name, i)


def logup_total_name(output):
    """board.LogupTotalName (board/logup.go:64-66)."""
    return output + "_total"


@dataclass
class ProverStep:
    kind: int
    ins: list
    outs: list
    module: str = ""
    # LOOKUP_MULTIPLICITY: dict(n_sources, n_targets, width, has_sel_s,
    # has_sel_t, source_modules); CYCLIC_LOGUP: dict(n_terms)
    ctx: dict = field(default_factory=dict)

    def exposed_outputs(self):
        """ProverStep.ExposedOutputs (board/logup.go:189-199)."""
        if self.kind == CYCLIC_LOGUP:
            return self.outs[1:2]
        return []


@dataclass
class LogupTerm:
    module: str
    E: object
    M: object


def logup_constraint_degree(terms):
    """board.LogupConstraintDegree (board/logup.go:46-56)."""
    sum_e = sum(E.degree(t.E) for t in terms)
    deg = 1 + sum_e
    for t in terms:
        deg = max(deg, E.degree(t.M) + sum_e - E.degree(t.E))
    return deg


@dataclass
class Module:
    name: str
    N: int = 0
    relations: list = field(default_factory=list)
    gencol: list = field(default_factory=list)

    def id_support_name(self, i):
        return "%s.ID_%d_%d" % (self.name, i, self.N)

    def lagrange_col(self, i):
        self.gencol.append(LagrangeGen(i))
        return E.Lagrange(lagrange_name(self.name, i))

    def lagrange_col_relative(self, i):
        self.gencol.append(LagrangeRelativeGen(i))
        return E.Lagrange(lagrange_name_relative(self.name, i))

    def assert_zero(self, rel):
        self.relations.append(rel)

    def assert_equal_at(self, a, b, i):
        self.assert_zero((a - b) * self.lagrange_col(i))

    def assert_equal_relative_at(self, a, b, i):
        self.assert_zero((a - b) * self.lagrange_col_relative(i))

    def assert_zero_at(self, rel, *idx):
        disj = self.lagrange_col(idx[0])
        for k in idx[1:]:
            disj = disj + self.lagrange_col(k)
        self.assert_zero(rel * disj)

    def assert_zero_except_at(self, rel, *idx):
        conj = E.Const(1) - self.lagrange_col(idx[0])
        for k in idx[1:]:
            conj = conj * (E.Const(1) - self.lagrange_col(k))
        self.assert_zero(rel * conj)

    def assert_zero_except_last(self, rel):
        conj = E.Const(1) - self.lagrange_col_relative(0)
        self.assert_zero(conj * rel)


@dataclass
class _BuilderRound:
    steps: list = field(default_factory=list)
    staged: list = field(default_factory=list)
    fshook: int = NO_FS_HOOK


class Builder:
    def __init__(self):
        self.modules = {}
        self.rounds = []
        self.names = 0
        self.logup_bus = []   # lists of total names that must sum to zero

    def fresh_name(self, prefix):
        self.names += 1
        return "%s%d" % (prefix, self.names)

    def add_module(self, m):
        self.modules[m.name] = m

    def _grow(self, r):
        while len(self.rounds) <= r:
            self.rounds.append(_BuilderRound())

    def stage_columns(self, r, *names):
        self._grow(r)
        for n in names:
            if n not in self.rounds[r].staged:
                self.rounds[r].staged.append(n)

    def stage_leaves(self, r, *exprs):
        for e in exprs:
            for leaf in E.leaves_full(e, E.COMMITTED_ONLY):
                self.stage_columns(r, leaf.name)

    def add_step_at(self, r, step):
        self._grow(r)
        self.rounds[r].steps.append(step)

    def set_fs_hook(self, r, hook):
        self._grow(r)
        self.rounds[r].fshook = hook

    def add_lookup_multiplicity_step(self, r, sel_s, sel_t, S, T, output, source_modules, target_modules):
        """board.AddLookupMultiplicityStep (board/board.go:235-249): one
        output column per target, named chunk_name(output, k), in the target's
        module; the step's module is the first target's."""
        ctx = dict(n_sources=len(S), n_targets=len(T), width=len(S[0]), has_sel_s=sel_s is not None,
                   has_sel_t=sel_t is not None, source_modules=list(source_modules),
                   target_modules=list(target_modules))
        target_module = target_modules[0]
        ins = list(sel_s or []) + list(sel_t or [])
        for t in S:
            ins += list(t)
        for t in T:
            ins += list(t)
        outs = [chunk_name(output, i) for i in range(len(T))]
        self.add_step_at(r, ProverStep(LOOKUP_MULTIPLICITY, ins, outs, target_module, ctx))

    def add_cyclic_logup_step(self, r, module, terms, output):
        """board.AddCyclicLogupStep (board/logup.go:68-136): the column L with
        L_i - L_{i-1} + T/N = sum_j M_j/E_j (row 0 reading L_{N-1}); returns
        the name of the exposed total T."""
        if not terms:
            raise ValueError("AddCyclicLogupStep %s: no terms" % output)
        for t in terms:
            if t.module != module:
                raise ValueError("AddCyclicLogupStep %s: term of module %r in a column of %r" % (output, t.module, module))
        d = logup_constraint_degree(terms)
        if d > MAX_LOGUP_DEGREE:
            raise ValueError("AddCyclicLogupStep %s: constraint degree %d exceeds %d" % (output, d, MAX_LOGUP_DEGREE))
        m = self.modules[module]
        total = logup_total_name(output)
        ins = []
        for t in terms:
            ins += [t.E, t.M]
        self.add_step_at(r, ProverStep(CYCLIC_LOGUP, ins, [output, total], module, dict(n_terms=len(terms))))

        def prod_e(skip):
            res = None
            for i, t in enumerate(terms):
                if i == skip:
                    continue
                e = E.clone(t.E)
                res = e if res is None else res * e
            return res

        increment = E.Col(output) - E.Col(output, -1) + E.ExposedAverage(total)
        relation = increment * prod_e(-1)
        for j, t in enumerate(terms):
            num = E.clone(t.M)
            if len(terms) > 1:
                num = num * prod_e(j)
            relation = relation - num
        m.assert_zero(relation)
        return total

    def add_grand_product_step(self, r, module, num, den, output):
        """board.AddGrandProductStep + addGrandProductConstraint (board.go:246-267)."""
        self.add_step_at(r, ProverStep(GRAND_PRODUCT, [num, den], [output], module))
        m = self.modules[module]
        gp = E.Col(output)
        gps = E.Col(output, 1)
        m.assert_zero(gps * den - gp * num)
        m.assert_zero_at(gp - E.Const(1), 0)


# ------------------------------------------------------------------ arguments
def permutation_within_module(b, module, A, B):
    """arguments/permutation.go:37-68."""
    b.stage_leaves(ROUND_FOLD, *A)
    b.stage_leaves(ROUND_FOLD, *B)
    gamma = Coin(ROUND_LOG_DERIVATIVE)
    am = [a - gamma for a in A]
    bm = [x - gamma for x in B]
    gp = b.fresh_name("gp_")
    amul = am[0]
    for x in am[1:]:
        amul = amul * x
    bmul = bm[0]
    for x in bm[1:]:
        bmul = bmul * x
    b.add_grand_product_step(ROUND_RUNNING_SUMS, module, amul, bmul, gp)
    m = b.modules[module]
    m.assert_equal_at(E.Const(1), E.Col(gp), 0)


def permutation_tuple_within_module(b, module, A, B):
    """arguments/permutation.go:82-96."""
    alpha = Coin(ROUND_FOLD)
    fa = [E.fold(alpha, row) for row in A]
    fb = [E.fold(alpha, row) for row in B]
    permutation_within_module(b, module, fa, fb)


def copy_constraint(b, module, A, S):
    """arguments/permutation.go:101-127."""
    m = b.modules[module]
    if m.N * len(A) != len(S.S):
        raise ValueError("m.N*len(A) must be equal to len(S.S)")
    m.gencol.append(S)
    aid = [[A[i], E.Col(m.id_support_name(i))] for i in range(len(A))]
    aperm = [[A[i], E.Col(S.chunk_name(i))] for i in range(len(A))]
    permutation_tuple_within_module(b, module, aid, aperm)


@dataclass
class Column:
    """board.Column: an expression of one module."""
    module: str
    expr: object


def range_check(b, S, bound):
    """arguments.Range (arguments/lookup.go:34-51): S takes values in
    [0, bound) (bound rounded up to a power of two), through a lookup into the
    shared range module of that bound."""
    bound = 1 << max(0, (int(bound) - 1).bit_length())
    name = range_module_name(bound)
    if name not in b.modules:
        rm = Module(name, bound)
        rm.gencol.append(RangeColumnGen(bound))
        b.add_module(rm)
    T = Column(name, E.Col(range_col_name(bound)))
    lookup(b, S, T)


@dataclass
class Table:
    """board.Table: a tuple of expressions of one module."""
    module: str
    cols: list


def _tables_of(cols):
    """columnsToTables (lookup.go:256-262)."""
    return [Table(c.module, [c.expr]) for c in cols]


def _check_widths(S, T):
    """checkWidths (lookup.go:246-254)."""
    w = len(S[0].cols)
    for t in list(S) + list(T):
        if len(t.cols) != w:
            raise ValueError("inconsistent width, expected %d, got %d" % (w, len(t.cols)))


def _fold_tables(tables):
    """foldTables (lookup.go:232-244): each table folded into one column with
    alpha = Coin(RoundFold); width-1 tables as they are."""
    alpha = Coin(ROUND_FOLD)
    return [Column(t.module, t.cols[0] if len(t.cols) == 1 else E.fold(alpha, t.cols)) for t in tables]


def lookup(b, S, T):
    """arguments.Lookup (lookup.go:111-113): S in T for two columns."""
    lookup_union(b, [S], [T])


def lookup_union(b, S, T):
    """arguments.LookupUnion (lookup.go:54-56): {S[0], S[1], ..} in {T[0], T[1], ..}."""
    lookup_union_tuple(b, _tables_of(S), _tables_of(T))


def lookup_tuple(b, S, T):
    """arguments.LookupTuple (lookup.go:116-121): row(S) in row(T), tables of
    equal width."""
    if len(S.cols) != len(T.cols):
        raise ValueError("[LookupTuple] S and T must have equal size, got %d and %d" % (len(S.cols), len(T.cols)))
    lookup_union_tuple(b, [S], [T])


def lookup_union_tuple(b, S, T):
    """arguments.LookupUnionTuple (lookup.go:62-80): multiplicities staged at
    RoundFold on the raw tuples, logups at RoundRunningSums with the tuples
    folded by alpha; every source counts 1, every target -multiplicity."""
    _check_widths(S, T)
    mult = _stage_lookup(b, None, None, S, T)
    one = E.Const(1)
    mult_t = [E.Col(chunk_name(mult, i)) for i in range(len(T))]
    _add_logups(b, _fold_tables(S), _fold_tables(T), [one] * len(S), mult_t)


def c_lookup(b, S, T, sel_s, sel_t):
    """arguments.CLookup (lookup.go:123-125)."""
    c_lookup_union(b, [sel_s], [sel_t], [S], [T])


def c_lookup_tuple(b, S, T, sel_s, sel_t):
    """arguments.CLookupTuple (lookup.go:129-131)."""
    c_lookup_union_tuple(b, [sel_s], [sel_t], [S], [T])


def c_lookup_union(b, sel_s, sel_t, S, T):
    """arguments.CLookupUnion (lookup.go:83-85)."""
    c_lookup_union_tuple(b, sel_s, sel_t, _tables_of(S), _tables_of(T))


def c_lookup_union_tuple(b, sel_s, sel_t, S, T):
    """arguments.CLookupUnionTuple (lookup.go:89-109): {row(S[i]) | selS[i] != 0}
    in {row(T[i]) | selT[i] != 0}; sources count selS, targets
    -multiplicity * selT."""
    if len(sel_s) != len(S):
        raise ValueError("selS has %d chunks, but S has %d chunks" % (len(sel_s), len(S)))
    if len(sel_t) != len(T):
        raise ValueError("selT has %d chunks, but T has %d chunks" % (len(sel_t), len(T)))
    _check_widths(S, T)
    mult = _stage_lookup(b, list(sel_s), list(sel_t), S, T)
    mult_t = [E.Col(chunk_name(mult, i)) * sel_t[i] for i in range(len(T))]
    _add_logups(b, _fold_tables(S), _fold_tables(T), list(sel_s), mult_t)


def _stage_lookup(b, sel_s, sel_t, S, T):
    """arguments.stageLookup (lookup.go:136-154); S, T: lists of Tables."""
    mult = b.fresh_name("Mult_")
    for tables in (S, T):
        for t in tables:
            b.stage_leaves(ROUND_FOLD, *t.cols)
    b.stage_leaves(ROUND_FOLD, *(sel_s or []))
    b.stage_leaves(ROUND_FOLD, *(sel_t or []))
    b.add_lookup_multiplicity_step(ROUND_FOLD, sel_s, sel_t, [list(t.cols) for t in S], [list(t.cols) for t in T],
                                   mult, [t.module for t in S], [t.module for t in T])
    for i in range(len(T)):
        b.stage_columns(ROUND_FOLD, chunk_name(mult, i))
    return mult


def _add_logups(b, S, T, num_s, num_t):
    """arguments.addLogups (lookup.go:163-184): sources contribute
    num_s/(s - gamma), targets -num_t/(t - gamma); totals go on a bus."""
    gamma = Coin(ROUND_LOG_DERIVATIVE)
    minus_one = E.Const(E.P - 1)
    terms = []
    for i, s in enumerate(S):
        terms.append(LogupTerm(s.module, s.expr - gamma, num_s[i]))
    for i, t in enumerate(T):
        terms.append(LogupTerm(t.module, t.expr - gamma, minus_one * num_t[i]))
    totals = add_logup_columns(b, "logup_", terms)
    b.logup_bus.append(totals)


def add_logup_columns(b, prefix, terms, max_degree=None):
    """arguments.AddLogupColumnsWithMaxDegree (lookup.go:197-230): terms
    grouped by module in order of first appearance, batched into columns of at
    most MAX_LOGUP_BATCH terms within the degree limit."""
    order = []
    by_module = {}
    for t in terms:
        if t.module not in by_module:
            order.append(t.module)
            by_module[t.module] = []
        by_module[t.module].append(t)
    totals = []
    for module in order:
        ts = by_module[module]
        limit = (max_degree or {}).get(module, MAX_LOGUP_DEGREE)
        while ts:
            n = 1
            while n < min(MAX_LOGUP_BATCH, len(ts)) and logup_constraint_degree(ts[:n + 1]) <= limit:
                n += 1
            name = chunk_name(b.fresh_name("%s.%s" % (module, prefix)), 0)
            totals.append(b.add_cyclic_logup_step(ROUND_RUNNING_SUMS, module, ts[:n], name))
            ts = ts[n:]
    return totals


# ------------------------------------------------------------------ compile
@dataclass
class ColumnRef:
    name: str
    module: str
    field: int = E.BASE


@dataclass
class ProverRound:
    steps: list
    staged: list
    fshook: int = NO_FS_HOOK


@dataclass
class CompiledModule:
    name: str
    N: int
    relation: object
    dag: object
    degree: int
    gencol: list


@dataclass
class Program:
    modules: dict
    rounds: list
    column_fields: dict
    logup_bus: list = field(default_factory=list)
    # precommitted setup columns (board.Program.SetupColumns, program.go:19):
    # sorted by (module, name) as collectSetupColumns (program.go:138-165)
    setup_columns: list = field(default_factory=list)
    # public-input columns referenced by the relations: (name, module), by name
    public_columns: list = field(default_factory=list)
    # public-input columns declared in the extension field (stored with six
    # limbs; their entries may be E6 values)
    public_ext: list = field(default_factory=list)


def _column_modules(modules):
    owner = {}
    for mn, m in modules.items():
        for rel in m.relations:
            for leaf in E.leaves_full(rel, E.COMMITTED_ONLY):
                owner[leaf.name] = mn
    return owner


def compile_program(b):
    owner = _column_modules(b.modules)
    declared = b.rounds if b.rounds else [_BuilderRound()]
    rounds = []
    staged = {}
    for r, br in enumerate(declared):
        hook = br.fshook
        if any(st.exposed_outputs() for st in br.steps):
            if hook not in (NO_FS_HOOK, BIND_EXPOSED_VALUES):
                raise ValueError("round %d exposes values but has FS hook %d" % (r, hook))
            hook = BIND_EXPOSED_VALUES
        rd = ProverRound(list(br.steps), [], hook)
        for name in br.staged:
            if name in staged:
                raise ValueError("column %r staged twice" % name)
            if name not in owner:
                raise ValueError("column %r staged but not referenced" % name)
            staged[name] = r
            rd.staged.append(ColumnRef(name, owner[name]))
        rounds.append(rd)
    last = len(rounds) - 1
    for name in sorted(owner):
        if name in staged:
            continue
        staged[name] = last
        rounds[last].staged.append(ColumnRef(name, owner[name]))

    # field inference (program.go:91-136): monotone joins.
    fields = {}

    def setf(n, f):
        fields[n] = E.join(fields.get(n, E.BASE), f)

    cfg = E.LeafConfig(wo_challenges=True)
    for m in b.modules.values():
        for rel in m.relations:
            for leaf in E.leaves_full(rel, cfg):
                setf(leaf.name, E.field_of(leaf, fields))
    for r, rd in enumerate(rounds):
        setf(challenge_name(r), E.EXT)
        for st in rd.steps:
            # inferStepOutputFields (board/program.go:181-222)
            f = E.BASE
            if st.kind != LOOKUP_MULTIPLICITY:
                for e in st.ins:
                    f = E.join(f, E.field_of(e, fields))
            for o in st.outs:
                setf(o, f)
    for rd in rounds:
        for c in rd.staged:
            c.field = E.join(c.field, fields.get(c.name, E.BASE))

    # setup columns: every setup leaf of a module's relations
    # (collectSetupColumns, board/program.go:138-165); public-input leaves
    setup = {}
    public = {}
    public_ext = set()  # public inputs declared in the extension field (expr.PublicInputExt)
    for mn, m in b.modules.items():
        for rel in m.relations:
            for leaf in E.leaves_full(rel):
                if leaf.type == E.SETUP:
                    setup[mn + "\x00" + leaf.name] = ColumnRef(leaf.name, mn, E.join(
                        setup.get(mn + "\x00" + leaf.name, ColumnRef("", "", E.BASE)).field, leaf.field_kind()))
                elif leaf.type == E.VERIFIER and leaf.hook == E.PUBLIC_INPUT_HOOK:
                    if public.get(leaf.name, mn) != mn:
                        raise ValueError("public input %r used by two modules" % leaf.name)
                    public[leaf.name] = mn
                    if leaf.field == E.EXT:
                        public_ext.add(leaf.name)
    setup_columns = [setup[k] for k in sorted(setup)]
    for c in setup_columns:
        setf(c.name, c.field)
        c.field = E.join(c.field, fields.get(c.name, E.BASE))

    fold_ch = Coin(len(rounds) - 1)
    mods = {}
    for k, m in b.modules.items():
        rel = E.fold(fold_ch, m.relations)
        dag = E.to_dag(rel, fields)
        mods[k] = CompiledModule(m.name, m.N, rel, dag, E.degree(rel), list(m.gencol))
    return Program(mods, rounds, fields, [list(x) for x in b.logup_bus], setup_columns,
                   sorted(public.items()), sorted(public_ext))
