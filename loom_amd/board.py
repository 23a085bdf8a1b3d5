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
  Coin(0), grand product with Coin(1) at round 2).

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


@dataclass
class ProverStep:
    kind: int
    ins: list
    outs: list
    module: str = ""


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
        rd = ProverRound(list(br.steps), [], br.fshook)
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
            f = E.BASE
            for e in st.ins:
                f = E.join(f, E.field_of(e, fields))
            for o in st.outs:
                setf(o, f)
    for rd in rounds:
        for c in rd.staged:
            c.field = E.join(c.field, fields.get(c.name, E.BASE))

    fold_ch = Coin(len(rounds) - 1)
    mods = {}
    for k, m in b.modules.items():
        rel = E.fold(fold_ch, m.relations)
        dag = E.to_dag(rel, fields)
        mods[k] = CompiledModule(m.name, m.N, rel, dag, E.degree(rel), list(m.gencol))
    return Program(mods, rounds, fields)
