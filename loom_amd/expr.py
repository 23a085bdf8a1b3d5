"""Symbolic expressions over trace columns (host-side mirror of loom's expr/).

Only what the prove path consumes: leaf kinds (committed column with a row
shift, challenge, constant, verifier column with a hook), binary Add/Sub/Mul,
Pow, ``fold`` (Horner, expr/utils.go:21-33), left-to-right leaf order
(expr/ast.go:591-640), degree (ast.go:303-315) and field inference
(ast.go:256-293).  Values are canonical integers mod p.
"""
from dataclasses import dataclass, field as dc_field

P = 2**31 - 2**24 + 1

BASE, EXT = 0, 1

# leaf types (expr/ast.go:28-38)
COMMITTED, CHALLENGE, CONSTANT, SETUP, VERIFIER = range(5)
# verifier hooks (expr/ast.go:44-58)
NO_HOOK, LAGRANGE_HOOK, PUBLIC_INPUT_HOOK, EXPOSED_VALUE_HOOK, EXPOSED_AVERAGE_HOOK = range(5)

NEG_INF = -(1 << 30)


def join(a, b):
    return EXT if (a == EXT or b == EXT) else BASE


class Expr:
    def __add__(self, o):
        return Add(self, _lift(o))

    def __sub__(self, o):
        return Sub(self, _lift(o))

    def __mul__(self, o):
        return Mul(self, _lift(o))

    def __radd__(self, o):
        return Add(_lift(o), self)

    def __rsub__(self, o):
        return Sub(_lift(o), self)

    def __rmul__(self, o):
        return Mul(_lift(o), self)

    def pow(self, n):
        return Pow(self, int(n))


def _lift(o):
    if isinstance(o, Expr):
        return o
    return Const(int(o))


@dataclass(eq=False)
class Leaf(Expr):
    type: int
    name: str = ""
    shift: int = 0
    field: int = BASE
    hook: int = NO_HOOK
    value: int = 0

    def field_kind(self):
        if self.type in (CHALLENGE, VERIFIER):
            return EXT
        return self.field

    def key(self):
        """leaf.String() (ast.go:295-301); constants print their value."""
        if self.type == CONSTANT:
            return str(self.value)
        if self.shift != 0:
            return "%s_shift_%d" % (self.name, self.shift)
        return self.name

    def degree(self):
        if self.type == CONSTANT:
            return NEG_INF if self.value % P == 0 else 0
        if self.type == CHALLENGE:
            return 0
        return 1


@dataclass(eq=False)
class Add(Expr):
    left: Expr
    right: Expr


@dataclass(eq=False)
class Sub(Expr):
    left: Expr
    right: Expr


@dataclass(eq=False)
class Mul(Expr):
    left: Expr
    right: Expr


@dataclass(eq=False)
class Pow(Expr):
    base: Expr
    exp: int


def Col(name, shift=0):
    return Leaf(COMMITTED, name, shift, BASE)


def ExtCol(name, shift=0):
    return Leaf(COMMITTED, name, shift, EXT)


def Setup(name, shift=0):
    """A precommitted setup column (expr.Setup, ast.go:211-213)."""
    return Leaf(SETUP, name, shift, BASE)


def ExtSetup(name, shift=0):
    """A precommitted setup column of the extension field (expr.ExtSetup,
    ast.go:215-217): its values are E6."""
    return Leaf(SETUP, name, shift, EXT)


def PublicInput(name):
    """A public-input column (expr.PublicInput, ast.go:230-232): a verifier
    column the prover materialises from the statement's entries."""
    return Leaf(VERIFIER, name, 0, BASE, PUBLIC_INPUT_HOOK)


def PublicInputExt(name):
    """A public-input column declared in the extension field
    (expr.PublicInputExt, ast.go:234-239): its entries may be E6 values
    (public.Entry.SetExt); base entries are lifted."""
    return Leaf(VERIFIER, name, 0, EXT, PUBLIC_INPUT_HOOK)


def Challenge(name):
    return Leaf(CHALLENGE, name, 0, EXT)


def Const(value):
    return Leaf(CONSTANT, "", 0, BASE, NO_HOOK, int(value) % P)


def Lagrange(name):
    return Leaf(VERIFIER, name, 0, BASE, LAGRANGE_HOOK)


def Exposed(name):
    return Leaf(VERIFIER, name, 0, BASE, EXPOSED_VALUE_HOOK)


def ExposedAverage(name):
    return Leaf(VERIFIER, name, 0, EXT, EXPOSED_AVERAGE_HOOK)


def clone(e):
    if isinstance(e, Leaf):
        return Leaf(e.type, e.name, e.shift, e.field, e.hook, e.value)
    if isinstance(e, Pow):
        return Pow(clone(e.base), e.exp)
    return type(e)(clone(e.left), clone(e.right))


def fold(v, cs):
    """Horner fold C[0] + v*(C[1] + v*(C[2] + ...)) (expr/utils.go:21-33)."""
    if not cs:
        return Const(0)
    res = cs[-1]
    for c in reversed(cs[:-1]):
        res = Add(c, Mul(clone(v), res))
    return res


@dataclass
class LeafConfig:
    wo_committed: bool = False
    wo_setup: bool = False
    wo_challenges: bool = False
    wo_verifier: bool = False
    wo_rotated: bool = False


COMMITTED_ONLY = LeafConfig(wo_setup=True, wo_challenges=True, wo_verifier=True)


def leaves_full(e, cfg=LeafConfig()):
    """Leaves in left-to-right order, constants excluded (ast.go:591-626)."""
    out = []

    def walk(x):
        if isinstance(x, Leaf):
            if x.type == CONSTANT:
                return
            if cfg.wo_rotated and x.shift != 0:
                return
            if x.type == COMMITTED and cfg.wo_committed:
                return
            if x.type == SETUP and cfg.wo_setup:
                return
            if x.type == CHALLENGE and cfg.wo_challenges:
                return
            if x.type == VERIFIER and cfg.wo_verifier:
                return
            out.append(x)
        elif isinstance(x, Pow):
            walk(x.base)
        else:
            walk(x.left)
            walk(x.right)

    walk(e)
    return out


def degree(e):
    if isinstance(e, Leaf):
        return e.degree()
    if isinstance(e, Pow):
        return degree(e.base) * e.exp
    if isinstance(e, Mul):
        return degree(e.left) + degree(e.right)
    return max(degree(e.left), degree(e.right))


def field_of(e, column_fields=None):
    if isinstance(e, Leaf):
        f = e.field_kind()
        if e.type in (COMMITTED, SETUP) and column_fields is not None:
            f = join(f, column_fields.get(e.name, BASE))
        return f
    if isinstance(e, Pow):
        return field_of(e.base, column_fields)
    return join(field_of(e.left, column_fields), field_of(e.right, column_fields))


# ------------------------------------------------------------------ DAG
# Hash-consed DAG in topological order (internal/dag/dag.go:113-312): the
# node list is what the prover evaluates on the device.  Commutative nodes are
# keyed on sorted child ids so (a+b) and (b+a) share one node.

OP_LEAF, OP_ADD, OP_SUB, OP_MUL, OP_POW = range(5)


@dataclass
class DagNode:
    op: int
    field: int
    children: tuple = ()
    leaf: Leaf = None
    exp: int = 0


@dataclass
class Dag:
    nodes: list = dc_field(default_factory=list)
    root: int = -1


def to_dag(e, column_fields=None):
    dag = Dag()
    cache = {}

    def leaf_field(l):
        return field_of(l, column_fields)

    def build(x):
        if isinstance(x, Leaf):
            f = leaf_field(x)
            key = ("L", x.type, x.shift, x.name, f, x.hook, x.value)
            if key not in cache:
                cache[key] = len(dag.nodes)
                dag.nodes.append(DagNode(OP_LEAF, f, (), x))
            return cache[key]
        if isinstance(x, Pow):
            c = build(x.base)
            key = ("P", c, x.exp)
            if key not in cache:
                cache[key] = len(dag.nodes)
                dag.nodes.append(DagNode(OP_POW, dag.nodes[c].field, (c,), None, x.exp))
            return cache[key]
        a = build(x.left)
        b = build(x.right)
        op = OP_ADD if isinstance(x, Add) else OP_SUB if isinstance(x, Sub) else OP_MUL
        ch = (min(a, b), max(a, b)) if op != OP_SUB else (a, b)
        key = (op,) + ch
        if key not in cache:
            cache[key] = len(dag.nodes)
            dag.nodes.append(DagNode(op, join(dag.nodes[a].field, dag.nodes[b].field), ch))
        return cache[key]

    dag.root = build(e)
    return dag
