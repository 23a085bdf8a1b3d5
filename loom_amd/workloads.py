"""Seeded synthetic statements for the prove-path workloads.

``plonk_instance(log_n, seed)``: one plonk-like module of N = 2^log_n rows over
KoalaBear, built like the reference's gnark PLONK modules
(integration_test/gnark_plonk/plonk_gnark_utils.go:221-232, bench/main.go:93-104):
gate qL*L + qR*R + qM*L*R + qO*O + qK committed as columns (no setup), and a
copy constraint over (L, R, O) (arguments.CopyConstraint).

Generator (numpy PCG64 seeded with `seed`):
* the copy permutation first: 10% of the 3N wire cells, in random order, are
  cut into cycles of length 2-4; sigma maps each cycle cell to the next one,
  every other cell to itself;
* L, R, O uniform in [0, p); each cycle then copies its first cell's value to
  its other cells, so tied cells are equal;
* per row a gate kind: 40% add (qL=qR=1, qO=-1, qM=0), 40% mul (qM=1, qO=-1,
  qL=qR=0), 20% constant (qL=1, others 0);
* qK = -(qL*L + qR*R + qM*L*R + qO*O), so every row satisfies the gate.
"""
import numpy as np

from . import board as B
from . import expr as E

P = E.P


def _names(i):
    s = lambda x: "%s_%d" % (x, i)
    return dict(module=s("plonk"), L=s("L"), R=s("R"), O=s("O"), QL=s("QL"), QR=s("QR"),
                QM=s("QM"), QO=s("QO"), QK=s("QK"), perm=s("plonk.S"))


def _plonk_gate(m, nm):
    qll = E.Col(nm["QL"]) * E.Col(nm["L"])
    qrr = E.Col(nm["QR"]) * E.Col(nm["R"])
    qmlr = E.Col(nm["QM"]) * E.Col(nm["L"]) * E.Col(nm["R"])
    qoo = E.Col(nm["QO"]) * E.Col(nm["O"])
    qk = E.Col(nm["QK"])
    m.assert_zero(qll + qrr + qmlr + qoo + qk)


def plonk_program(log_n, idx=0):
    """The builder of one plonk-like module (same for every seed)."""
    nm = _names(idx)
    b = B.Builder()
    m = B.Module(nm["module"], 1 << log_n)
    _plonk_gate(m, nm)
    b.add_module(m)
    return b, nm


def _plonk_values(n, rng, l_bound=P):
    """sigma, wires and selectors of one module, drawn from rng: L uniform in
    [0, l_bound), R and O uniform in [0, p).  A copy cycle that holds an L
    cell carries the value of its first L cell (so a range-checked L stays in
    its table), any other cycle the value of its first cell."""
    cells = 3 * n
    sigma = np.arange(cells, dtype=np.int64)
    tied = rng.permutation(cells)[: cells // 10]
    cycles = []
    pos = 0
    while pos + 2 <= len(tied):
        ln = int(rng.integers(2, 5))
        ln = min(ln, len(tied) - pos)
        cyc = tied[pos:pos + ln]
        cycles.append(cyc)
        for k in range(ln):
            sigma[cyc[k]] = cyc[(k + 1) % ln]
        pos += ln
    w = rng.integers(0, P, size=cells, dtype=np.uint64)
    if l_bound != P:
        w[:n] = rng.integers(0, l_bound, size=n, dtype=np.uint64)
    for cyc in cycles:
        src = cyc[0]
        if l_bound != P:
            in_l = cyc[cyc < n]
            if len(in_l):
                src = in_l[0]
        w[cyc] = w[src]
    L, R, O = w[:n], w[n:2 * n], w[2 * n:]
    kind = rng.choice(3, size=n, p=[0.4, 0.4, 0.2])
    one, mone = np.uint64(1), np.uint64(P - 1)
    zero = np.uint64(0)
    ql = np.where(kind == 0, one, np.where(kind == 2, one, zero)).astype(np.uint64)
    qr = np.where(kind == 0, one, zero).astype(np.uint64)
    qm = np.where(kind == 1, one, zero).astype(np.uint64)
    qo = np.where(kind == 2, zero, mone).astype(np.uint64)
    pp = np.uint64(P)
    acc = (ql * L % pp + qr * R % pp + qm * (L * R % pp) % pp + qo * O % pp) % pp
    qk = (pp - acc) % pp
    return sigma, dict(L=L.copy(), R=R.copy(), O=O.copy(), QL=ql, QR=qr, QM=qm, QO=qo, QK=qk)


def _add_plonk_module(b, log_n, idx, rng, range_bound=None):
    n = 1 << log_n
    sigma, vals = _plonk_values(n, rng, P if range_bound is None else range_bound)
    nm = _names(idx)
    m = B.Module(nm["module"], n)
    _plonk_gate(m, nm)
    b.add_module(m)
    S = B.PermutationGen(sigma, nm["perm"])
    B.copy_constraint(b, nm["module"], [E.Col(nm["L"]), E.Col(nm["R"]), E.Col(nm["O"])], S)
    if range_bound is not None:
        B.range_check(b, B.Column(nm["module"], E.Col(nm["L"])), range_bound)
    return {nm[k]: v for k, v in vals.items()}


def plonk_instance(log_n, seed, idx=0):
    """Returns (program, trace) with trace: name -> canonical uint64 array."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = B.Builder()
    trace = _add_plonk_module(b, log_n, idx, rng)
    return B.compile_program(b), trace


def uneven_instance(log_ns, seed):
    """One program with len(log_ns) plonk-like modules plonk_0, plonk_1, ...
    of sizes 2^log_ns[i] (the uneven-batch shape: modules of different sizes
    proved together); module i draws from PCG64([seed, i])."""
    b = B.Builder()
    trace = {}
    for i, ln in enumerate(log_ns):
        rng = np.random.Generator(np.random.PCG64([seed, i]))
        trace.update(_add_plonk_module(b, ln, i, rng))
    return B.compile_program(b), trace


def uneven_range_instance(log_ns, seed, bound=256):
    """The uneven batch with lookups: every module as in uneven_instance, its
    L wire drawn from [0, bound) (R and O uniform in [0, p)), plus
    arguments.Range(L, bound) -- a logup lookup of its L column into the
    shared range_<bound> module."""
    b = B.Builder()
    trace = {}
    for i, ln in enumerate(log_ns):
        rng = np.random.Generator(np.random.PCG64([seed, i]))
        trace.update(_add_plonk_module(b, ln, i, rng, bound))
    return B.compile_program(b), trace


def uneven_sizes(count, seed, lo=10, hi=20):
    """log2 sizes drawn uniformly from [lo, hi] (PCG64([seed, 1 << 20]))."""
    rng = np.random.Generator(np.random.PCG64([seed, 1 << 20]))
    return [int(x) for x in rng.integers(lo, hi + 1, size=count)]


def _fibonacci_table(n):
    """F_0 .. F_n mod p by block doubling: with F_0 .. F_{m+1} known,
    F_{m+1+j} = F_{m+1} F_{j+1} + F_m F_j for j = 0 .. m."""
    pp = np.uint64(P)
    F = np.zeros(2 * (n + 2), dtype=np.uint64)
    F[1] = F[2] = 1
    m = 1
    while m + 1 < n:
        j = np.arange(m + 1, dtype=np.int64)
        F[m + 1 + j] = (F[m + 1] * F[j + 1] % pp + F[m] * F[j] % pp) % pp
        m *= 2
    return F[:n + 1].copy()


def wide_instance(log_n, seed, n_fib=32, n_cubic=32):
    """The wide single-module trace: N = 2^log_n rows, 2*n_fib + n_cubic base
    columns.  Pair k: u_{i+1} = v_i, v_{i+1} = u_i + v_i; chain k:
    w_{i+1} = w_i^3 + c_k (c_k uniform in [0, p)); seeds uniform in [0, p).
    Every transition is asserted except at the last row
    (AssertZeroExceptAt(N-1)), so the folded relation has degree 4."""
    n = 1 << log_n
    rng = np.random.Generator(np.random.PCG64(seed))
    pp = np.uint64(P)
    u0 = rng.integers(0, P, size=n_fib, dtype=np.uint64)
    v0 = rng.integers(0, P, size=n_fib, dtype=np.uint64)
    w0 = rng.integers(0, P, size=n_cubic, dtype=np.uint64)
    cs = rng.integers(0, P, size=n_cubic, dtype=np.uint64)
    F = _fibonacci_table(n)
    fprev = np.concatenate(([np.uint64(1)], F[:n - 1]))  # F_{i-1}, with F_{-1} = 1
    fcur = F[:n]
    fnext = F[1:n + 1]
    trace = {}
    for k in range(n_fib):
        # (u_i, v_i) = M^i (u0, v0), M^i = [[F_{i-1}, F_i], [F_i, F_{i+1}]]
        trace["U_%d" % k] = (fprev * u0[k] % pp + fcur * v0[k] % pp) % pp
        trace["V_%d" % k] = (fcur * u0[k] % pp + fnext * v0[k] % pp) % pp
    W = np.empty((n, n_cubic), dtype=np.uint64)
    w = w0.copy()
    for i in range(n):
        W[i] = w
        w = (w * w % pp * w + cs) % pp
    for k in range(n_cubic):
        trace["W_%d" % k] = np.ascontiguousarray(W[:, k])
    b = B.Builder()
    m = B.Module("wide", n)
    for k in range(n_fib):
        u, v = E.Col("U_%d" % k), E.Col("V_%d" % k)
        m.assert_zero_except_at(E.Col("U_%d" % k, 1) - v, n - 1)
        m.assert_zero_except_at(E.Col("V_%d" % k, 1) - u - v, n - 1)
    for k in range(n_cubic):
        wk = E.Col("W_%d" % k)
        m.assert_zero_except_at(E.Col("W_%d" % k, 1) - wk * wk * wk - E.Const(int(cs[k])), n - 1)
    b.add_module(m)
    return B.compile_program(b), trace


def plonk_setup_instance(log_n, setup_seed, witness_seed, n_public=0):
    """A plonk-like module whose selectors are setup columns
    (PrepareIthPlonkWithSetup, integration_test/gnark_plonk/
    plonk_gnark_utils.go:217-232): qL*L + qR*R + qM*L*R + qO*O + qK (+ PI)
    with QL..QK precommitted, and a copy constraint over (L, R, O).

    The circuit (selectors, copy cycles) comes from setup_seed, so instances
    of one setup_seed share one proving key; the witness from witness_seed.
    Gates as in plonk_instance (40% add, 40% mul, 20% constant), qK uniform in
    [0, p); copy cycles tie 10% of the free cells (L and R of add and mul rows,
    R of constant rows); each cycle carries one random value, R and the free L
    are random, then O (add, mul rows) or L (constant rows) is solved from the
    gate.  With n_public > 0 the relation adds the public-input column PI
    (entries at rows 0 .. n_public-1, random values), and the gates absorb it.
    Returns (program, trace, public_inputs) -- trace holds the witness and the
    setup columns; public_inputs maps PI -> (module, [(row, value)])."""
    n = 1 << log_n
    pp = np.uint64(P)
    rs = np.random.Generator(np.random.PCG64(setup_seed))
    kind = rs.choice(3, size=n, p=[0.4, 0.4, 0.2])
    one, mone, zero = np.uint64(1), np.uint64(P - 1), np.uint64(0)
    ql = np.where(kind == 1, zero, one).astype(np.uint64)
    qr = np.where(kind == 0, one, zero).astype(np.uint64)
    qm = np.where(kind == 1, one, zero).astype(np.uint64)
    qo = np.where(kind == 2, zero, mone).astype(np.uint64)
    qk = rs.integers(0, P, size=n, dtype=np.uint64)
    free = np.concatenate([np.nonzero(kind != 2)[0], n + np.arange(n)])  # L of add/mul rows, every R
    tied = rs.permutation(free)[: 3 * n // 10]
    sigma = np.arange(3 * n, dtype=np.int64)
    cycles = []
    pos = 0
    while pos + 2 <= len(tied):
        ln = min(int(rs.integers(2, 5)), len(tied) - pos)
        cyc = tied[pos:pos + ln]
        cycles.append(cyc)
        for k in range(ln):
            sigma[cyc[k]] = cyc[(k + 1) % ln]
        pos += ln
    rw = np.random.Generator(np.random.PCG64(witness_seed))
    w = rw.integers(0, P, size=3 * n, dtype=np.uint64)
    for cyc in cycles:
        w[cyc] = w[cyc[0]]
    pi = np.zeros(n, dtype=np.uint64)
    pi_entries = []
    if n_public:
        vals = rw.integers(0, P, size=n_public, dtype=np.uint64)
        pi[:n_public] = vals
        pi_entries = [(i, int(vals[i])) for i in range(n_public)]
    L, R = w[:n].copy(), w[n:2 * n].copy()
    c = (qk + pi) % pp
    O = np.where(kind == 0, (L + R + c) % pp, np.where(kind == 1, (L * R % pp + c) % pp, zero))
    L = np.where(kind == 2, (pp - c) % pp, L)
    trace = dict(L=L, R=R, O=O.astype(np.uint64), QL=ql, QR=qr, QM=qm, QO=qo, QK=qk)
    b = B.Builder()
    m = B.Module("plonk", n)
    gate = (E.Setup("QL") * E.Col("L") + E.Setup("QR") * E.Col("R") + E.Setup("QM") * E.Col("L") * E.Col("R")
            + E.Setup("QO") * E.Col("O") + E.Setup("QK"))
    if n_public:
        gate = gate + E.PublicInput("PI")
    m.assert_zero(gate)
    b.add_module(m)
    B.copy_constraint(b, "plonk", [E.Col("L"), E.Col("R"), E.Col("O")], B.PermutationGen(sigma, "plonk.S"))
    public = {"PI": ("plonk", pi_entries)} if n_public else {}
    return B.compile_program(b), trace, public


def lookup_instance(kind, seed, log_n=8, log_t=6, bad=False):
    """Programs with the general lookups of arguments/lookup.go (tuples,
    selectors, unions), seeded with PCG64(seed).  kind:
    * "tuple":  LookupTuple of (A, B) of module "main" (2^log_n rows) into
      (X, Y) of module "tab" (2^log_t rows); the table repeats some pairs, so
      the first matching row takes the weight;
    * "cond":   CLookupTuple with selectors: source rows with selector 0 hold
      pairs outside the table, others selector 1 or 2; target selectors 0 on
      some rows (a later copy of that pair then takes its weight), else
      values 1..5; two unselected-in-target source rows of one pair whose
      selectors sum to 0 mod p (no error: their weight cancels);
    * "self":   as "tuple" with the table in module "main" itself (2^log_n
      rows, one module);
    * "union":  LookupUnion of columns of two source modules (2^log_n and
      2^(log_n-1) rows) into columns of two target modules (2^log_t and
      2^(log_t-1) rows);
    * "wide":   LookupTuple of 11 columns of "main" into 11 columns of "tab",
      two table rows of which differ in their last column only.
    bad: one source row takes a value in no target (the prove fails)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n, nt = 1 << log_n, 1 << log_t
    b = B.Builder()
    tr = {}

    def module(name, rows):
        m = B.Module(name, rows)
        b.add_module(m)
        return m

    if kind == "self":  # the table in the sources' own module (one module: one DEEP class)
        nt = n
    if kind in ("tuple", "cond", "self"):
        module("main", n)
        if kind != "self":
            module("tab", nt)
        X = rng.integers(0, P, size=nt, dtype=np.uint64)
        Y = rng.integers(0, P, size=nt, dtype=np.uint64)
        dup = 8 + rng.choice(nt - 9, size=nt // 8, replace=False)
        X[dup[1:]], Y[dup[1:]] = X[dup[0]], Y[dup[0]]  # one pair on several rows
        X[nt - 1], Y[nt - 1] = X[3], Y[3]
        if kind == "cond":
            X[5] = (X[5] + 1) % P  # a pair on no selected target row (below)
        rows = np.array([i for i in range(nt) if not (kind == "cond" and i == 5)])
        r = rows[rng.integers(0, len(rows), size=n)]
        A, Bv = X[r].copy(), Y[r].copy()
        tr.update(X=X, Y=Y, A=A, B=Bv)
        S = B.Table("main", [E.Col("A"), E.Col("B")])
        T = B.Table("main" if kind == "self" else "tab", [E.Col("X"), E.Col("Y")])
        if kind in ("tuple", "self"):
            B.lookup_tuple(b, S, T)
        else:
            SA = rng.choice(np.array([0, 1, 2], dtype=np.uint64), size=n, p=[0.2, 0.6, 0.2])
            off = SA == 0
            A[off] = rng.integers(0, P, size=int(off.sum()), dtype=np.uint64)  # junk, unselected
            ST = rng.integers(1, 6, size=nt).astype(np.uint64)
            ST[3] = 0  # row 3's pair is also on row nt-1, which takes its weight
            # row 5's pair is on no selected target row: two source rows of it
            # whose selectors cancel
            ST[5] = 0
            A[0], Bv[0], SA[0] = X[5], Y[5], 1
            A[1], Bv[1], SA[1] = X[5], Y[5], P - 1
            tr.update(SA=SA, ST=ST)
            B.c_lookup_tuple(b, S, T, E.Col("SA"), E.Col("ST"))
        if bad:
            A[n - 1] = (X[0] + 12345) % P
            Bv[n - 1] = (Y[0] + 777) % P
            if kind == "cond":
                tr["SA"][n - 1] = 1
    elif kind == "union":
        module("src0", n)
        module("src1", n // 2)
        module("tab0", nt)
        module("tab1", nt // 2)
        T0 = rng.integers(0, P, size=nt, dtype=np.uint64)
        T1 = rng.integers(0, P, size=nt // 2, dtype=np.uint64)
        T1[2] = T0[7]  # a value in both targets: tab0's row takes it
        pool = np.concatenate([T0, T1])
        S0 = pool[rng.integers(0, len(pool), size=n)]
        S1 = pool[rng.integers(0, len(pool), size=n // 2)]
        S1[0] = T1[2]
        if bad:
            S1[-1] = (int(T0.max()) + 1) % P
            while S1[-1] in set(pool.tolist()):
                S1[-1] = (S1[-1] + 1) % P
        tr.update(S0=S0, S1=S1, T0=T0, T1=T1)
        B.lookup_union(b, [B.Column("src0", E.Col("S0")), B.Column("src1", E.Col("S1"))],
                       [B.Column("tab0", E.Col("T0")), B.Column("tab1", E.Col("T1"))])
    elif kind == "wide":
        WD = 11
        module("main", n)
        module("tab", nt)
        X = rng.integers(0, P, size=(WD, nt), dtype=np.uint64)
        dup = 8 + rng.choice(nt - 9, size=nt // 8, replace=False)
        X[:, dup[1:]] = X[:, dup[0]][:, None]
        X[:, 5] = X[:, 4]
        X[WD - 1, 5] = (X[WD - 1, 4] + 1) % P  # rows 4 and 5 differ in the last column only
        r = rng.integers(0, nt, size=n)
        r[0], r[1] = 4, 5
        A = X[:, r].copy()
        if bad:  # row 4's tuple but for its last column
            A[:, n - 1] = X[:, 4]
            A[WD - 1, n - 1] = (X[WD - 1, 4] + 2) % P
        for c in range(WD):
            tr["X%d" % c], tr["A%d" % c] = X[c].copy(), A[c].copy()
        B.lookup_tuple(b, B.Table("main", [E.Col("A%d" % c) for c in range(WD)]),
                       B.Table("tab", [E.Col("X%d" % c) for c in range(WD)]))
    else:
        raise ValueError(kind)
    return B.compile_program(b), tr


def ext_public_instance(log_n, seed, bad=False):
    """A module with a base public input "PB" (expr.PublicInput, base
    entries) and an extension public input "PX" (expr.PublicInputExt) whose
    entries are E6 values except one base entry (lifted), used in relations
    that hold on every row: PB * Q = 0 and PX * Q = 0 with the base column Q
    zero at every public row and random elsewhere, and A * B - C = 0.
    Returns (program, trace, public_inputs); public_inputs: name -> (module,
    [(row, value)]), value an int or 6 limbs; the rows depend on log_n only.
    bad: Q is non-zero at one public row of PX (the relation fails there)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    n = 1 << log_n
    b = B.Builder()
    m = B.Module("pub", n)
    b.add_module(m)
    m.assert_zero(E.Col("A") * E.Col("B") - E.Col("C"))
    m.assert_zero(E.PublicInput("PB") * E.Col("Q"))
    m.assert_zero(E.PublicInputExt("PX") * E.Col("Q"))
    A = rng.integers(0, P, size=n, dtype=np.uint64)
    Bv = rng.integers(0, P, size=n, dtype=np.uint64)
    C = A * Bv % np.uint64(P)
    Q = rng.integers(1, P, size=n, dtype=np.uint64)
    # the public rows are part of the statement's shape (the plan): they
    # depend on log_n only, so statements of one size batch together
    rows = sorted(int(x) for x in np.random.Generator(np.random.PCG64(log_n)).choice(n, size=min(n, 6), replace=False))
    rb, rx = rows[:2], rows[2:]
    Q[rows] = 0
    if bad:
        Q[rx[0]] = 5
    pb = [(r, int(rng.integers(0, P))) for r in rb]
    px = [(r, [int(x) for x in rng.integers(0, P, size=6)]) for r in rx[:-1]] + [(rx[-1], int(rng.integers(0, P)))]
    tr = dict(A=A, B=Bv, C=C, Q=Q)
    return B.compile_program(b), tr, {"PB": ("pub", pb), "PX": ("pub", px)}


def ext_setup_instance(log_n, setup_seed, witness_seed):
    """A module with a base setup column S1 (expr.Setup) and an extension
    setup column SX (expr.ExtSetup, N x 6 limbs), both committed once in the
    setup tree: relations A*B - C = 0, S1*A - D = 0 and SX*Z = 0, with SX
    non-zero on a few rows and the witness column Z zero there and random
    elsewhere.  The setup columns depend on setup_seed only (one proving key
    for many witnesses).  Returns (program, trace) with the setup columns in
    the trace (SX as N x 6 limbs)."""
    n = 1 << log_n
    rs = np.random.Generator(np.random.PCG64(setup_seed))
    rw = np.random.Generator(np.random.PCG64(witness_seed))
    b = B.Builder()
    m = B.Module("ks", n)
    b.add_module(m)
    m.assert_zero(E.Col("A") * E.Col("B") - E.Col("C"))
    m.assert_zero(E.Setup("S1") * E.Col("A") - E.Col("D"))
    m.assert_zero(E.ExtSetup("SX") * E.Col("Z"))
    S1 = rs.integers(0, P, size=n, dtype=np.uint64)
    SX = np.zeros((n, 6), dtype=np.uint64)
    rows = rs.choice(n, size=max(1, n // 8), replace=False)
    SX[rows] = rs.integers(0, P, size=(len(rows), 6), dtype=np.uint64)
    A = rw.integers(0, P, size=n, dtype=np.uint64)
    Bv = rw.integers(0, P, size=n, dtype=np.uint64)
    Z = rw.integers(1, P, size=n, dtype=np.uint64)
    Z[rows] = 0
    pp = np.uint64(P)
    tr = dict(A=A, B=Bv, C=A * Bv % pp, D=S1 * A % pp, Z=Z, S1=S1, SX=SX)
    return B.compile_program(b), tr
