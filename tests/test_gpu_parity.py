"""Bit-exact parity of the HIP prove path against the oracle (GPU).

Every test calls libloomgpu.so through its C ABI and compares with the numpy
oracle on the same seeded inputs: RS encoding, leaf hashing and Merkle trees
(with injections) as building blocks, then whole proofs (canonical proof
words and every transcript digest) for the plonk-like workload, alone and
batched.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

P = 2**31 - 2**24 + 1


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def torch():
    import torch as _t
    if not _t.cuda.is_available():
        pytest.skip("no HIP device")
    return _t


@pytest.fixture(scope="module")
def ctx(torch):
    from loom_amd.prover import Context
    c = Context(0)
    yield c
    c.close()


def dev_u32(torch, arr):
    a = np.ascontiguousarray(arr, dtype=np.uint32)
    t = torch.from_numpy(a.view(np.int32).copy()).cuda()
    return t


def host_u32(t):
    return t.cpu().numpy().view(np.uint32)


def ptr(t):
    from loom_amd import _lib
    return ctypes.cast(ctypes.c_void_p(t.data_ptr()), _lib.u32p)


def mont(a):
    from oracle import field as F
    return F.to_mont(np.asarray(a, dtype=np.uint64))


def canon(a):
    from oracle import field as F
    return F.from_mont(np.asarray(a, dtype=np.uint32))


# 2^15 .. 2^21 take the fused four-step LDE: every row size of its middle
# pass (2^8 .. 2^11) and both factor splits (a = b and a = b + 1)
@pytest.mark.parametrize("logn", [1, 4, 10, 14, 15, 16, 17, 18, 19, 20, 21])
def test_rs_encode(torch, ctx, logn):
    from loom_amd import _lib
    from oracle import ntt
    log("rs_encode logn=%d" % logn)
    n = 1 << logn
    ncol = 3 if logn <= 17 else 1
    rng = np.random.default_rng(logn)
    cols = rng.integers(0, P, size=(ncol, n), dtype=np.uint64)
    din = dev_u32(torch, mont(cols).reshape(-1))
    dout = torch.zeros(ncol * 4 * n, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().lg_rs_encode(ctx.h, ptr(din), n, ncol, ptr(dout)), "lg_rs_encode")
    got = canon(host_u32(dout)).reshape(ncol, 4 * n)
    want = ntt.encode(cols)
    assert np.array_equal(got, want)


def test_rs_encode_max_size(torch, ctx):
    """N = 2^22 (the largest module): the first n rows of the bit-reversed LDE
    are the Lagrange values in bit-reversed order (w_{4n}^{4 bitrev_n(r)} =
    w_n^{bitrev_n(r)}), and sampled rows equal the barycentric value
    P(x) = (x^n - 1)/n * sum_i v_i w^i / (x - w^i) at x = w_{4n}^{bitrev(r)}."""
    from loom_amd import _lib
    from oracle import field as F
    logn = 22
    n = 1 << logn
    log("rs_encode logn=22 (properties)")
    rng = np.random.default_rng(22)
    col = rng.integers(0, P, size=n, dtype=np.uint64)
    din = dev_u32(torch, mont(col))
    dout = torch.zeros(4 * n, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().lg_rs_encode(ctx.h, ptr(din), n, 1, ptr(dout)), "lg_rs_encode")
    got = canon(host_u32(dout))
    br = F.bitrev_indices(n)
    assert np.array_equal(got[:n], col[br])
    wi = F.powers(F.omega(n), n)
    w4 = F.omega(4 * n)
    for r in [n + 5, 2 * n + 12345, 3 * n + 999_999, 4 * n - 1]:
        rb = int(format(r, "0%db" % (logn + 2))[::-1], 2)
        x = pow(w4, rb, P)
        den = (np.uint64(x) + np.uint64(P) - wi) % np.uint64(P)
        terms = col * wi % np.uint64(P) * F.vinv(den) % np.uint64(P)
        val = int(np.sum(terms, dtype=np.uint64) % np.uint64(P))  # < 2^22 terms below 2^31: no overflow
        val = val * (pow(x, n, P) - 1) % P * pow(n, P - 2, P) % P
        assert int(got[r]) == val, r


@pytest.mark.parametrize("nb,ne,logrows", [(9, 0, 3), (5, 1, 6), (0, 4, 8), (1, 0, 12), (0, 1, 15)])
def test_leaf_hash(torch, ctx, nb, ne, logrows):
    from loom_amd import _lib
    from oracle import poseidon2 as p2
    log("leaf_hash nb=%d ne=%d rows=2^%d" % (nb, ne, logrows))
    rows = 1 << logrows
    rng = np.random.default_rng(nb * 100 + ne)
    base = [rng.integers(0, P, size=rows, dtype=np.uint64) for _ in range(nb)]
    ext = [rng.integers(0, P, size=(rows, 6), dtype=np.uint64) for _ in range(ne)]
    db = [dev_u32(torch, mont(c)) for c in base]
    de = [dev_u32(torch, mont(c.T.copy()).reshape(-1)) for c in ext]
    bp = (_lib.u32p * max(1, nb))(*[ptr(t) for t in db])
    ep = (_lib.u32p * max(1, ne))(*[ptr(t) for t in de])
    dig = torch.zeros(rows // 2 * 8, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().lg_hash_leaf_pairs(ctx.h, bp, nb, ep, ne, rows, ptr(dig)), "lg_hash_leaf_pairs")
    got = canon(host_u32(dig)).reshape(rows // 2, 8).T
    want = p2.hash_leaf_pairs(base, ext)
    assert np.array_equal(got, want)


def test_leaf_hash_extreme_values(torch, ctx):
    """Leaf sponges of rows at the ends of the field (all p - 1, zeros,
    alternating, p - 2 and 1): the device permutation's lazy ranges (S-box
    outputs below 2p into the 64-bit external layer, scaled partial-round
    lanes below 2p) meet their largest first-round sums; equal to the oracle."""
    from loom_amd import _lib
    from oracle import poseidon2 as p2
    log("leaf_hash extreme values")
    rows = 64
    pats = [np.full(rows, P - 1, dtype=np.uint64), np.zeros(rows, dtype=np.uint64),
            np.where(np.arange(rows) % 2 == 0, P - 1, 0).astype(np.uint64),
            np.full(rows, P - 2, dtype=np.uint64), np.ones(rows, dtype=np.uint64)]
    base = [pats[k % len(pats)].copy() for k in range(16)]
    base[3][::3] = P - 1
    ext = [np.full((rows, 6), P - 1, dtype=np.uint64)]
    db = [dev_u32(torch, mont(c)) for c in base]
    de = [dev_u32(torch, mont(c.T.copy()).reshape(-1)) for c in ext]
    bp = (_lib.u32p * len(db))(*[ptr(t) for t in db])
    ep = (_lib.u32p * len(de))(*[ptr(t) for t in de])
    dig = torch.zeros(rows // 2 * 8, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().lg_hash_leaf_pairs(ctx.h, bp, len(db), ep, len(de), rows, ptr(dig)), "lg_hash_leaf_pairs")
    got = canon(host_u32(dig)).reshape(rows // 2, 8).T
    assert np.array_equal(got, p2.hash_leaf_pairs(base, ext))
    # Merkle nodes over all-(p - 1) and all-zero leaves
    for fill in (P - 1, 0):
        n = 64
        leaves = np.full((8, n), fill, dtype=np.uint64)
        nodes = np.zeros((2 * n, 8), dtype=np.uint32)
        nodes[n:] = mont(leaves.T)
        dn = dev_u32(torch, nodes.reshape(-1))
        ip = (_lib.u32p * 1)()
        iw = (ctypes.c_size_t * 1)()
        _lib.check(_lib.lib().lg_merkle_build(ctx.h, ptr(dn), n, ip, iw, 0), "lg_merkle_build")
        from oracle.merkle import Tree
        assert np.array_equal(canon(host_u32(dn)).reshape(2 * n, 8)[1:], Tree(leaves, []).nodes.T[1:])


@pytest.mark.parametrize("logn,inj", [(0, []), (3, []), (5, [32, 8, 8, 1]), (10, [1024, 64]), (15, [])])
def test_merkle(torch, ctx, logn, inj):
    from loom_amd import _lib
    from oracle.merkle import Tree
    log("merkle leaves=2^%d inj=%s" % (logn, inj))
    n = 1 << logn
    rng = np.random.default_rng(logn + 7)
    leaves = rng.integers(0, P, size=(8, n), dtype=np.uint64)
    injs = [rng.integers(0, P, size=(8, w), dtype=np.uint64) for w in inj]
    nodes = np.zeros((2 * n, 8), dtype=np.uint32)
    nodes[n:] = mont(leaves.T)
    dn = dev_u32(torch, nodes.reshape(-1))
    di = [dev_u32(torch, mont(x.T.copy()).reshape(-1)) for x in injs]
    ip = (_lib.u32p * max(1, len(di)))(*[ptr(t) for t in di])
    iw = (ctypes.c_size_t * max(1, len(inj)))(*inj)
    _lib.check(_lib.lib().lg_merkle_build(ctx.h, ptr(dn), n, ip, iw, len(inj)), "lg_merkle_build")
    got = canon(host_u32(dn)).reshape(2 * n, 8)
    t = Tree(leaves, injs)
    assert np.array_equal(got[1:], t.nodes.T[1:])


@pytest.mark.parametrize("logn", [0, 3, 6, 10, 12, 13, 14, 16, 17])
def test_grand_product(torch, ctx, logn):
    from loom_amd import _lib
    from oracle import field as F
    log("grand product n=2^%d" % logn)
    n = 1 << logn
    rng = np.random.default_rng(logn + 3)
    num = rng.integers(1, P, size=(n, 6), dtype=np.uint64)
    den = rng.integers(1, P, size=(n, 6), dtype=np.uint64)
    dn = dev_u32(torch, mont(num.T.copy()).reshape(-1))
    dd = dev_u32(torch, mont(den.T.copy()).reshape(-1))
    dg = torch.zeros(6 * n, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().lg_grand_product(ctx.h, ptr(dn), ptr(dd), ptr(dg), n), "lg_grand_product")
    got = canon(host_u32(dg)).reshape(6, n).T
    ratio = F.e6_mul(num, F.e6_inv(den))
    want = np.zeros((n, 6), dtype=np.uint64)
    want[0, 0] = 1
    for i in range(1, n):
        want[i] = F.e6_mul(want[i - 1], ratio[i - 1])
    assert np.array_equal(got, want)


@pytest.mark.parametrize("logn", [20, 22])
def test_grand_product_large(torch, ctx, logn):
    """Multi-workgroup scans (tiles of 2048 rows, tile prefixes scanned per
    instance): gp[0] = 1 and gp[i+1] den[i] = gp[i] num[i] for every row."""
    from loom_amd import _lib
    from oracle import field as F
    log("grand product n=2^%d" % logn)
    n = 1 << logn
    rng = np.random.default_rng(logn + 5)
    num = rng.integers(1, P, size=(n, 6), dtype=np.uint64)
    den = rng.integers(1, P, size=(n, 6), dtype=np.uint64)
    dn = dev_u32(torch, mont(num.T.copy()).reshape(-1))
    dd = dev_u32(torch, mont(den.T.copy()).reshape(-1))
    dg = torch.zeros(6 * n, dtype=torch.int32, device="cuda")
    _lib.check(_lib.lib().lg_grand_product(ctx.h, ptr(dn), ptr(dd), ptr(dg), n), "lg_grand_product")
    got = canon(host_u32(dg)).reshape(6, n).T
    assert list(got[0]) == [1, 0, 0, 0, 0, 0]
    assert np.array_equal(F.e6_mul(got[1:], den[:-1]), F.e6_mul(got[:-1], num[:-1]))


def _oracle_words(prog, trace, nq):
    from oracle import prover as OP, field as F
    pf, enc, logs = OP.prove(prog, trace, nq)
    tw = np.array([F.to_mont(v) for _, d in logs for v in d], dtype=np.uint32)
    return enc, tw


@pytest.mark.parametrize("logn,nq", [(1, 3), (3, 8), (6, 16), (10, 40)])
def test_prove_plonk_small(ctx, logn, nq):
    from loom_amd import workloads, prover
    log("prove plonk logn=%d nq=%d" % (logn, nq))
    prog, tr = workloads.plonk_instance(logn, 11 + logn)
    enc, tw = _oracle_words(prog, tr, nq)
    words, trans = prover.prove(ctx, prog, tr, nq, want_transcript=True)
    assert np.array_equal(trans, tw), "transcript digests differ"
    assert words.shape == enc.shape
    assert np.array_equal(words, enc)


@pytest.mark.parametrize("logn,nq", [(15, 30), (16, 40)])
def test_prove_plonk_large(ctx, logn, nq):
    """Modules above 2^14 take the four-step NTT path."""
    from loom_amd import workloads, prover
    log("prove plonk logn=%d nq=%d" % (logn, nq))
    prog, tr = workloads.plonk_instance(logn, 7 + logn)
    enc, tw = _oracle_words(prog, tr, nq)
    words, trans = prover.prove(ctx, prog, tr, nq, want_transcript=True)
    assert np.array_equal(trans, tw), "transcript digests differ"
    assert np.array_equal(words, enc)


@pytest.mark.parametrize("log_ns,nq", [((4, 6, 5, 6, 3), 12), ((10, 12, 8, 14), 30), ((15, 9, 12, 12), 20)])
def test_prove_uneven_modules(ctx, log_ns, nq):
    """One program with modules of different sizes: groups become Merkle
    level injections (equal sizes fold into the leaves), one DEEP class per
    module, FRI levels entering at different rounds."""
    from loom_amd import workloads, prover
    log("prove uneven modules %s nq=%d" % (log_ns, nq))
    prog, tr = workloads.uneven_instance(list(log_ns), 5)
    enc, tw = _oracle_words(prog, tr, nq)
    words, trans = prover.prove(ctx, prog, tr, nq, want_transcript=True)
    assert np.array_equal(trans, tw), "transcript digests differ"
    assert np.array_equal(words, enc)


def test_prove_uneven_batched(ctx):
    """Three instances of one uneven shape in one batched call."""
    from loom_amd import workloads, prover
    log("prove uneven batched 3 x (7, 9, 6)")
    inst = [workloads.uneven_instance([7, 9, 6], 40 + s) for s in range(3)]
    batch = prover.prove_batch(ctx, inst, 16)
    for (prog, tr), w in zip(inst, batch):
        enc, _ = _oracle_words(prog, tr, 16)
        assert np.array_equal(w, enc)


@pytest.mark.parametrize("log_ns,nq", [((4, 6, 3), 12), ((10, 12, 9), 20), ((16, 9), 16)])
def test_prove_range_lookups(ctx, log_ns, nq):
    """Modules with a logup range lookup into the shared range_256 module:
    multiplicity histograms, cyclic logup columns, exposed totals bound to
    the transcript and written to the proof."""
    from loom_amd import workloads, prover
    log("prove range lookups %s nq=%d" % (log_ns, nq))
    prog, tr = workloads.uneven_range_instance(list(log_ns), 8)
    enc, tw = _oracle_words(prog, tr, nq)
    words, trans = prover.prove(ctx, prog, tr, nq, want_transcript=True)
    assert np.array_equal(trans, tw), "transcript digests differ"
    assert np.array_equal(words, enc)


def test_prove_range_lookups_repeated_sizes(ctx):
    """Modules of one size run their prover steps as one class (one launch per
    phase for the grand products, the multiplicity sources and the logups)."""
    from loom_amd import workloads, prover
    log("prove range lookups, repeated sizes (7, 5, 7, 7, 5)")
    prog, tr = workloads.uneven_range_instance([7, 5, 7, 7, 5], 9)
    enc, tw = _oracle_words(prog, tr, 12)
    words, trans = prover.prove(ctx, prog, tr, 12, want_transcript=True)
    assert np.array_equal(trans, tw), "transcript digests differ"
    assert np.array_equal(words, enc)


def test_prove_range_lookups_batched(ctx):
    from loom_amd import workloads, prover
    log("prove range lookups batched 3 x (6, 8)")
    inst = [workloads.uneven_range_instance([6, 8], 60 + s) for s in range(3)]
    batch = prover.prove_batch(ctx, inst, 10)
    for (prog, tr), w in zip(inst, batch):
        enc, _ = _oracle_words(prog, tr, 10)
        assert np.array_equal(w, enc)


def test_range_lookup_value_outside_table_is_an_error(ctx):
    from loom_amd import workloads, prover, _lib
    log("range lookup with a value outside the table")
    prog, tr = workloads.uneven_range_instance([5], 3)
    tr = dict(tr)
    tr["L_0"] = tr["L_0"].copy()
    tr["L_0"][7] = 300
    with pytest.raises(_lib.LoomGpuError, match="not in the range table"):
        prover.prove(ctx, prog, tr, 4)


@pytest.mark.parametrize("logn,nf,nc,nq", [(10, 4, 4, 16), (15, 3, 2, 12)])
def test_prove_wide(ctx, logn, nf, nc, nq):
    """The wide Fibonacci/cubic trace (shifted columns, AssertZeroExceptAt
    through a Lagrange column, one round, first FRI round at 22 - log N)."""
    from loom_amd import workloads, prover
    log("prove wide logn=%d fib=%d cubic=%d nq=%d" % (logn, nf, nc, nq))
    prog, tr = workloads.wide_instance(logn, 3, nf, nc)
    enc, tw = _oracle_words(prog, tr, nq)
    words, trans = prover.prove(ctx, prog, tr, nq, want_transcript=True)
    assert np.array_equal(trans, tw), "transcript digests differ"
    assert np.array_equal(words, enc)


def test_prove_plonk_full_size(ctx):
    """Config-0 shape: N = 2^14, 200 queries."""
    from loom_amd import workloads, prover
    log("prove plonk logn=14 nq=200")
    prog, tr = workloads.plonk_instance(14, 0)
    enc, tw = _oracle_words(prog, tr, 200)
    words, trans = prover.prove(ctx, prog, tr, 200, want_transcript=True)
    assert np.array_equal(trans, tw)
    assert np.array_equal(words, enc)


def test_batched_equals_alone(ctx):
    """Instances proved together in one batched call give exactly the proofs
    they give alone (config-1 pattern at reduced size)."""
    from loom_amd import workloads, prover
    log("batched vs alone, 5 x 2^8")
    inst = [workloads.plonk_instance(8, 100 + s) for s in range(5)]
    batch = prover.prove_batch(ctx, inst, 20)
    for (prog, tr), w in zip(inst, batch):
        alone = prover.prove(ctx, prog, tr, 20)
        assert np.array_equal(w, alone)
    enc, _ = _oracle_words(inst[2][0], inst[2][1], 20)
    assert np.array_equal(batch[2], enc)


def test_batch_split_over_lanes_equals_alone(ctx):
    """A batch of 17 is proved as two concurrent sub-batches (lanes, own
    streams and host pools); every proof equals the instance's proof alone."""
    from loom_amd import workloads, prover
    log("batch of 17 x 2^6 over lanes vs alone")
    inst = [workloads.plonk_instance(6, 700 + s) for s in range(17)]
    items = [prover.Prepared(p, t, 10) for p, t in inst]
    batch = prover.prove_prepared(ctx, items)
    for it, w in zip(items, batch):
        assert np.array_equal(w, prover.prove_prepared(ctx, [it])[0])
    enc, _ = _oracle_words(inst[16][0], inst[16][1], 10)
    assert np.array_equal(batch[16], enc)


def test_device_inputs_equal_host_inputs(ctx):
    """lg_prove_device (trace already in HBM) gives the proofs lg_prove gives
    from host buffers, and a host pointer handed to it is refused."""
    from loom_amd import workloads, prover, _lib
    log("device inputs vs host inputs, 3 x 2^9")
    inst = [workloads.plonk_instance(9, 300 + s) for s in range(3)]
    items = [prover.Prepared(p, t, 24) for p, t in inst]
    dv = prover.DeviceInputs(items, 0)
    a = prover.prove_prepared(ctx, items, dev_inputs=dv)
    b = prover.prove_prepared(ctx, items)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    enc, _ = _oracle_words(inst[1][0], inst[1][1], 24)
    assert np.array_equal(a[1], enc)

    class HostAsDevice:
        def __init__(self, it):
            self.items = [it]
            self.columns = [[_HostPtr(c) for c in it.columns]]
            self.perms = [[_HostPtr(s) for s in it.perms]]

    with pytest.raises(_lib.LoomGpuError, match="not device memory"):
        prover.prove_prepared(ctx, items[:1], dev_inputs=HostAsDevice(items[0]))


class _HostPtr:
    def __init__(self, arr):
        self.arr = arr

    def data_ptr(self):
        return self.arr.ctypes.data


@pytest.mark.gpu
def test_graph_replay_equals_eager():
    """HIP graphs per shape bucket: the same shape proved repeatedly with new
    witnesses (eager, then recorded, then replayed) gives the proofs and
    transcripts of a context without graphs -- one lane (batch 3), two lanes
    (batch 16), one instance per call from device inputs held at different
    addresses at once, and a plan with range lookups whose replayed batch
    holds an invalid instance that fails alone.  No replay may fall back to
    an eager prove."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, workloads
    g = prover.Context(0)
    e = prover.Context(0)
    e.set_graphs(False)
    try:
        held = []  # device inputs stay alive, so every call's addresses differ
        for B, logn in ((3, 9), (16, 8)):
            for rep in range(4):
                log("graphs: batch %d, 2^%d rows, call %d" % (B, logn, rep))
                items = [prover.Prepared(*workloads.plonk_instance(logn, 1000 * rep + 10 * B + k), 24)
                         for k in range(B)]
                a, ta = prover.prove_prepared(g, items, want_transcript=True)
                b, tb = prover.prove_prepared(e, items, want_transcript=True)
                for k in range(B):
                    assert np.array_equal(a[k], b[k]), (B, rep, k)
                    assert np.array_equal(ta[k], tb[k]), (B, rep, k)
                # the same shape from device inputs is a bucket of its own
                held.append(prover.DeviceInputs(items))
                d = prover.prove_prepared(g, items, dev_inputs=held[-1])
                for k in range(B):
                    assert np.array_equal(d[k], b[k]), (B, rep, k)
            st = prover.graph_stats(g, items)
            log("graphs: batch %d stats %s" % (B, st))
            # per lane and input kind: call 0 eager, call 1 records, calls 2-3 replay
            assert st["fallbacks"] == 0 and st["records"] >= 2 and st["replays"] == 2 * st["records"], st
        # one instance per call, each from its own device inputs (the plain path)
        items = [prover.Prepared(*workloads.plonk_instance(8, 500 + k), 24) for k in range(5)]
        dev = [prover.DeviceInputs([it]) for it in items]
        want = [prover.prove_prepared(e, [it])[0] for it in items]
        for rep in range(3):
            log("graphs: one instance per call from device inputs, pass %d" % rep)
            for k in range(5):
                got = prover.prove_prepared(g, [items[k]], dev_inputs=dev[k])[0]
                assert np.array_equal(got, want[k]), (rep, k)
        st = prover.graph_stats(g, items)
        log("graphs: single-instance stats %s" % st)
        assert st["fallbacks"] == 0 and st["replays"] >= 12, st
        for rep in range(4):
            log("graphs: range lookups, call %d" % rep)
            items = [prover.Prepared(*workloads.uneven_range_instance([8], 70 + 5 * rep + k), 16) for k in range(4)]
            if rep == 3:
                bad = items[1]
                bad.columns = [c.copy() for c in bad.columns]
                bad.columns[bad.plan.input_names().index("L_0")][4] = prover.to_mont([4000])[0]
            a = prover.prove_prepared(g, items, allow_failures=True)
            b = prover.prove_prepared(e, items, allow_failures=True)
            for k in range(4):
                if rep == 3 and k == 1:
                    assert a[k] is None and b[k] is None
                else:
                    assert np.array_equal(a[k], b[k]), (rep, k)
    finally:
        g.close()
        e.close()


@pytest.mark.gpu
def test_tuning_choices_keep_the_proofs(tmp_path, monkeypatch):
    """Launch choices change how a prove runs, not what it proves: the proofs
    under extreme settings of every knob equal those under the defaults; the
    autotuner returns a valid choice, keeps it for the batch size and appends
    it to LOOMGPU_TUNE_FILE."""
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from loom_amd import prover, workloads
    table = tmp_path / "tune.txt"
    monkeypatch.setenv("LOOMGPU_TUNE_FILE", str(table))
    items = [prover.Prepared(*workloads.plonk_instance(11, 300 + k), 16) for k in range(16)]
    ctx = prover.Context(0)
    try:
        want = prover.prove_prepared(ctx, items)
        for tun in (dict(lanes=1, merkle_wide_min=1 << 13, hash_threads=64, fri_coop_pairs=2048, claims_rows=4096),
                    dict(lanes=2, merkle_wide_min=1 << 17, hash_threads=128, fri_coop_pairs=128, claims_rows=512)):
            log("tuning %s" % tun)
            prover.set_tuning(ctx, items, tun)
            for _ in range(3):  # eager, recorded, replayed
                got = prover.prove_prepared(ctx, items)
                for a, b in zip(got, want):
                    assert np.array_equal(a, b), tun
        r = prover.autotune(ctx, items, reps=1)
        log("autotune: %s" % r)
        t = r["tuning"]
        assert t["lanes"] in (1, 2) and t["hash_threads"] in (64, 128, 256) and r["trials"] >= 10
        assert r["best_ms"] <= r["start_ms"] * 1.05
        got = prover.prove_prepared(ctx, items)
        for a, b in zip(got, want):
            assert np.array_equal(a, b)
        lines = table.read_text().split("\n")
        assert any(ln.startswith("LGTUNE1 ") and " 16 " in ln for ln in lines)
    finally:
        ctx.close()
