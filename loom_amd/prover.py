"""Host-side ``Prove`` (mirror of prover.Prove, prover/prover.go:655-719) over
the HIP library.

    ctx = Context(0)
    proof_words = prove(ctx, program, trace, num_queries=200)
    many = prove_batch(ctx, [(program_i, trace_i), ...], num_queries=200)

``trace`` maps column names to canonical values in [0, p) (numpy arrays);
they are converted to Montgomery u32, the reference's in-memory layout.  The
result is the canonical proof word stream (include/loomgpu.h); ``proof.decode``
turns it back into a structure.  Errors from the library raise
``LoomGpuError`` -- there is no CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from .plan import Plan
from . import protocol as PR

P = 2**31 - 2**24 + 1
R = (1 << 32) % P


def to_mont(a):
    a = np.asarray(a, dtype=np.uint64) % np.uint64(P)
    return np.ascontiguousarray((a * np.uint64(R) % np.uint64(P)).astype(np.uint32))


class Context:
    def __init__(self, device=0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.lg_ctx_create(device, ctypes.byref(h)), "lg_ctx_create")
        self.h = h
        self.plans = {}
        self.setups = {}

    def close(self):
        if self.h:
            for p in self.plans.values():
                _lib.lib().lg_plan_destroy(p)
            self.plans.clear()
            _lib.lib().lg_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def issue_probe(self):
        """Measured issue rates (lg_probe_issue_rate): lane operations per ns
        of v_add_u32, v_mul_lo_u32, v_mad_u64_u32, Montgomery products, and
        the clock in MHz."""
        v = (ctypes.c_double * 5)()
        _lib.check(_lib.lib().lg_probe_issue_rate(self.h, v), "lg_probe_issue_rate")
        return dict(add=v[0], mul_lo=v[1], mad_u64=v[2], montgomery=v[3], clock_mhz=v[4])

    def set_graphs(self, on):
        _lib.check(_lib.lib().lg_ctx_set_graphs(self.h, 1 if on else 0), "lg_ctx_set_graphs")

    def set_timing(self, on):
        _lib.check(_lib.lib().lg_ctx_set_timing(self.h, 1 if on else 0), "lg_ctx_set_timing")

    def timing_report(self):
        """{kernel: (launches, device ms, algorithmic bytes, permutations)};
        clears the library's totals."""
        n = ctypes.c_size_t()
        buf = ctypes.create_string_buffer(1 << 16)
        _lib.check(_lib.lib().lg_ctx_timing_report(self.h, buf, len(buf), ctypes.byref(n)), "lg_ctx_timing_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms, by, pm = line.split()
            out[name] = (int(cnt), float(ms), float(by), float(pm))
        return out

    def plan_handle(self, words, setup=None):
        """The plan for these words; with setup columns (Montgomery arrays in
        plan setup order) they are committed once (lg_plan_set_setup) and kept
        on the device until other setup columns are given."""
        key = words.tobytes()
        if key not in self.plans:
            h = ctypes.c_void_p()
            w = np.ascontiguousarray(words, dtype=np.uint32)
            _lib.check(_lib.lib().lg_plan_create(self.h, w.ctypes.data_as(_lib.u32p), w.size, ctypes.byref(h)),
                       "lg_plan_create")
            self.plans[key] = h
        if setup:
            import hashlib
            digest = hashlib.sha256(b"".join(np.ascontiguousarray(c).tobytes() for c in setup)).digest()
            if self.setups.get(key) != digest:
                set_setup(self.plans[key], setup)
                self.setups[key] = digest
        return self.plans[key]


def set_setup(plan_h, setup):
    """Commits a plan's setup columns (setup.Setup): returns the setup root
    (the verification key), canonical values."""
    cols = [np.ascontiguousarray(c, dtype=np.uint32) for c in setup]
    ptrs = (_lib.u32p * len(cols))(*[c.ctypes.data_as(_lib.u32p) for c in cols])
    root = np.zeros(8, dtype=np.uint32)
    _lib.check(_lib.lib().lg_plan_set_setup(plan_h, ptrs, len(cols), root.ctypes.data_as(_lib.u32p)),
               "lg_plan_set_setup")
    return from_mont(root)


def from_mont(a):
    rinv = pow(R, P - 2, P)
    return (np.asarray(a, dtype=np.uint64) * np.uint64(rinv) % np.uint64(P)).astype(np.uint64)


def check_plan(words):
    """Parse + hiprtc-compile a plan without a device; returns code bytes."""
    w = np.ascontiguousarray(words, dtype=np.uint32)
    n = ctypes.c_size_t()
    _lib.check(_lib.lib().lg_plan_check(w.ctypes.data_as(_lib.u32p), w.size, ctypes.byref(n)), "lg_plan_check")
    return n.value


class Prepared:
    """An instance ready for lg_prove: the plan words and the Montgomery
    input arrays (kept alive for the duration of the call).  The trace holds
    the witness columns and, for a program with setup columns, the proving
    key's setup columns; public_inputs maps a public-input column to
    (module, [(row, value), ...]) (base-field entries)."""

    def __init__(self, program, trace, num_queries=PR.NUM_QUERIES_DEFAULT, public_inputs=None):
        public_inputs = dict(public_inputs or {})
        self.public_inputs = public_inputs
        self.plan = Plan(program, num_queries, {k: [(int(i), PR.public_entry_field(v)) for i, v in e[1]]
                                                for k, e in public_inputs.items()})
        self.words = self.plan.words()
        self.columns = [to_mont(trace[name]) for name in self.plan.input_names()]
        for name in self.plan.public:
            ents = sorted(((int(i), v) for i, v in public_inputs[name][1]), key=lambda e: e[0])
            self.columns.append(to_mont([x for _, v in ents for x in PR.public_entry_words(v)]))
        self.perms = self.plan.perm_arrays()
        # setup columns: N words, or for an extension column (N x 6 limbs) the
        # 6 limb arrays one after another (lg_plan_set_setup)
        self.setup = [to_mont(np.asarray(trace[name]).T.reshape(-1) if np.ndim(trace[name]) == 2 else trace[name])
                      for name in self.plan.setup]
        # the plan's identity, computed once (per call it would cost more than
        # the batched prove's host work)
        self.words_key = self.words.tobytes()
        self._addr_arrays = None

    def host_addresses(self):
        """Host addresses of the input columns and permutations.  Cached
        while .columns and .perms hold the same array objects (the cache keeps
        them referenced); replacing an array recomputes them."""
        cache = self._addr_arrays
        if (cache is None or len(cache[0]) != len(self.columns) or len(cache[1]) != len(self.perms)
                or any(a is not b for a, b in zip(cache[0], self.columns))
                or any(a is not b for a, b in zip(cache[1], self.perms))):
            cols, perms = list(self.columns), list(self.perms)
            perm_c = [np.ascontiguousarray(q) for q in perms]  # referenced with the addresses
            cache = (cols, perms, [c.ctypes.data for c in cols], [q.ctypes.data for q in perm_c], perm_c)
            self._addr_arrays = cache
        return cache[2], cache[3]


_proof_cap = {}


class _OutputPool:
    """Reusable host buffers for proof words.  A proof is returned as a view
    of one buffer; the buffer is handed out again only once no view of it is
    referenced any more, so results stay valid for as long as the caller keeps
    them.  Reuse avoids the page faults of fresh multi-megabyte allocations
    on every call (first-touch costs ~0.4 ms per MB on the host)."""

    MAX_PER_CAP = 256

    def __init__(self):
        import threading
        self._lock = threading.Lock()
        self._bufs = {}

    def take(self, cap, n):
        import sys
        with self._lock:
            pool = self._bufs.setdefault(cap, [])
            out = []
            for a in pool:
                if len(out) == n:
                    break
                # referenced by the pool list, the loop variable and the call
                # argument only: no view of it is alive
                if sys.getrefcount(a) <= 3:
                    out.append(a)
            while len(out) < n:
                a = np.empty(cap, dtype=np.uint32)
                if len(pool) < self.MAX_PER_CAP:
                    pool.append(a)
                out.append(a)
            return out


_out_pool = _OutputPool()


class DeviceInputs:
    """Prepared instances with their trace columns and permutations copied
    into device memory (torch tensors on `device`), for lg_prove_device."""

    def __init__(self, items, device=0):
        import torch
        dev = torch.device("cuda", device)
        self.items = items
        self.columns = [[torch.from_numpy(c.view(np.int32)).to(dev) for c in it.columns] for it in items]
        self.perms = [[torch.from_numpy(np.ascontiguousarray(s).view(np.int32)).to(dev) for s in it.perms]
                      for it in items]
        torch.cuda.synchronize(dev)
        self.tables = _address_tables([[c.data_ptr() for c in cs] for cs in self.columns],
                                      [[q.data_ptr() for q in ps] for ps in self.perms])


class _AddressTable:
    """A C array of addresses passed as uint32_t* const* (the array stays
    referenced by the object)."""

    def __init__(self, addrs):
        self.arr = (ctypes.c_void_p * max(1, len(addrs)))(*addrs)
        self.ptr = ctypes.cast(self.arr, _lib.u32pp)


def _address_tables(col_addrs, perm_addrs):
    return (_AddressTable([a for row in col_addrs for a in row]),
            _AddressTable([a for row in perm_addrs for a in row]))


def _pointer_tables(items, dev_inputs):
    """(columns, permutations) tables of the call: instance-major, plan
    input order."""
    if dev_inputs is not None:
        # a DeviceInputs keeps its tables; any object with per-instance lists
        # of device tensors (.columns, .perms) also serves
        tabs = getattr(dev_inputs, "tables", None)
        if tabs is not None:
            return tabs
        return _address_tables([[c.data_ptr() for c in cs] for cs in dev_inputs.columns],
                               [[q.data_ptr() for q in ps] for ps in dev_inputs.perms])
    addrs = [it.host_addresses() for it in items]
    return _address_tables([a[0] for a in addrs], [a[1] for a in addrs])


def prove_prepared(ctx, items, want_transcript=False, phases=None, dev_inputs=None, allow_failures=False,
                   fs_prev=None):
    """Proves prepared instances sharing one plan shape in one batched call.
    With `dev_inputs` (a DeviceInputs of these items) the inputs are read
    from device memory (lg_prove_device), else from host memory (lg_prove).

    An instance with an invalid witness (a lookup value outside its table, a
    permutation entry outside its cells) fails alone: by default the call
    raises LoomGpuError naming it; with `allow_failures` its result is None
    and the other instances' proofs are returned.

    fs_prev: per instance the last digest (8 canonical values) of a running
    transcript the proof's challenges follow (prover.WithTranscript,
    lg_prove_chained); None for fresh transcripts."""
    if not items:
        return []
    words = items[0].words
    key = items[0].words_key
    for it in items[1:]:
        if it.words_key != key:
            raise ValueError("batched instances must share one plan (same program shape)")
    if dev_inputs is not None and len(dev_inputs.items) != len(items):
        raise ValueError("dev_inputs does not match items")
    if items[0].setup:
        for it in items[1:]:
            if len(it.setup) != len(items[0].setup) or any(not np.array_equal(a, b)
                                                          for a, b in zip(it.setup, items[0].setup)):
                raise ValueError("batched instances must share one proving key (setup columns)")
    ph = ctx.plan_handle(words, items[0].setup)
    B = len(items)
    col_tab, perm_tab = _pointer_tables(items, dev_inputs)
    col_ptrs, perm_ptrs = col_tab.ptr, perm_tab.ptr
    entry = _lib.lib().lg_prove if dev_inputs is None else _lib.lib().lg_prove_device
    if fs_prev is not None:
        prev = np.ascontiguousarray(to_mont(np.asarray(fs_prev, dtype=np.uint64).reshape(-1)), dtype=np.uint32)
        if prev.size != 8 * len(items):
            raise ValueError("fs_prev: 8 values per instance")
        dev_flag = 0 if dev_inputs is None else 1

        def entry(ph_, B_, cp_, pp_, *rest, _p=prev, _d=dev_flag):
            return _lib.lib().lg_prove_chained(ph_, B_, cp_, pp_, _d, _p.ctypes.data_as(_lib.u32p), *rest)
    cap = _proof_cap.get(key, 1 << 21)
    while True:
        outs = _out_pool.take(cap, B)
        touts = [np.zeros(1 << 16, dtype=np.uint32) for _ in range(B)] if want_transcript else None
        out_tab = _AddressTable([o.ctypes.data for o in outs])
        out_ptrs = out_tab.ptr
        caps = (ctypes.c_size_t * B)(*([cap] * B))
        lens = (ctypes.c_size_t * B)()
        if want_transcript:
            t_ptrs = (_lib.u32p * B)(*[o.ctypes.data_as(_lib.u32p) for o in touts])
            tcaps = (ctypes.c_size_t * B)(*([1 << 16] * B))
            tlens = (ctypes.c_size_t * B)()
        else:
            t_ptrs, tcaps, tlens = None, None, None
        rc = entry(ph, B, col_ptrs, perm_ptrs, out_ptrs, caps, lens, t_ptrs, tcaps, tlens)
        if rc == -2 and max(lens) > cap:
            cap = max(lens)
            continue
        if rc == -3 and allow_failures:
            rc = 0
        _lib.check(rc, "lg_prove")
        _proof_cap[key] = max(cap if max(lens) == 0 else max(lens), 1)
        break
    if phases is not None:
        arr = (ctypes.c_double * 3)()
        _lib.check(_lib.lib().lg_plan_last_phases(ph, arr), "lg_plan_last_phases")
        phases[:] = list(arr)
    res = [outs[b][:lens[b]] if lens[b] else None for b in range(B)]
    if want_transcript:
        return res, [touts[b][:tlens[b]].copy() if lens[b] else None for b in range(B)]
    return res


TUNING_FIELDS = ["lanes", "merkle_wide_min", "hash_threads", "fri_coop_pairs", "claims_rows"]


def autotune(ctx, items, dev_inputs=None, reps=3):
    """Runs the library's autotuner on these prepared instances (one plan);
    returns {"tuning": {field: value}, "trials": n, "start_ms": ..., "best_ms": ...}.
    The choice applies to later proves of this batch size on this context."""
    words = items[0].words
    ph = ctx.plan_handle(words, items[0].setup)
    col_tab, perm_tab = _pointer_tables(items, dev_inputs)
    col_ptrs, perm_ptrs = col_tab.ptr, perm_tab.ptr
    t = (ctypes.c_int32 * 6)()
    ms = (ctypes.c_double * 2)()
    _lib.check(_lib.lib().lg_plan_autotune(ph, len(items), col_ptrs, perm_ptrs, 0 if dev_inputs is None else 1, reps,
                                           t, ms), "lg_plan_autotune")
    return {"tuning": dict(zip(TUNING_FIELDS, list(t)[:5])), "trials": t[5], "start_ms": ms[0], "best_ms": ms[1]}


VERIFY_CHECKS = {0: "ok", 1: "format", 2: "logup bus", 3: "AIR identity", 4: "opening", 5: "DEEP", 6: "FRI"}


def verify(ctx, words, proofs, setup_roots=(), publics=None, skip_fri=False, fs_prev=None):
    """lg_verify: checks proofs of the plan `words` in one batch; returns
    [(status, message)] per proof (status 0: valid; VERIFY_CHECKS names the
    others).  setup_roots: the verification key, canonical 8-value roots;
    publics: per proof, the public-input value arrays in plan order
    (Montgomery); fs_prev: per proof the running transcript's last digest
    (lg_verify_chained), None for fresh transcripts."""
    ph = ctx.plan_handle(words)
    arrs = [np.ascontiguousarray(p, dtype=np.uint32) for p in proofs]
    n = len(arrs)
    ptrs = (_lib.u32p * max(1, n))(*[a.ctypes.data_as(_lib.u32p) for a in arrs])
    lens = (ctypes.c_size_t * max(1, n))(*[a.size for a in arrs])
    key = np.ascontiguousarray(to_mont(np.asarray(setup_roots, dtype=np.uint64).reshape(-1)), dtype=np.uint32)
    pubs = None
    keep = []
    if publics is not None:
        flat = [np.ascontiguousarray(v, dtype=np.uint32) for pp in publics for v in pp]
        keep = flat
        pubs = (_lib.u32p * max(1, len(flat)))(*[v.ctypes.data_as(_lib.u32p) for v in flat])
    st = (ctypes.c_int32 * max(1, n))()
    keyp = key.ctypes.data_as(_lib.u32p) if key.size else None
    if fs_prev is None:
        _lib.check(_lib.lib().lg_verify(ph, n, ptrs, lens, keyp, key.size // 8, pubs, 1 if skip_fri else 0, st),
                   "lg_verify")
    else:
        prev = np.ascontiguousarray(to_mont(np.asarray(fs_prev, dtype=np.uint64).reshape(-1)), dtype=np.uint32)
        if prev.size != 8 * n:
            raise ValueError("fs_prev: 8 values per proof")
        _lib.check(_lib.lib().lg_verify_chained(ph, n, ptrs, lens, keyp, key.size // 8, pubs, 1 if skip_fri else 0,
                                                prev.ctypes.data_as(_lib.u32p), st), "lg_verify_chained")
    out = []
    for i in range(n):
        msg = ""
        if st[i]:
            ln = ctypes.c_size_t()
            buf = ctypes.create_string_buffer(512)
            _lib.check(_lib.lib().lg_plan_verify_message(ph, i, buf, 512, ctypes.byref(ln)), "lg_plan_verify_message")
            msg = buf.value.decode()
        out.append((int(st[i]), msg))
    del keep
    return out


def verify_prepared(ctx, items, proofs, setup_roots=(), skip_fri=False, fs_prev=None):
    """verify() for the proofs of these prepared instances (one plan): the
    public-input values are taken from the instances."""
    npub = len(items[0].plan.public)
    publics = [it.columns[len(it.columns) - npub:] for it in items] if npub else None
    return verify(ctx, items[0].words, proofs, setup_roots, publics, skip_fri, fs_prev)


def graph_stats(ctx, items):
    """{"bytes", "records", "replays", "fallbacks"} of the plan of these
    prepared instances (lg_plan_graph_stats)."""
    ph = ctx.plan_handle(items[0].words, items[0].setup)
    v = (ctypes.c_uint64 * 4)()
    _lib.check(_lib.lib().lg_plan_graph_stats(ph, v), "lg_plan_graph_stats")
    return dict(zip(("bytes", "records", "replays", "fallbacks"), list(v)))


def set_options(ctx, items, skip_fri=False):
    """Prove options of the plan of these prepared instances
    (lg_plan_set_options): skip_fri is prover.SkipFRI."""
    ph = ctx.plan_handle(items[0].words, items[0].setup)
    _lib.check(_lib.lib().lg_plan_set_options(ph, 1 if skip_fri else 0), "lg_plan_set_options")


def set_grinding(ctx, items, bits):
    """FRI proof-of-work bits of the plan of these prepared instances
    (lg_plan_set_grinding; prover.WithFriOption(fri.WithGrinding(bits))):
    later proves and verifications of the plan use them; 0 turns it off."""
    ph = ctx.plan_handle(items[0].words, items[0].setup)
    _lib.check(_lib.lib().lg_plan_set_grinding(ph, int(bits)), "lg_plan_set_grinding")


def set_tuning(ctx, items, tuning):
    ph = ctx.plan_handle(items[0].words, items[0].setup)
    v = (ctypes.c_int32 * 5)(*[int(tuning[f]) for f in TUNING_FIELDS])
    _lib.check(_lib.lib().lg_plan_set_tuning(ph, len(items), v), "lg_plan_set_tuning")


def prove(ctx, program, trace, num_queries=PR.NUM_QUERIES_DEFAULT, want_transcript=False):
    r = prove_prepared(ctx, [Prepared(program, trace, num_queries)], want_transcript)
    if want_transcript:
        return r[0][0], r[1][0]
    return r[0]


def prove_batch(ctx, instances, num_queries=PR.NUM_QUERIES_DEFAULT, want_transcript=False):
    return prove_prepared(ctx, [Prepared(p, t, num_queries) for p, t in instances], want_transcript)
