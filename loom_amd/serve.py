"""Serving loop over the HIP prover (lg_server_*, include/loomgpu.h).

    srv = Server(device=0, slots=3)
    t = srv.submit([Prepared(program, trace, 200), ...])   # one request
    proofs = srv.wait(t)                                    # proof words per instance

Requests are admitted by HBM footprint onto concurrent proving slots, and
queued requests of one plan are batched together when a slot takes them;
each proof is the one ``prover.prove`` gives for that instance alone.
"""
import ctypes

import numpy as np

from . import _lib

STAT_NAMES = ["submitted", "completed", "failed", "queued", "running", "peak_running", "peak_bytes", "proves",
              "merged", "budget_bytes"]


def plan_bytes_per_instance(words):
    w = np.ascontiguousarray(words, dtype=np.uint32)
    v = ctypes.c_uint64()
    _lib.check(_lib.lib().lg_plan_bytes_per_instance(w.ctypes.data_as(_lib.u32p), w.size, ctypes.byref(v)),
               "lg_plan_bytes_per_instance")
    return v.value


class Server:
    def __init__(self, device=0, slots=2, budget_bytes=0, max_batch=64):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lg_server_create(device, slots, budget_bytes, max_batch, ctypes.byref(h)),
                   "lg_server_create")
        self.h = h
        self._batch = {}

    def close(self):
        if self.h:
            _lib.lib().lg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, items):
        """items: prover.Prepared instances of one plan; returns a ticket."""
        words = items[0].words
        for it in items[1:]:
            if not np.array_equal(it.words, words):
                raise ValueError("a request's instances must share one plan")
        B = len(items)
        nc, npm = len(items[0].columns), len(items[0].perms)
        cols = (_lib.u32p * max(1, B * nc))()
        perms = (_lib.u32p * max(1, B * npm))()
        for b, it in enumerate(items):
            for k, c in enumerate(it.columns):
                cols[b * nc + k] = c.ctypes.data_as(_lib.u32p)
            for k, s in enumerate(it.perms):
                perms[b * npm + k] = s.ctypes.data_as(_lib.u32p)
        w = np.ascontiguousarray(words, dtype=np.uint32)
        t = ctypes.c_uint64()
        _lib.check(_lib.lib().lg_server_submit(self.h, w.ctypes.data_as(_lib.u32p), w.size, B, cols, perms,
                                               ctypes.byref(t)), "lg_server_submit")
        self._batch[t.value] = B
        return t.value

    def wait(self, ticket, timeout_ms=-1):
        """Proof words of each instance of the request, or None on timeout."""
        B = self._batch[ticket]
        cap = 1 << 20
        while True:
            outs = [np.empty(cap, dtype=np.uint32) for _ in range(B)]
            ptrs = (_lib.u32p * B)(*[o.ctypes.data_as(_lib.u32p) for o in outs])
            caps = (ctypes.c_size_t * B)(*([cap] * B))
            lens = (ctypes.c_size_t * B)()
            rc = _lib.lib().lg_server_wait(self.h, ticket, timeout_ms, ptrs, caps, lens)
            if rc == 1:
                return None
            if rc == -2:
                cap = max(lens)
                continue
            _lib.check(rc, "lg_server_wait")
            del self._batch[ticket]
            return [outs[b][:lens[b]] for b in range(B)]

    def stats(self):
        v = (ctypes.c_uint64 * len(STAT_NAMES))()
        _lib.check(_lib.lib().lg_server_stats(self.h, v, len(STAT_NAMES)), "lg_server_stats")
        return dict(zip(STAT_NAMES, [int(x) for x in v]))
