"""Fiat-Shamir transcript (pure Python around the numpy sponge).

Follows internal/fiat-shamir/transcript.go:102-315: challenges are registered
in order; ``bind`` appends elements to a challenge; ``compute`` hashes
StringToElements(FSID, name) || previous challenge's 8-element value (when the
challenge is not the first) || the bindings, with the Poseidon2 sponge
(grinding off, the default).  StringToElements is internal/hash/hash.go:36-53
(tag, byte length, then little-endian 3-byte chunks).  Pinned by a recording
test on exact element sequences (the transcript_test.go:13-78 pattern).
"""
from .poseidon2 import sponge1

FSID_TAG = 0x46534944


def string_to_elements(tag, s):
    b = s.encode()
    out = [tag, len(b)]
    for i in range(0, len(b), 3):
        limb = 0
        for j in range(min(3, len(b) - i)):
            limb |= b[i + j] << (8 * j)
        out.append(limb)
    return out


class Transcript:
    def __init__(self, names=(), hasher=sponge1):
        self.order = []
        self.pos = {}
        self.bindings = {}
        self.values = {}
        self.hasher = hasher
        for n in names:
            self.new_challenge(n)

    def new_challenge(self, name):
        if name in self.pos:
            raise ValueError("challenge already exists: " + name)
        self.pos[name] = len(self.order)
        self.order.append(name)
        self.bindings[name] = []

    def bind(self, name, elems):
        if name not in self.pos:
            raise KeyError(name)
        if name in self.values:
            raise ValueError("challenge already computed: " + name)
        self.bindings[name].extend(int(e) for e in elems)

    def message(self, name):
        """The exact element sequence ``compute`` hashes."""
        p = self.pos[name]
        msg = string_to_elements(FSID_TAG, name)
        if p != 0:
            prev = self.order[p - 1]
            if prev not in self.values:
                raise ValueError("previous challenge not computed")
            msg += self.values[prev]
        msg += self.bindings[name]
        return msg

    def compute(self, name):
        if name in self.values:
            return self.values[name]
        v = self.hasher(self.message(name))
        self.values[name] = v
        return v
