"""Fiat-Shamir transcript (pure Python around the numpy sponge).

Follows internal/fiat-shamir/transcript.go:102-315: challenges are registered
in order; ``bind`` appends elements to a challenge; ``compute`` hashes
StringToElements(FSID, name) || previous challenge's 8-element value (when the
challenge is not the first) || the bindings, with the Poseidon2 sponge; with
grinding (ComputeChallenge WithGrinding, :216-332) the message ends with
[FSPW, bits, salt] for the smallest salt >= 0 whose digest has `bits` zero
low bits.  StringToElements is internal/hash/hash.go:36-53
(tag, byte length, then little-endian 3-byte chunks).  Pinned by a recording
test on exact element sequences (the transcript_test.go:13-78 pattern).
"""
import numpy as np

from .poseidon2 import sponge, sponge1

FSID_TAG = 0x46534944
FSPW_TAG = 0x46535057  # "FSPW": before a proof of work's bit count and salt
KOALABEAR_BITS = 31
P = 2**31 - 2**24 + 1


def has_zero_grinding_bits(digest, bits):
    """hasZeroGrindingBits (transcript.go:349-365): the low `bits` bits of the
    canonical digest elements, 31 per element from element 0, are zero."""
    remaining = bits
    i = 0
    while remaining > 0 and i < 8:
        k = min(remaining, KOALABEAR_BITS)
        if int(digest[i]) & ((1 << k) - 1):
            return False
        remaining -= k
        i += 1
    return True


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
        self.pow = {}  # name -> (bits, salt) of challenges computed with grinding
        self.prev0 = None  # the running transcript's last digest (chain_from)
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

    def chain_from(self, digest):
        """A running transcript (prover/verifier.WithTranscript) whose last
        challenge is computed: this transcript's challenges follow it, so the
        first one's message carries that digest (8 canonical elements)."""
        if any(n in self.values for n in self.order):
            raise ValueError("chained after a computation")
        self.prev0 = [int(x) for x in digest]

    def message(self, name):
        """The exact element sequence ``compute`` hashes."""
        p = self.pos[name]
        msg = string_to_elements(FSID_TAG, name)
        if p != 0:
            prev = self.order[p - 1]
            if prev not in self.values:
                raise ValueError("previous challenge not computed")
            msg += self.values[prev]
        elif self.prev0 is not None:
            msg += self.prev0
        msg += self.bindings[name]
        return msg

    def compute(self, name):
        if name in self.values:
            return self.values[name]
        v = self.hasher(self.message(name))
        self.values[name] = v
        return v

    def compute_grind(self, name, bits, chunk=256):
        """grindChallenge (transcript.go:317-332): the first salt 0, 1, ... whose
        digest passes has_zero_grinding_bits; salts are tried in chunks of
        `chunk` at once (the batched sponge), the smallest passing one wins.
        Records the proof of work (bits, salt); returns the digest."""
        if name in self.values:
            return self.values[name]
        if not 0 < bits <= KOALABEAR_BITS:
            raise ValueError("invalid grinding bit count")
        msg = self.message(name) + [FSPW_TAG, bits]
        for base in range(0, P, chunk):
            salts = np.arange(base, min(base + chunk, P), dtype=np.uint64)
            m = np.repeat(np.asarray(msg, dtype=np.uint64).reshape(-1, 1), len(salts), axis=1)
            dig = sponge(np.vstack([m, salts.reshape(1, -1)]))
            for k in range(len(salts)):
                if has_zero_grinding_bits(dig[:, k], bits):
                    v = [int(x) for x in dig[:, k]]
                    self.values[name] = v
                    self.pow[name] = (bits, int(salts[k]))
                    return v
        raise ValueError("invalid proof of work: no salt below p")

    def compute_with_pow(self, name, bits, salt):
        """The verifier's side (SetProofOfWork + ComputeChallenge WithGrinding):
        the digest with the given salt, which must pass the grinding check."""
        if name in self.values:
            return self.values[name]
        if not 0 < bits <= KOALABEAR_BITS or not 0 <= salt < P:
            raise ValueError("invalid proof of work")
        v = self.hasher(self.message(name) + [FSPW_TAG, bits, salt])
        if not has_zero_grinding_bits(v, bits):
            raise ValueError("invalid proof of work")
        self.values[name] = v
        self.pow[name] = (bits, salt)
        return v
