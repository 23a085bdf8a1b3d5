"""Poseidon2 over KoalaBear, the loom sponge, leaf and node hashers (numpy).

Follows:
* permutation structure -- gnark-crypto ``poseidon2.Permutation`` as restated
  in-tree by recursion/p2_air.go:78-117 and by the zkc port
  integration_test/zkc/riscv/zkc/lib/poseidon2/poseidon2.zkc:33-176:
  M_E first, 3 full rounds, 21 partial rounds, 3 full rounds, S-box x^3,
  M_E = circ(2*M4, M4, ...) (M4 applied per 4-lane chunk then the chunk sums
  added), M_I = 1*1^T + diag(d);
* internal diagonal -- recursion/poseidon2.go:48-90 (``p2Diag``);
* round keys -- gnark ``NewParameters(width, 6, 21)``: a legacy Keccak-256
  chain seeded by "Poseidon2-koalabear[t=W,rF=6,rP=21,d=3]", each 32-byte
  output read big-endian and reduced mod p.  Pinned: the derivation
  reproduces every width-16 key of the zkc table
  (integration_test/zkc/riscv/zkc/lib/poseidon2/constants.zkc:34-62) and
  ``internal_diag`` there (constants.zkc:77-80) equals ``p2_diag(16)``;
* sponge -- internal/hash/poseidon2.go:165-255 (width 24, rate 16, overwrite
  mode; a partial last block overwrites only its own lanes; no padding; empty
  input gives the zero digest; digest = state[0:8]);
* leaf hash -- internal/fri/hash.go:39-56 ([LEAF, 2nb, 2ne, lo.base, hi.base,
  lo.ext limbs, hi.ext limbs]);
* node compress -- internal/hash/poseidon2.go:409-422
  (P([NODE, 0*7, L, R])[0:8]).

The width-24 permutation itself has no in-tree known-answer vector: its
structure is pinned through the width-16 path (same code, width 16 vs the zkc
table) and its constants through the derivation above.
"""
import numpy as np

from .field import P, U64

NB_FULL = 6
NB_PARTIAL = 21
WIDTH = 24
RATE = 16
DIGEST = 8
LEAF_TAG = 0x4C454146
NODE_TAG = 0x4E4F4445

# ------------------------------------------------------------ legacy keccak
_RC = [0x0000000000000001, 0x0000000000008082, 0x800000000000808A, 0x8000000080008000,
       0x000000000000808B, 0x0000000080000001, 0x8000000080008081, 0x8000000000008009,
       0x000000000000008A, 0x0000000000000088, 0x0000000080008009, 0x000000008000000A,
       0x000000008000808B, 0x800000000000008B, 0x8000000000008089, 0x8000000000008003,
       0x8000000000008002, 0x8000000000000080, 0x000000000000800A, 0x800000008000000A,
       0x8000000080008081, 0x8000000000008080, 0x0000000080000001, 0x8000000080008008]
_ROT = [[0, 36, 3, 41, 18], [1, 44, 10, 45, 2], [62, 6, 43, 15, 61],
        [28, 55, 25, 21, 56], [27, 20, 39, 8, 14]]
_M64 = (1 << 64) - 1


def _rol(x, n):
    n %= 64
    return ((x << n) | (x >> (64 - n))) & _M64 if n else x


def _keccak_f(a):
    for rc in _RC:
        c = [a[x][0] ^ a[x][1] ^ a[x][2] ^ a[x][3] ^ a[x][4] for x in range(5)]
        d = [c[(x - 1) % 5] ^ _rol(c[(x + 1) % 5], 1) for x in range(5)]
        a = [[a[x][y] ^ d[x] for y in range(5)] for x in range(5)]
        b = [[0] * 5 for _ in range(5)]
        for x in range(5):
            for y in range(5):
                b[y][(2 * x + 3 * y) % 5] = _rol(a[x][y], _ROT[x][y])
        a = [[b[x][y] ^ ((~b[(x + 1) % 5][y]) & b[(x + 2) % 5][y]) for y in range(5)]
             for x in range(5)]
        a[0][0] ^= rc
    return a


def keccak256(data, pad=0x01):
    """Keccak-256; pad=0x01 is the legacy (pre-SHA3) padding gnark uses,
    pad=0x06 gives SHA3-256 (used by the tests to check this routine against
    hashlib)."""
    rate = 136
    msg = bytearray(data) + bytes([pad])
    while len(msg) % rate:
        msg.append(0)
    msg[-1] |= 0x80
    a = [[0] * 5 for _ in range(5)]
    for off in range(0, len(msg), rate):
        blk = msg[off:off + rate]
        for i in range(rate // 8):
            a[i % 5][i // 5] ^= int.from_bytes(blk[8 * i:8 * i + 8], 'little')
        a = _keccak_f(a)
    return b''.join(a[i % 5][i // 5].to_bytes(8, 'little') for i in range(4))


def round_keys(width):
    """Round keys of gnark ``poseidon2.NewParameters(width, 6, 21)`` as a list of
    27 rows (full rows have `width` keys, partial rows one)."""
    seed = "Poseidon2-koalabear[t=%d,rF=%d,rP=%d,d=3]" % (width, NB_FULL, NB_PARTIAL)
    cur = keccak256(seed.encode())

    def nxt():
        nonlocal cur
        cur = keccak256(cur)
        return int.from_bytes(cur, 'big') % P

    rows = []
    for _ in range(NB_FULL // 2):
        rows.append([nxt() for _ in range(width)])
    for _ in range(NB_PARTIAL):
        rows.append([nxt()])
    for _ in range(NB_FULL // 2):
        rows.append([nxt() for _ in range(width)])
    return rows


def p2_diag(width):
    """Internal-matrix diagonal (recursion/poseidon2.go:48-90)."""
    if width == 16:
        pos, neg = [8, 3, 24], [8, 3, 4, 24]
    elif width == 24:
        pos, neg = [8, 2, 3, 4, 5, 6, 24], [8, 3, 4, 5, 6, 7, 9, 24]
    else:
        raise ValueError(width)
    inv2k = lambda k: pow(1 << k, P - 2, P)
    d = [P - 2, 1, 2, inv2k(1), 3, 4, P - inv2k(1), P - 3, P - 4]
    d += [inv2k(k) for k in pos]
    d += [(P - inv2k(k)) % P for k in neg]
    assert len(d) == width
    return d


class Permutation:
    """Vectorised Poseidon2: state is an array of shape (width, batch)."""

    def __init__(self, width=WIDTH):
        self.width = width
        self.rk = round_keys(width)
        self.diag = np.array(p2_diag(width), dtype=np.uint64)

    def _ext(self, s):
        w = self.width
        out = np.empty_like(s)
        for c in range(w // 4):
            x0, x1, x2, x3 = s[4 * c], s[4 * c + 1], s[4 * c + 2], s[4 * c + 3]
            # M4 = [[2,3,1,1],[1,2,3,1],[1,1,2,3],[3,1,1,2]]
            out[4 * c] = (2 * x0 + 3 * x1 + x2 + x3) % U64(P)
            out[4 * c + 1] = (x0 + 2 * x1 + 3 * x2 + x3) % U64(P)
            out[4 * c + 2] = (x0 + x1 + 2 * x2 + 3 * x3) % U64(P)
            out[4 * c + 3] = (3 * x0 + x1 + x2 + 2 * x3) % U64(P)
        sums = [out[j::4].sum(axis=0) % U64(P) for j in range(4)]
        for i in range(w):
            out[i] = (out[i] + sums[i % 4]) % U64(P)
        return out

    def _int(self, s):
        tot = s.sum(axis=0) % U64(P)
        return (tot[None, :] + self.diag[:, None] * s) % U64(P)

    def __call__(self, state):
        s = np.asarray(state, dtype=np.uint64).copy()
        squeeze = s.ndim == 1
        if squeeze:
            s = s[:, None]
        s = self._ext(s)
        h = NB_FULL // 2
        for r in range(h):
            s = (s + np.array(self.rk[r], dtype=np.uint64)[:, None]) % U64(P)
            s = s * s % U64(P) * s % U64(P)
            s = self._ext(s)
        for r in range(h, h + NB_PARTIAL):
            s[0] = (s[0] + U64(self.rk[r][0])) % U64(P)
            s[0] = s[0] * s[0] % U64(P) * s[0] % U64(P)
            s = self._int(s)
        for r in range(h + NB_PARTIAL, NB_FULL + NB_PARTIAL):
            s = (s + np.array(self.rk[r], dtype=np.uint64)[:, None]) % U64(P)
            s = s * s % U64(P) * s % U64(P)
            s = self._ext(s)
        return s[:, 0] if squeeze else s


_PERM = None


def perm24():
    global _PERM
    if _PERM is None:
        _PERM = Permutation(WIDTH)
    return _PERM


def sponge(msgs):
    """Width-24 rate-16 overwrite sponge over a batch of equal-length messages.

    msgs: array (L, B) of canonical elements (B independent messages of length
    L).  Returns digests (8, B).  Follows Poseidon2SpongeHasher
    (internal/hash/poseidon2.go:165-255)."""
    msgs = np.asarray(msgs, dtype=np.uint64)
    L, B = msgs.shape
    if L == 0:
        return np.zeros((DIGEST, B), dtype=np.uint64)
    st = np.zeros((WIDTH, B), dtype=np.uint64)
    perm = perm24()
    for off in range(0, L, RATE):
        blk = msgs[off:off + RATE]
        st[:blk.shape[0]] = blk
        st = perm(st)
    return st[:DIGEST]


def sponge1(msg):
    """Single-message sponge; returns a list of 8 ints."""
    m = np.asarray(msg, dtype=np.uint64).reshape(-1, 1)
    return [int(v) for v in sponge(m)[:, 0]]


def node_compress(left, right):
    """left/right: arrays (8, B) -> (8, B)."""
    B = left.shape[1]
    st = np.zeros((WIDTH, B), dtype=np.uint64)
    st[0] = NODE_TAG
    st[8:16] = left
    st[16:24] = right
    return perm24()(st)[:DIGEST]


def hash_leaf_pairs(base_cols, ext_cols):
    """Leaf digests of an RS-encoded group.

    base_cols: list of arrays (rows,), ext_cols: list of arrays (rows, 6);
    pair leaf k absorbs rows 2k and 2k+1 (internal/fri/hash.go:39-56 and
    commitment.go:512-521).  Returns (8, rows/2)."""
    rows = len(base_cols[0]) if base_cols else len(ext_cols[0])
    nb, ne = len(base_cols), len(ext_cols)
    K = rows // 2
    parts = [np.full((1, K), LEAF_TAG, dtype=np.uint64),
             np.full((1, K), 2 * nb, dtype=np.uint64),
             np.full((1, K), 2 * ne, dtype=np.uint64)]
    for c in base_cols:
        parts.append(np.asarray(c, dtype=np.uint64)[0::2][None, :])
    for c in base_cols:
        parts.append(np.asarray(c, dtype=np.uint64)[1::2][None, :])
    for c in ext_cols:
        parts.append(np.asarray(c, dtype=np.uint64)[0::2].T)
    for c in ext_cols:
        parts.append(np.asarray(c, dtype=np.uint64)[1::2].T)
    return sponge(np.concatenate(parts, axis=0))
