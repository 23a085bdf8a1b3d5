#!/usr/bin/env python3
"""Emit poseidon2_constants.h: KoalaBear Poseidon2 round keys and internal
diagonals in Montgomery form, for widths 16 and 24.

Round keys follow gnark-crypto's ``poseidon2.NewParameters(width, 6, 21)``
(called at internal/hash/poseidon2.go:52-53 of the reference): a legacy
Keccak-256 chain seeded by ``"Poseidon2-koalabear[t=W,rF=6,rP=21,d=3]"``, the
first output is the hash of the seed, each key is the hash of the previous
output read big-endian mod p.  Full rounds take `width` keys, partial rounds
one.  The width-16 table reproduces the constants the reference ships in
integration_test/zkc/riscv/zkc/lib/poseidon2/constants.zkc:34-62.  The internal
diagonal follows recursion/poseidon2.go:48-90.

Run: python3 loom_amd/csrc/gen_poseidon2_constants.py > loom_amd/csrc/poseidon2_constants.h
"""
import sys

P = 2**31 - 2**24 + 1
R = (1 << 32) % P

_RC = [0x0000000000000001, 0x0000000000008082, 0x800000000000808A, 0x8000000080008000,
       0x000000000000808B, 0x0000000080000001, 0x8000000080008081, 0x8000000000008009,
       0x000000000000008A, 0x0000000000000088, 0x0000000080008009, 0x000000008000000A,
       0x000000008000808B, 0x800000000000008B, 0x8000000000008089, 0x8000000000008003,
       0x8000000000008002, 0x8000000000000080, 0x000000000000800A, 0x800000008000000A,
       0x8000000080008081, 0x8000000000008080, 0x0000000080000001, 0x8000000080008008]
_PI_ROT = {(0, 0): 0, (0, 1): 36, (0, 2): 3, (0, 3): 41, (0, 4): 18,
           (1, 0): 1, (1, 1): 44, (1, 2): 10, (1, 3): 45, (1, 4): 2,
           (2, 0): 62, (2, 1): 6, (2, 2): 43, (2, 3): 15, (2, 4): 61,
           (3, 0): 28, (3, 1): 55, (3, 2): 25, (3, 3): 21, (3, 4): 56,
           (4, 0): 27, (4, 1): 20, (4, 2): 39, (4, 3): 8, (4, 4): 14}
M = (1 << 64) - 1


def rotl(v, n):
    return ((v << n) | (v >> (64 - n))) & M if n % 64 else v


def permute(lanes):
    for rc in _RC:
        col = [lanes[(x, 0)] ^ lanes[(x, 1)] ^ lanes[(x, 2)] ^ lanes[(x, 3)] ^ lanes[(x, 4)]
               for x in range(5)]
        for x in range(5):
            dx = col[(x + 4) % 5] ^ rotl(col[(x + 1) % 5], 1)
            for y in range(5):
                lanes[(x, y)] ^= dx
        moved = {}
        for (x, y), v in lanes.items():
            moved[(y, (2 * x + 3 * y) % 5)] = rotl(v, _PI_ROT[(x, y)])
        for x in range(5):
            for y in range(5):
                lanes[(x, y)] = moved[(x, y)] ^ (~moved[((x + 1) % 5, y)] & moved[((x + 2) % 5, y)] & M)
        lanes[(0, 0)] ^= rc
    return lanes


def keccak256_legacy(data):
    rate = 136
    buf = bytearray(data)
    buf.append(0x01)
    buf.extend(b"\x00" * ((-len(buf)) % rate))
    buf[-1] |= 0x80
    lanes = {(x, y): 0 for x in range(5) for y in range(5)}
    for off in range(0, len(buf), rate):
        for i in range(rate // 8):
            lanes[(i % 5, i // 5)] ^= int.from_bytes(buf[off + 8 * i: off + 8 * i + 8], "little")
        lanes = permute(lanes)
    return b"".join(lanes[(i % 5, i // 5)].to_bytes(8, "little") for i in range(4))


def round_keys(width, rf=6, rp=21):
    state = keccak256_legacy(("Poseidon2-koalabear[t=%d,rF=%d,rP=%d,d=3]" % (width, rf, rp)).encode())
    keys = []

    def draw():
        nonlocal state
        state = keccak256_legacy(state)
        return int.from_bytes(state, "big") % P

    for r in range(rf + rp):
        full = r < rf // 2 or r >= rf // 2 + rp
        keys.append([draw() for _ in range(width)] if full else [draw()])
    return keys


def diag(width):
    tails = {16: ([8, 3, 24], [8, 3, 4, 24]),
             24: ([8, 2, 3, 4, 5, 6, 24], [8, 3, 4, 5, 6, 7, 9, 24])}[width]
    half = pow(2, P - 2, P)
    d = [P - 2, 1, 2, half, 3, 4, P - half, P - 3, P - 4]
    d += [pow(1 << k, P - 2, P) for k in tails[0]]
    d += [P - pow(1 << k, P - 2, P) for k in tails[1]]
    return d


def mont(x):
    return x * R % P


def emit(out):
    out.write("// Generated by gen_poseidon2_constants.py -- do not edit.\n")
    out.write("// KoalaBear Poseidon2 (x^3, 6 full + 21 partial rounds) constants in\n")
    out.write("// Montgomery form (R = 2^32).  *_CANON arrays hold the canonical values.\n")
    out.write("#pragma once\n#include <cstdint>\n\nnamespace lg {\n")
    for w in (16, 24):
        ks = round_keys(w)
        full = [row for row in ks if len(row) == w]
        part = [row[0] for row in ks if len(row) == 1]
        for tag, conv in (("", mont), ("_CANON", lambda v: v)):
            out.write("constexpr uint32_t P2_W%d_FULL_RK%s[6][%d] = {\n" % (w, tag, w))
            for row in full:
                out.write("  {" + ", ".join("0x%08xu" % conv(v) for v in row) + "},\n")
            out.write("};\n")
            out.write("constexpr uint32_t P2_W%d_PARTIAL_RK%s[21] = {" % (w, tag))
            out.write(", ".join("0x%08xu" % conv(v) for v in part) + "};\n")
            out.write("constexpr uint32_t P2_W%d_DIAG%s[%d] = {" % (w, tag, w))
            out.write(", ".join("0x%08xu" % conv(v) for v in diag(w)) + "};\n")
    out.write("}  // namespace lg\n")


if __name__ == "__main__":
    emit(sys.stdout)
