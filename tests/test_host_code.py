"""The device field/Poseidon2 headers built for the host with g++ (CPU):
specialised divisions by 2^k equal Montgomery products, and the width-24
permutation (its host form: lazy 64-bit partial rounds) equals the oracle's
on random and extreme states; the host transcript equals the oracle's."""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "loom_amd", "csrc")


@pytest.fixture(scope="module")
def p2_bin():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "p2_host_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(HERE, "host", "p2_host_check.cpp"),
                    "-o", exe], check=True)
    yield exe
    shutil.rmtree(d)


def test_div_pow2_and_permutation_match_oracle(p2_bin):
    """Both forms of the partial rounds (the host's lazy signed 64-bit lanes and
    the device's scaled lanes) give the oracle's permutation."""
    from oracle import poseidon2 as p2
    rng = np.random.default_rng(5)
    P = 2**31 - 2**24 + 1
    # the host permutation carries the partial rounds in lazy 64-bit form:
    # extreme and alternating states as well as random ones
    states = rng.integers(0, P, size=(200, 24), dtype=np.uint64)
    states[0] = 0
    states[1] = P - 1
    states[2, 0::2] = P - 1
    states[2, 1::2] = 0
    states[3, 0::2] = 0
    states[3, 1::2] = P - 1
    states[4] = np.arange(24) * (P // 24)
    inp = "%d\n" % len(states) + "\n".join(" ".join(str(int(v)) for v in st) for st in states) + "\n"
    r = subprocess.run([p2_bin], input=inp, capture_output=True, text=True)
    lines = r.stdout.strip().splitlines()
    assert lines[0] == "div_pow2_mismatches 0", lines[0]
    assert r.returncode == 0
    got = np.array([[int(x) for x in ln.split()] for ln in lines[1:]], dtype=np.uint64)
    perm = p2.Permutation(24)
    want = perm(states.T.copy()).T
    # host form of the partial rounds, then the device form (scaled lanes,
    # one Montgomery reduction per lane and round)
    assert np.array_equal(got[0::2], want)
    assert np.array_equal(got[1::2], want)


@pytest.fixture(scope="module")
def ntt_bin():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    exe = os.path.join(d, "ntt16_host_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(HERE, "host", "ntt16_host_check.cpp"),
                    "-o", exe], check=True)
    yield exe
    shutil.rmtree(d)


@pytest.mark.parametrize("logn", [8, 9, 10, 11, 12, 13, 14])
def test_radix16_passes_match_dft(ntt_bin, logn):
    """The radix-16 pass schedule of the LDS NTT: forward output equals the
    DFT in bit-reversed order, and the inverse passes bring it back (times n)."""
    from oracle import ntt, field as F
    P = F.P
    n = 1 << logn
    rng = np.random.default_rng(logn)
    x = rng.integers(0, P, size=n, dtype=np.uint64)
    r = subprocess.run([ntt_bin], input="%d\n" % logn + " ".join(str(int(v)) for v in x) + "\n",
                       capture_output=True, text=True, check=True)
    fwd, inv = [np.array([int(v) for v in ln.split()], dtype=np.uint64) for ln in r.stdout.strip().splitlines()]
    want = ntt.evaluate(x)
    assert np.array_equal(fwd, want[F.bitrev_indices(n)])
    assert np.array_equal(inv, x * np.uint64(n) % np.uint64(P))


def test_montgomery_product_matches_closed_form():
    """kb_mul (one 64-bit multiply-add Montgomery reduction) equals
    a*b*2^-32 mod p computed in 128-bit integers, and the host S-box cube
    equals a^3*2^-64 mod p."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    try:
        exe = os.path.join(d, "kb_mul_check")
        subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(HERE, "host", "kb_mul_check.cpp"),
                        "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.stdout.split() == ["kb_mul_mismatches", "0", "kb_cube_mismatches", "0"], r.stdout
        assert r.returncode == 0
    finally:
        shutil.rmtree(d)


def test_e6_product_matches_oracle():
    """e6_mul / e2_mul of kb.h (Karatsuba-style cross terms) equal the
    oracle's schoolbook E6 product on random operands and edge values."""
    from oracle import field as F
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    try:
        exe = os.path.join(d, "e6_mul_check")
        subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(HERE, "host", "e6_mul_check.cpp"),
                        "-o", exe], check=True)
        rng = np.random.default_rng(11)
        P = F.P
        x = rng.integers(0, P, size=(400, 6), dtype=np.uint64)
        y = rng.integers(0, P, size=(400, 6), dtype=np.uint64)
        x[0] = P - 1
        y[0] = P - 1
        x[1] = 0
        y[2, :] = [1, 0, 0, 0, 0, 0]
        inp = "%d\n" % len(x) + "\n".join(" ".join(str(int(v)) for v in np.concatenate([a, b]))
                                          for a, b in zip(x, y)) + "\n"
        r = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True)
        got = np.array([[int(v) for v in ln.split()] for ln in r.stdout.strip().splitlines()], dtype=np.uint64)
        assert np.array_equal(got, F.e6_mul(x, y))
    finally:
        shutil.rmtree(d)


def test_transcript_forms_match_oracle():
    """transcript.h over a shared challenge schedule (by name and by
    position), with names registered one by one, and with accepted digests
    gives the oracle transcript's digests for the prove path's challenge
    names and random bindings."""
    from oracle.transcript import Transcript as OT
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    rng = np.random.default_rng(17)
    P = 2**31 - 2**24 + 1
    names = ["challenge@loom_0", "challenge@loom_1", "__zeta", "alpha_DEEP", "fri_level_1_gamma"]
    names += ["fri_fold_%d" % j for j in range(3)] + ["fri_query_%d" % k for k in range(12)]
    binds = [[int(v) for v in rng.integers(0, P, size=int(rng.integers(0, 40)))] for _ in names]
    binds[0] = [P - 1] * 17
    binds[2] = []
    d = tempfile.mkdtemp()
    try:
        exe = os.path.join(d, "transcript_check")
        subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(HERE, "host", "transcript_check.cpp"),
                        "-o", exe], check=True)
        inp = "%d\n" % len(names) + "".join("%s %d %s\n" % (n, len(b), " ".join(map(str, b)))
                                            for n, b in zip(names, binds))
        r = subprocess.run([exe], input=inp, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
        got = [[int(x) for x in ln.split()] for ln in r.stdout.strip().splitlines()]
    finally:
        shutil.rmtree(d)
    t = OT(names)
    want = []
    for n, b in zip(names, binds):
        t.bind(n, b)
        want.append([int(v) for v in t.compute(n)])
    assert got == want



def test_lds_index_decomposition():
    """The radix-16 NTT kernel computes each LDS index of a pass as
    lds_swz(base) ^ lds_swz(k << lq) (and the scatter and row indices
    likewise); the host check compares that with lds_swz of the full index for
    every size, pass, thread and value, and checks the map is a bijection."""
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    d = tempfile.mkdtemp()
    try:
        exe = os.path.join(d, "ntt16_swizzle_check")
        subprocess.run([gxx, "-O2", "-std=c++17", "-I" + CSRC, os.path.join(HERE, "host", "ntt16_swizzle_check.cpp"),
                        "-o", exe], check=True)
        r = subprocess.run([exe], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout
        assert r.stdout.split()[-1] == "0"
    finally:
        shutil.rmtree(d)
