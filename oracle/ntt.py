"""Number-theoretic transforms and the RS encoder (numpy).

Follows internal/reedsolomon/rs.go:52-87: a column in Lagrange form on the
size-n subgroup is interpolated (iFFT), zero-padded to rate*n coefficients
and evaluated on the size-rate*n subgroup; the output is in bit-reversed
evaluation order, row r = P(w_{rn}^{bitrev(r)}).  The transforms here are
written as plain matrix-free radix-2 evaluations in natural order; the order
conventions are applied explicitly.  Pinned by a naive O(n^2) DFT and by
Horner evaluation of the interpolant in the tests (the rs_eval_test.go
pattern).  The values depend on the subgroup generator, which is "parity
unpinned" (see field.py).
"""
import numpy as np

from .field import P, U64, omega, powers, bitrev_indices, inv


def _ntt_natural(a, w):
    """Evaluate coefficient vectors (natural order, last axis) at w^0..w^(n-1),
    natural order output.  Iterative radix-2 Cooley-Tukey on the last axis."""
    a = np.asarray(a, dtype=np.uint64)
    n = a.shape[-1]
    if n == 1:
        return a.copy()
    rev = bitrev_indices(n)
    x = a[..., rev].copy()
    m = 1
    while m < n:
        wm = pow(int(w), n // (2 * m), P)
        tw = powers(wm, m)
        x = x.reshape(a.shape[:-1] + (n // (2 * m), 2, m))
        u = x[..., 0, :].copy()
        v = x[..., 1, :] * tw % U64(P)
        x[..., 0, :] = (u + v) % U64(P)
        x[..., 1, :] = (u + U64(P) - v) % U64(P)
        x = x.reshape(a.shape)
        m *= 2
    return x


def evaluate(coeffs, n_or_w=None):
    """Coefficients (natural) -> evaluations on <w_n>, natural order."""
    n = coeffs.shape[-1]
    return _ntt_natural(coeffs, omega(n))


def interpolate(values):
    """Evaluations on <w_n> (natural order) -> coefficients (natural)."""
    n = values.shape[-1]
    c = _ntt_natural(values, inv(omega(n)))
    return c * U64(inv(n)) % U64(P)


def encode(values, rate=4):
    """RS encode (rs.go:52-70) along the last axis: Lagrange(n) -> bit-reversed
    evaluations on the size rate*n subgroup."""
    values = np.asarray(values, dtype=np.uint64)
    n = values.shape[-1]
    c = interpolate(values)
    pad = np.zeros(values.shape[:-1] + (rate * n,), dtype=np.uint64)
    pad[..., :n] = c
    ev = evaluate(pad)
    return ev[..., bitrev_indices(rate * n)]


def encode_ext(values, rate=4):
    """values: (n, 6) E6 column -> (rate*n, 6)."""
    return encode(np.asarray(values).T, rate).T.copy()


def coset_evaluate(coeffs, shift, m):
    """Evaluate coefficients (natural, length <= m) at shift*w_m^k, natural k."""
    coeffs = np.asarray(coeffs, dtype=np.uint64)
    k = coeffs.shape[-1]
    pad = np.zeros(coeffs.shape[:-1] + (m,), dtype=np.uint64)
    pad[..., :k] = coeffs * powers(shift, k) % U64(P)
    return evaluate(pad)


def horner(coeffs, x):
    """Evaluate one coefficient vector at a base-field point (pure Python)."""
    r = 0
    for c in reversed([int(v) for v in coeffs]):
        r = (r * x + c) % P
    return r
