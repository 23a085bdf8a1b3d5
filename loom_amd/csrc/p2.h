// Poseidon2 over KoalaBear (x^3 S-box, 6 full + 21 partial rounds) on a state
// held in registers, one state per lane.  Structure: external layer, 3 full
// rounds, 21 partial rounds, 3 full rounds (gnark-crypto Permutation, restated
// in recursion/p2_air.go:78-117).  External layer = M4 on each 4-lane chunk
// followed by adding the chunk-wise sums (circ(2*M4, M4, ...)); internal layer
// s_i <- sum(s) + d_i * s_i with the diagonal of recursion/poseidon2.go:48-90.
//
// Also the loom hashing formats built on the width-24 permutation:
//   node  = P([NODE, 0 x 7, L[0..8], R[0..8]])[0..8]   (internal/hash/poseidon2.go:409-422)
//   leaf  = sponge([LEAF, 2nb, 2ne, lo.base, hi.base, lo.ext, hi.ext])   (internal/fri/hash.go:39-56)
//   sponge: rate 16, overwrite mode, partial last block overwrites only its
//   lanes, digest = state[0..8] (internal/hash/poseidon2.go:165-255).
#pragma once
#include <cstddef>

#include "kb.h"
#include "poseidon2_constants.h"

namespace lg {

constexpr uint32_t TAG_LEAF = 0x4c454146u;
constexpr uint32_t TAG_NODE = 0x4e4f4445u;
constexpr uint32_t TAG_LEAF_MONT = 0x6d420179u;  // Montgomery form of TAG_LEAF
constexpr uint32_t TAG_NODE_MONT = 0x6425c2feu;  // Montgomery form of TAG_NODE

template <int W>
struct P2Consts;

// acc + x (a 32-bit term into a 64-bit sum).  On the device it is one
// v_mad_u64_u32 with an opaque multiplier 1 (a plain 64-bit addition of a
// zero-extended value costs a register move and an add); the asm input I
// keeps the multipliers of one sum distinct, so that the compiler neither
// merges them nor re-factors the sum.
template <int I>
LG_HD uint64_t p2_acc(uint32_t x, uint64_t acc) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t one = 1;
  asm("" : "+s"(one) : "i"(I));
  return (uint64_t)x * one + acc;
#else
  return acc + x;
#endif
}
template <int I>
LG_HD uint64_t p2_acc2(uint32_t x, uint64_t acc) {  // acc + 2 x
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t two = 2;
  asm("" : "+s"(two) : "i"(I));
  return (uint64_t)x * two + acc;
#else
  return acc + 2 * (uint64_t)x;
#endif
}

// v mod p below 2p for v < 2^37: v = hi 2^31 + lo and 2^31 = 2^24 - 1
// (mod p), so r = lo + hi (2^24 - 1) < 2^31 + 2^30 (hi < 64); on the device
// the product and sum are one multiply-add.
LG_HD uint32_t p2_fold31(uint64_t v) {
  const uint32_t lo = (uint32_t)v & 0x7fffffffu, hi = (uint32_t)(v >> 31);
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul24(hi, 0xffffffu) + lo;
#else
  return lo + (hi << 24) - hi;
#endif
}

// M4 of chunks C.. in 64 bits (the chunk index is a template parameter so
// that every multiply-add has its own opaque multiplier).  Inputs below p
// (WIDE false) take 32-bit pair sums x0 + x1 < 2p; inputs below 2p (WIDE:
// the unreduced S-box outputs) enter the 64-bit sums one by one.
template <int W, int C, bool WIDE>
LG_HD void p2_m4_chunks(const uint32_t* s, uint64_t* y) {
  if constexpr (C < W / 4) {
    const uint32_t x0 = s[4 * C], x1 = s[4 * C + 1], x2 = s[4 * C + 2], x3 = s[4 * C + 3];
    constexpr int K = 12 * C;
    if constexpr (!WIDE) {
      const uint32_t a = x0 + x1, b = x2 + x3;
      const uint64_t t0123 = p2_acc<K>(b, p2_acc<K + 1>(a, 0));
      const uint64_t t01123 = p2_acc<K + 2>(x1, t0123), t01233 = p2_acc<K + 3>(x3, t0123);
      y[4 * C + 0] = p2_acc<K + 4>(a, t01123);
      y[4 * C + 1] = p2_acc2<K + 5>(x2, t01123);
      y[4 * C + 2] = p2_acc<K + 6>(b, t01233);
      y[4 * C + 3] = p2_acc2<K + 7>(x0, t01233);
    } else {
      const uint64_t t01 = p2_acc<K>(x1, p2_acc<K + 1>(x0, 0));
      const uint64_t t0123 = p2_acc<K + 2>(x3, p2_acc<K + 3>(x2, t01));
      const uint64_t t01123 = p2_acc<K + 4>(x1, t0123), t01233 = p2_acc<K + 5>(x3, t0123);
      y[4 * C + 0] = p2_acc<K + 6>(x1, p2_acc<K + 7>(x0, t01123));
      y[4 * C + 1] = p2_acc2<K + 8>(x2, t01123);
      y[4 * C + 2] = p2_acc<K + 9>(x3, p2_acc<K + 10>(x2, t01233));
      y[4 * C + 3] = p2_acc2<K + 11>(x0, t01233);
    }
    p2_m4_chunks<W, C + 1, WIDE>(s, y);
  }
}

// External layer; with RK >= 0 the round keys of full round RK (which the
// next S-boxes take) are added before the reduction.  One reduction per
// output: the M4 outputs (coefficient sum 7) and the chunk-wise sums are
// formed in 64 bits, so every output (plus a key) is below 50p < 50 * 2^31
// for inputs below p, or 99p for inputs below 2p (WIDE); with v = hi 2^31 +
// lo and 2^31 = 2^24 - 1 (mod p), lo + hi 2^24 - hi is below 2^31 + 99 * 2^24
// < 2p, and one unsigned minimum gives the output below p.
template <int W, int RK = -1, bool WIDE = false>
LG_HD void p2_external(uint32_t* s) {
  uint64_t y[W];
  p2_m4_chunks<W, 0, WIDE>(s, y);
  uint64_t sum[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    uint64_t acc = y[j];
#pragma unroll
    for (int c = 1; c < W / 4; ++c) acc += y[4 * c + j];
    sum[j] = acc;
  }
#pragma unroll
  for (int i = 0; i < W; ++i) {
    uint64_t v = y[i] + sum[i & 3];
    if constexpr (RK >= 0) v += P2Consts<W>::full(RK, i);
    const uint32_t r = p2_fold31(v);
    s[i] = kb_umin(r, r - KB_P);
  }
}

template <>
struct P2Consts<24> {
  static LG_HD constexpr uint32_t full(int r, int i) { return P2_W24_FULL_RK[r][i]; }
  static LG_HD constexpr uint32_t partial(int r) { return P2_W24_PARTIAL_RK[r]; }
  static LG_HD constexpr uint32_t diag(int i) { return P2_W24_DIAG[i]; }
};

template <>
struct P2Consts<16> {
  static LG_HD constexpr uint32_t full(int r, int i) { return P2_W16_FULL_RK[r][i]; }
  static LG_HD constexpr uint32_t partial(int r) { return P2_W16_PARTIAL_RK[r]; }
  static LG_HD constexpr uint32_t diag(int i) { return P2_W16_DIAG[i]; }
};

// x^3 for x < p.  The square skips its final correction: it is below 2p, and
// a Montgomery product of a < 2p and b < p stays exact (t = ab < 2p^2 < 2^63,
// t + m p < 2^64) and below 2p before kb_mul's own correction.
LG_HD uint32_t kb_cube(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t t = (uint64_t)x * x;
  const uint32_t m = (uint32_t)t * 0x7effffffu;
  const uint32_t sq = (uint32_t)((t + (uint64_t)m * KB_P) >> 32);
  return kb_mul(sq, x);
#else
  // On the host one reduction by 2^64 of the 93-bit cube (x^3 R^3 -> x^3 R):
  // m = t (-p^-1) mod 2^64, t + m p < 2^95 is divisible by 2^64, and the
  // quotient is below p^3/2^64 + p < 1.25 p.  Four multiplies instead of six.
  constexpr uint64_t NEG_PINV64 = [] {
    uint64_t inv = KB_P;  // Newton: correct to 3, 6, 12, 24, 48, 96 bits
    for (int i = 0; i < 6; ++i) inv *= 2 - (uint64_t)KB_P * inv;
    return 0 - inv;
  }();
  const unsigned __int128 t = (unsigned __int128)((uint64_t)x * x) * x;
  const uint64_t m = (uint64_t)t * NEG_PINV64;
  const uint32_t r = (uint32_t)((t + (unsigned __int128)m * KB_P) >> 64);
  return r >= KB_P ? r - KB_P : r;
#endif
}

// x^3 below 2p for x < p: the cube without kb_mul's final correction (the
// product of sq < 2p and x < p is exact and its Montgomery quotient is below
// 2p).  Its consumers take operands below 2p: the external layer's wide form
// and the first operand of a Montgomery product.  The host returns x^3 < p.
LG_HD uint32_t kb_cube_lazy(uint32_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  const uint64_t t = (uint64_t)x * x;
  const uint32_t m = (uint32_t)t * 0x7effffffu;
  const uint32_t sq = (uint32_t)((t + (uint64_t)m * KB_P) >> 32);
  const uint64_t u = (uint64_t)sq * x;
  const uint32_t n = (uint32_t)u * 0x7effffffu;
  return (uint32_t)((u + (uint64_t)n * KB_P) >> 32);
#else
  return kb_cube(x);
#endif
}

// S-boxes of a full round whose keys are already in the state; outputs
// below 2p (for the wide external layer).
template <int W>
LG_HD void p2_cube_all(uint32_t* s) {
#pragma unroll
  for (int i = 0; i < W; ++i) s[i] = kb_cube_lazy(s[i]);
}

// The internal diagonal entries are small integers and (negated) inverse
// powers of two; each lane's product d_i * s_i is specialised at compile time
// (additions, or kb_div_pow2) and falls back to a Montgomery product.
struct DiagOp {
  int neg;   // 1: subtract the term from the sum
  int kind;  // 0: Montgomery product, 1: small integer c, 2: divide by 2^k
  int c;     // small integer or k
};

constexpr uint32_t p2_pow_canon(uint32_t b, uint32_t e) {
  uint64_t r = 1, x = b;
  while (e) {
    if (e & 1) r = r * x % KB_P;
    x = x * x % KB_P;
    e >>= 1;
  }
  return (uint32_t)r;
}

constexpr DiagOp p2_diag_op(uint32_t d) {
  for (int neg = 0; neg < 2; ++neg) {
    const uint32_t v = neg ? (d ? KB_P - d : 0u) : d;
    for (uint32_t c = 1; c <= 4; ++c)
      if (v == c) return DiagOp{neg, 1, (int)c};
    for (int k = 1; k <= 24; ++k)
      if (v == p2_pow_canon(p2_pow_canon(2, (uint32_t)k), KB_P - 2)) return DiagOp{neg, 2, k};
  }
  return DiagOp{0, 0, 0};
}

template <int W>
struct P2Diag;
template <>
struct P2Diag<24> {
  static constexpr uint32_t canon(int i) { return P2_W24_DIAG_CANON[i]; }
};
template <>
struct P2Diag<16> {
  static constexpr uint32_t canon(int i) { return P2_W16_DIAG_CANON[i]; }
};

template <int W, int I>
LG_HD uint32_t p2_diag_lane(uint32_t sum, uint32_t x) {
  constexpr DiagOp op = p2_diag_op(P2Diag<W>::canon(I));
  uint32_t t;
  if constexpr (op.kind == 1) {
    if constexpr (op.c == 1) t = x;
    else if constexpr (op.c == 2) t = kb_dbl(x);
    else if constexpr (op.c == 3) t = kb_add(kb_dbl(x), x);
    else t = kb_dbl(kb_dbl(x));
  } else if constexpr (op.kind == 2) {
    t = kb_div_pow2<op.c>(x);
  } else {
    return kb_add(sum, kb_mul(x, P2Consts<W>::diag(I)));
  }
  if constexpr (op.neg) return kb_sub(sum, t);
  else return kb_add(sum, t);
}

template <int W, int I>
LG_HD void p2_diag_all(uint32_t* s, uint32_t sum) {
  if constexpr (I < W) {
    s[I] = p2_diag_lane<W, I>(sum, s[I]);
    p2_diag_all<W, I + 1>(s, sum);
  }
}

// Sum of the W state elements mod p: pair sums of two elements below p fit
// 32 bits, the W/2 pair sums accumulate in 64 bits (v_lshl_add_u64 on the
// device), and one reduction with 2^31 = 2^24 - 1 (mod p) follows: acc =
// hi 2^31 + lo with hi < W, so lo + hi 2^24 - hi < 2^31 + 2^29 < 2p.
template <int W>
LG_HD uint32_t p2_state_sum(const uint32_t* s) {
  static_assert(W % 2 == 0 && W <= 32, "pairwise sum needs an even width up to 32");
  uint64_t acc = 0;
#pragma unroll
  for (int i = 0; i < W; i += 2) acc += (uint64_t)(s[i] + s[i + 1]);
  const uint32_t r = p2_fold31(acc);
  return kb_umin(r, r - KB_P);
}

template <int W>
LG_HD void p2_partial_round(uint32_t* s, int r) {
  s[0] = kb_cube(kb_add(s[0], P2Consts<W>::partial(r)));
  p2_diag_all<W, 0>(s, p2_state_sum<W>(s));
}

// Device form of the 21 partial rounds: the internal layer without a
// reduction per lane.  Round r keeps lane values u_i = sigma_r s_i (mod p),
// below 2p, for the field scale sigma_r = 2^(-8r) (s: the Montgomery state).
// The layer s_i <- S + d_i s_i becomes, times 2^24,
//   T_i = 2^24 sum_j u_j + e_i u_i,   e_i = 2^24 d_i,
// where every diagonal entry here is a small integer or +-2^-k (k <= 24), so
// e_i is an integer up to 2^26 (a negative one multiplies 2p - u_i instead);
// T_i < 48p 2^24 + 2^26 2p < 2^62, and one Montgomery reduction
// (T_i + m p) / 2^32, m = T_i (-p^-1) mod 2^32, gives the next lane below
// (2^62 + 2^63) / 2^32 < 2p with scale sigma_r 2^24 2^-32 = sigma_(r+1): two
// multiply-adds per lane instead of a modular addition and a reduced product.
// Lane 0's S-box works on the canonical scaled value plus the scaled key and
// scales its cube back (times sigma_r^-2).  After the last round one product
// by 2^168 restores the scale and reduces.
constexpr uint32_t p2_mulmod(uint32_t a, uint32_t b) { return (uint32_t)((uint64_t)a * b % KB_P); }

template <int W>
struct P2Scaled {
  // kind of each diagonal entry as a multiple of 2^-24: e = 2^24 d as a
  // signed integer (0 if d is neither a small integer nor +-2^-k)
  static constexpr int64_t e(int i) {
    const DiagOp op = p2_diag_op(P2Diag<W>::canon(i));
    const int64_t m = op.kind == 1 ? ((int64_t)op.c << 24) : (op.kind == 2 ? ((int64_t)1 << (24 - op.c)) : 0);
    return op.neg ? -m : m;
  }
  static constexpr bool ok() {
    for (int i = 0; i < W; ++i)
      if (e(i) == 0) return false;
    return true;
  }
  // sigma_r times the partial-round key r (Montgomery form is linear)
  static constexpr uint32_t rk(int r) {
    return p2_mulmod(P2Consts<W>::partial(r), p2_pow_canon(p2_pow_canon(2, KB_P - 2), 8u * (uint32_t)r));
  }
  // Montgomery form of sigma_r^-2 = 2^(16r)
  static constexpr uint32_t cfix(int r) { return p2_mulmod(p2_pow_canon(2, 16u * (uint32_t)r), KB_R); }
  // Montgomery form of sigma_21^-1 = 2^168
  static constexpr uint32_t cend() { return p2_mulmod(p2_pow_canon(2, 168), KB_R); }
};

// x * C + acc as one 32 x 32 + 64-bit multiply-add (v_mad_u64_u32): on the
// device the constant passes through an empty asm, so that a power of two
// does not become a zero extension, a 64-bit shift and a 64-bit add.
template <uint32_t C>
LG_HD uint64_t p2_mad_c(uint32_t x, uint64_t acc) {
#if defined(__HIP_DEVICE_COMPILE__)
  uint32_t c = C;
  asm("" : "+s"(c));
  return (uint64_t)x * c + acc;
#else
  return (uint64_t)x * C + acc;
#endif
}

// Sum of the W lane values (each below 2p) in 64 bits, one multiply-add by an
// opaque 1 per lane: plain additions of zero-extended values cost a register
// move and a 64-bit add each.  The asm input I keeps the W multipliers
// distinct, so that they are not merged and the sum not re-factored.
template <int W, int I>
LG_HD uint64_t p2_scaled_sum(const uint32_t* u, uint64_t acc) {
  if constexpr (I < W) {
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t one = 1;
    asm("" : "+s"(one) : "i"(I));
    acc = (uint64_t)u[I] * one + acc;
#else
    acc += u[I];
#endif
    return p2_scaled_sum<W, I + 1>(u, acc);
  }
  return acc;
}

template <int W, int I>
LG_HD void p2_scaled_lane(uint32_t* u, uint64_t s24) {
  if constexpr (I < W) {
    constexpr int64_t e = P2Scaled<W>::e(I);
    const uint32_t x = e < 0 ? 2u * KB_P - u[I] : u[I];
    const uint64_t t = p2_mad_c<(uint32_t)(e < 0 ? -e : e)>(x, s24);
    const uint32_t m = (uint32_t)t * 0x7effffffu;
    u[I] = (uint32_t)((t + (uint64_t)m * KB_P) >> 32);
    p2_scaled_lane<W, I + 1>(u, s24);
  }
}

// Round R of the scaled partial rounds (the round is a template parameter so
// that its key and scale constants are compile-time literals).
template <int W, int R>
LG_HD void p2_partial_scaled(uint32_t* u) {
  if constexpr (R < 21) {
    constexpr uint32_t rk = P2Scaled<W>::rk(R), cfix = P2Scaled<W>::cfix(R);
    uint32_t x0 = kb_umin(u[0], u[0] - KB_P);
    x0 = kb_add(x0, rk);
    // cube below 2p times cfix < p: the Montgomery quotient, below 2p, is a
    // lane value as it stands
    const uint64_t t = (uint64_t)kb_cube_lazy(x0) * cfix;
    const uint32_t m = (uint32_t)t * 0x7effffffu;
    u[0] = (uint32_t)((t + (uint64_t)m * KB_P) >> 32);
    p2_scaled_lane<W, 0>(u, p2_scaled_sum<W, 0>(u, 0) << 24);
    p2_partial_scaled<W, R + 1>(u);
  }
}

template <int W>
LG_HD void p2_partial_rounds_scaled(uint32_t* u) {
  static_assert(P2Scaled<W>::ok(), "diagonal entries must be small integers or +-2^-k");
  p2_partial_scaled<W, 0>(u);
  constexpr uint32_t cend = P2Scaled<W>::cend();
#pragma unroll
  for (int i = 0; i < W; ++i) u[i] = kb_mul(u[i], cend);
}

#if !defined(__HIP_DEVICE_COMPILE__)
// Host form of the partial rounds.  The internal layer is linear with small
// diagonal entries, so lanes 1..W-1 are carried as signed 64-bit integers
// congruent to the state and reduced only every fourth round: the sum is a
// plain integer sum, c*x a plain product, and x*2^-k is (x + m p) >> k with
// m = -x mod 2^k (p = 1 mod 2^24 makes x + m p divisible by 2^k; >> of a
// negative value is the arithmetic shift in GCC and Clang).  Lane 0 is
// reduced every round for its S-box.  Bound: a lane starts a run below
// M = 2^31, one round gives |S| <= p + (W-1) M and |lane| <= |S| + 4 M + p
// <= 27 M + 2p (W = 24), so four rounds stay below 2^50.2, the round sums
// below 2^55, well inside the 2^62 that kb_reduce_i64 accepts.
inline uint32_t kb_reduce_i64(int64_t v) {
  constexpr uint64_t OFF = ((1ull << 62) / KB_P + 1) * KB_P;  // multiple of p, >= 2^62
  return (uint32_t)(((uint64_t)v + OFF) % KB_P);
}

template <int W, int I>
inline int64_t p2_diag_i64(int64_t sum, int64_t x) {
  constexpr DiagOp op = p2_diag_op(P2Diag<W>::canon(I));
  int64_t t;
  if constexpr (op.kind == 1) {
    t = (int64_t)op.c * x;
  } else if constexpr (op.kind == 2) {
    const uint64_t m = (0u - (uint64_t)x) & ((1ull << op.c) - 1u);
    t = (x + (int64_t)(m * KB_P)) >> op.c;
  } else {
    t = kb_mul(kb_reduce_i64(x), P2Consts<W>::diag(I));
  }
  return op.neg ? sum - t : sum + t;
}

template <int W, int I>
inline void p2_diag_rest_i64(int64_t* v, int64_t sum) {
  if constexpr (I < W) {
    v[I] = p2_diag_i64<W, I>(sum, v[I]);
    p2_diag_rest_i64<W, I + 1>(v, sum);
  }
}

template <int W>
inline void p2_partial_rounds_host(uint32_t* s) {
  static_assert(W <= 24 && W % 2 == 0, "lazy bound derived for even widths up to 24");
  int64_t v[W];  // v[0] holds lane 0 between rounds
  for (int i = 0; i < W; ++i) v[i] = s[i];
  // not unrolled: the unrolled rounds overflow the instruction cache
#pragma nounroll
  for (int r = 0; r < 21; ++r) {
    const uint32_t x0 = kb_cube(kb_add(kb_reduce_i64(v[0]), P2Consts<W>::partial(r)));
    // the sum of lanes 1..W-1 does not wait for the S-box (pairwise: short chain)
    int64_t pr[W / 2];
    pr[0] = v[1];
    for (int i = 1; i < W / 2; ++i) pr[i] = v[2 * i] + v[2 * i + 1];
    for (int w = 1; w < W / 2; w *= 2)
      for (int i = 0; i + w < W / 2; i += 2 * w) pr[i] += pr[i + w];
    const int64_t sum = pr[0] + x0;
    v[0] = p2_diag_i64<W, 0>(sum, x0);
    p2_diag_rest_i64<W, 1>(v, sum);
    if ((r & 3) == 3)
      for (int i = 1; i < W; ++i) v[i] = kb_reduce_i64(v[i]);
  }
  for (int i = 0; i < W; ++i) s[i] = kb_reduce_i64(v[i]);
}
#endif

// The permutation with the partial rounds in their device form (DEV) or
// their host form; p2_permute picks the one of the compilation target (the
// host build of the device form is for tests).
template <int W, bool DEV>
LG_HD void p2_permute_form(uint32_t* s) {
  // full round r = keys, S-boxes, external layer; the keys of a full round
  // that follows an external layer are folded into that layer's reduction
  p2_external<W, 0>(s);
  p2_cube_all<W>(s);
  p2_external<W, 1, true>(s);
  p2_cube_all<W>(s);
  p2_external<W, 2, true>(s);
  p2_cube_all<W>(s);
  p2_external<W, -1, true>(s);
#if defined(__HIP_DEVICE_COMPILE__)
  if constexpr (P2Scaled<W>::ok()) p2_partial_rounds_scaled<W>(s);
  else {
#pragma unroll
    for (int r = 0; r < 21; ++r) p2_partial_round<W>(s, r);
  }
#else
  if constexpr (DEV) p2_partial_rounds_scaled<W>(s);
  else p2_partial_rounds_host<W>(s);
#endif
#pragma unroll
  for (int i = 0; i < W; ++i) s[i] = kb_add(s[i], P2Consts<W>::full(3, i));
  p2_cube_all<W>(s);
  p2_external<W, 4, true>(s);
  p2_cube_all<W>(s);
  p2_external<W, 5, true>(s);
  p2_cube_all<W>(s);
  p2_external<W, -1, true>(s);
}

template <int W>
LG_HD void p2_permute(uint32_t* s) {
#if defined(__HIP_DEVICE_COMPILE__)
  p2_permute_form<W, true>(s);
#else
  p2_permute_form<W, false>(s);
#endif
}

// node = P([NODE, 0 x 7, L, R])[0..8]; l, r, out: 8 Montgomery elements each.
LG_HD void p2_node(const uint32_t* l, const uint32_t* r, uint32_t* out) {
  uint32_t s[24];
  s[0] = TAG_NODE_MONT;
#pragma unroll
  for (int i = 1; i < 8; ++i) s[i] = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    s[8 + i] = l[i];
    s[16 + i] = r[i];
  }
  p2_permute<24>(s);
#pragma unroll
  for (int i = 0; i < 8; ++i) out[i] = s[i];
}

// Host-side sponge over an explicit message (transcript, tests).
inline void p2_sponge_host(const uint32_t* msg, size_t len, uint32_t* digest) {
  uint32_t s[24] = {0};
  if (len == 0) {
    for (int i = 0; i < 8; ++i) digest[i] = 0;
    return;
  }
  for (size_t off = 0; off < len; off += 16) {
    size_t k = len - off < 16 ? len - off : 16;
    for (size_t i = 0; i < k; ++i) s[i] = msg[off + i];
    p2_permute<24>(s);
  }
  for (int i = 0; i < 8; ++i) digest[i] = s[i];
}

}  // namespace lg

#if defined(__HIPCC__)
namespace lg {

// ---------------------------------------------------------------- cooperative
// One width-24 state spread over a DPP row (16 consecutive lanes of a
// wavefront): lane l < 12 holds state elements l (register a) and l + 12
// (register b), lanes 12..15 hold zeros.  Chunks 0-2 of the external layer sit
// in a and chunks 3-5 in b, on lanes 4c..4c+3 in both, so M4 is a quad
// permutation and the chunk-wise and full sums are row rotations (DPP, no
// LDS).  The dependent chain of one permutation is a few times shorter than
// with one state per lane; used where few independent permutations are
// available (tree tops, FRI challenges) and latency bounds the time.

template <int CTRL>
__device__ __forceinline__ uint32_t dpp(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
}
// quad_perm [1,2,3,0], [2,3,0,1], [3,0,1,2], [1,0,3,2]; row_ror:4, row_ror:8
constexpr int DPP_QUAD_ROT1 = 0x39, DPP_QUAD_ROT2 = 0x4E, DPP_QUAD_ROT3 = 0x93, DPP_QUAD_XOR1 = 0xB1;
constexpr int DPP_ROW_ROR4 = 0x124, DPP_ROW_ROR8 = 0x128;

// Sum over the lanes of a row with the same index mod 4, to every lane.
__device__ __forceinline__ uint32_t row_col_sum(uint32_t v) {
  v = kb_add(v, dpp<DPP_ROW_ROR4>(v));
  return kb_add(v, dpp<DPP_ROW_ROR8>(v));
}

// Sum over all lanes of a row, to every lane.
__device__ __forceinline__ uint32_t row_sum(uint32_t v) {
  v = kb_add(v, dpp<DPP_QUAD_XOR1>(v));
  v = kb_add(v, dpp<DPP_QUAD_ROT2>(v));
  return row_col_sum(v);
}

// M4 of the lane's quad, row j of the circulant [[2,3,1,1],[1,2,3,1],...]:
// 2 x_j + 3 x_{j+1} + x_{j+2} + x_{j+3}, formed in 64 bits (below 7p) and
// reduced once with 2^31 = 2^24 - 1 (mod p).
__device__ __forceinline__ uint32_t quad_m4(uint32_t x) {
  const uint32_t x1 = dpp<DPP_QUAD_ROT1>(x), x2 = dpp<DPP_QUAD_ROT2>(x), x3 = dpp<DPP_QUAD_ROT3>(x);
  const uint64_t v = 2 * (uint64_t)(x + x1) + x1 + (uint64_t)(x2 + x3);
  const uint32_t r = p2_fold31(v);
  return kb_umin(r, r - KB_P);
}

// (t + m p) / 2^32 for t < 2^63 + 2^57, reduced below p (t's quotient is below 2p)
__device__ __forceinline__ uint32_t coop_redc(uint64_t t) {
  const uint32_t m = (uint32_t)t * 0x7effffffu;
  const uint32_t r = (uint32_t)((t + (uint64_t)m * KB_P) >> 32);
  return kb_umin(r, r - KB_P);
}

struct P2Coop16 {
  uint32_t rka[6], rkb[6];  // full-round keys of elements l and l + 12 (zeros on lanes 12..15)
  uint32_t da, db;          // internal diagonal of elements l and l + 12
  uint32_t amask;           // all ones on lanes l < 12
  uint32_t l0mask;          // all ones on lane 0

  __device__ __forceinline__ void init(int l) {
    const bool active = l < 12;
    const int i = active ? l : 0;
    amask = active ? 0xffffffffu : 0u;
    l0mask = l == 0 ? 0xffffffffu : 0u;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      rka[r] = active ? P2_W24_FULL_RK[r][i] : 0u;
      rkb[r] = active ? P2_W24_FULL_RK[r][i + 12] : 0u;
    }
    da = active ? P2_W24_DIAG[i] : 0u;
    db = active ? P2_W24_DIAG[i + 12] : 0u;
  }

  // lanes 12..15 hold zeros, which M4 keeps; the chunk-wise sums reach them
  // too, so they are masked off there
  __device__ __forceinline__ void external(uint32_t& a, uint32_t& b) const {
    const uint32_t ya = quad_m4(a), yb = quad_m4(b);
    const uint32_t s = row_col_sum(kb_add(ya, yb)) & amask;
    a = kb_add(ya, s);
    b = kb_add(yb, s);
  }

  __device__ __forceinline__ void permute(uint32_t& a, uint32_t& b) const {
    external(a, b);
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      a = kb_cube(kb_add(a, rka[r]));
      b = kb_cube(kb_add(b, rkb[r]));
      external(a, b);
    }
#pragma unroll
    for (int r = 0; r < 21; ++r) {
      const uint32_t c = kb_cube(kb_add(a, P2_W24_PARTIAL_RK[r]));
      a = (c & l0mask) | (a & ~l0mask);
      // inactive lanes: a = b = 0, da = db = 0 and the sum masked to 0
      const uint32_t s = row_sum(kb_add(a, b)) & amask;
      // s + x d as one Montgomery reduction of x d + s R (R = 2^32 mod p, so
      // the quotient is x d R^-1 + s): x d < p^2, s R < 2^56, and the sum
      // plus m p stays below 2^64; the quotient is below 2p
      const uint64_t sr = (uint64_t)s * KB_R;
      a = coop_redc((uint64_t)a * da + sr);
      b = coop_redc((uint64_t)b * db + sr);
    }
#pragma unroll
    for (int r = 3; r < 6; ++r) {
      a = kb_cube(kb_add(a, rka[r]));
      b = kb_cube(kb_add(b, rkb[r]));
      external(a, b);
    }
  }
};

}  // namespace lg
#endif
