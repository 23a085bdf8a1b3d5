// KoalaBear field (p = 2^31 - 2^24 + 1) in Montgomery form, and the degree-6
// extension E6 = E2[v]/(v^3 - (u+1)), E2 = F_p[u]/(u^2 - 3).
//
// Elements are one u32 in Montgomery form (x*2^32 mod p), the in-memory layout
// of gnark-crypto's koalabear.Element that the reference prover uses.  E6
// limbs follow the reference order (B0.A0, B0.A1, B1.A0, B1.A1, B2.A0, B2.A1)
// (internal/hash/hash.go:59-68); the multiplication is the schoolbook product
// restated in recursion/e6.go:21-53.
//
// Usable from host code (g++ / hipcc host pass), device code and hiprtc.
#pragma once
#ifndef LG_HIPRTC
#include <cstdint>
#endif

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define LG_HD __host__ __device__ __forceinline__
#else
#define LG_HD inline
#endif

namespace lg {

constexpr uint32_t KB_P = 0x7f000001u;
constexpr uint32_t KB_PINV = 0x81000001u;   // p^-1 mod 2^32
constexpr uint32_t KB_R = 0x01fffffeu;      // 2^32 mod p  (Montgomery one)
constexpr uint32_t KB_R2 = 0x17f7efe4u;     // 2^64 mod p

// Unsigned minimum (v_min_u32 on the device).
LG_HD uint32_t kb_umin(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_elementwise_min(a, b);
#else
  return a < b ? a : b;
#endif
}

// On the device the conditional corrections are unsigned minima: a + b < 2p,
// and min(s, s - p) is s - p exactly when s >= p (else s - p wraps above s);
// for a - b, min(d, d + p) is d + p exactly when the difference wrapped.  This
// keeps them to three VALU instructions without a compare into VCC (and the
// hazard wait a compare-then-select pair needs).
LG_HD uint32_t kb_add(uint32_t a, uint32_t b) {
  const uint32_t r = a + b;
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_elementwise_min(r, r - KB_P);
#else
  return r >= KB_P ? r - KB_P : r;
#endif
}

LG_HD uint32_t kb_sub(uint32_t a, uint32_t b) {
  const uint32_t r = a - b;
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_elementwise_min(r, r + KB_P);
#else
  return a < b ? r + KB_P : r;
#endif
}

LG_HD uint32_t kb_neg(uint32_t a) { return a ? KB_P - a : 0u; }

// Montgomery product a*b*2^-32 mod p for a, b < p.  With m = -t p^-1 mod 2^32
// (-p^-1 = 0x7effffff), t + m p is divisible by 2^32 and below
// 2^62 + 2^63, and (t + m p) / 2^32 < p^2/2^32 + p < 2p: one 64-bit
// multiply-add and one conditional subtraction (on gfx950 a v_mad_u64_u32 with
// the accumulator, then v_add + v_min).  The result is also reduced for
// a < 2p, b < p: t + m p < 2p^2 + 2^32 p < 2^64 and the quotient is below
// 2p^2/2^32 + p < 2p.
LG_HD uint32_t kb_mul(uint32_t a, uint32_t b) {
  const uint64_t t = (uint64_t)a * b;
  const uint32_t m = (uint32_t)t * 0x7effffffu;
  const uint32_t r = (uint32_t)((t + (uint64_t)m * KB_P) >> 32);
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_elementwise_min(r, r - KB_P);
#else
  return r >= KB_P ? r - KB_P : r;
#endif
}

LG_HD uint32_t kb_dbl(uint32_t a) { return kb_add(a, a); }

// a * 2^-K mod p (a < p, 1 <= K <= 24), on the Montgomery representation
// directly (the map is linear): with m = -a mod 2^K, a + m p is divisible by
// 2^K, and since p = 2^31 - 2^24 + 1 the quotient is
// (a + m) / 2^K + m 2^(31-K) - m 2^(24-K), below 2p.
template <int K>
LG_HD uint32_t kb_div_pow2(uint32_t a) {
  static_assert(K >= 1 && K <= 24, "K out of range");
  const uint32_t m = (0u - a) & ((1u << K) - 1u);
#if defined(__HIP_DEVICE_COMPILE__)
  // m (2^(31-K) - 2^(24-K)) = (127 m) 2^(24-K) with 127 m < 2^24 for K <= 17:
  // a full-rate 24-bit multiply and a shift-add.  The empty asm keeps the
  // compiler from folding them back into one 64-bit multiply-add (a
  // quarter-rate instruction) when the constant exceeds 24 bits (K < 7).
  uint32_t y;
  if constexpr (K < 7) {
    uint32_t q = __umul24(m, 127u);
    asm("" : "+v"(q));
    y = (q << (24 - K)) + ((a + m) >> K);
  } else {
    y = ((a + m) >> K) + (m << (31 - K)) - (m << (24 - K));
  }
  return __builtin_elementwise_min(y, y - KB_P);
#else
  const uint32_t y = ((a + m) >> K) + (m << (31 - K)) - (m << (24 - K));
  return y >= KB_P ? y - KB_P : y;
#endif
}

LG_HD uint32_t kb_to_mont(uint32_t x) { return kb_mul(x % KB_P, KB_R2); }
LG_HD uint32_t kb_from_mont(uint32_t x) { return kb_mul(x, 1u); }

LG_HD uint32_t kb_pow(uint32_t a, uint64_t e) {
  uint32_t r = KB_R;
  while (e) {
    if (e & 1) r = kb_mul(r, a);
    a = kb_mul(a, a);
    e >>= 1;
  }
  return r;
}

LG_HD uint32_t kb_inv(uint32_t a) { return kb_pow(a, KB_P - 2); }

// ------------------------------------------------------------------ E2 / E6
struct E2 {
  uint32_t a0, a1;
};

struct E6 {
  uint32_t v[6];
};

LG_HD E2 e2_add(E2 x, E2 y) { return {kb_add(x.a0, y.a0), kb_add(x.a1, y.a1)}; }
LG_HD E2 e2_sub(E2 x, E2 y) { return {kb_sub(x.a0, y.a0), kb_sub(x.a1, y.a1)}; }
LG_HD uint32_t kb_mul3(uint32_t a) { return kb_add(kb_dbl(a), a); }
// (a0 + a1 u)(b0 + b1 u) = (a0 b0 + 3 a1 b1) + (a0 b1 + a1 b0) u, with the
// u coefficient as (a0 + a1)(b0 + b1) - a0 b0 - a1 b1 (three products).
LG_HD E2 e2_mul(E2 x, E2 y) {
  const uint32_t t0 = kb_mul(x.a0, y.a0), t1 = kb_mul(x.a1, y.a1);
  const uint32_t s = kb_mul(kb_add(x.a0, x.a1), kb_add(y.a0, y.a1));
  return {kb_add(t0, kb_mul3(t1)), kb_sub(s, kb_add(t0, t1))};
}
// times xi = 1 + u: (p + q u)(1 + u) = (p + 3q) + (p + q) u
LG_HD E2 e2_mul_xi(E2 x) { return {kb_add(x.a0, kb_mul3(x.a1)), kb_add(x.a0, x.a1)}; }
LG_HD E2 e2_scale(E2 x, uint32_t s) { return {kb_mul(x.a0, s), kb_mul(x.a1, s)}; }

LG_HD E2 e6_c(const E6& x, int i) { return {x.v[2 * i], x.v[2 * i + 1]}; }

LG_HD E6 e6_make(E2 c0, E2 c1, E2 c2) {
  E6 r;
  r.v[0] = c0.a0; r.v[1] = c0.a1; r.v[2] = c1.a0; r.v[3] = c1.a1; r.v[4] = c2.a0; r.v[5] = c2.a1;
  return r;
}

LG_HD E6 e6_zero() {
  E6 r;
  for (int i = 0; i < 6; ++i) r.v[i] = 0;
  return r;
}

LG_HD E6 e6_one() {
  E6 r = e6_zero();
  r.v[0] = KB_R;
  return r;
}

LG_HD E6 e6_from_base(uint32_t a) {
  E6 r = e6_zero();
  r.v[0] = a;
  return r;
}

LG_HD E6 e6_add(const E6& x, const E6& y) {
  E6 r;
  for (int i = 0; i < 6; ++i) r.v[i] = kb_add(x.v[i], y.v[i]);
  return r;
}

LG_HD E6 e6_sub(const E6& x, const E6& y) {
  E6 r;
  for (int i = 0; i < 6; ++i) r.v[i] = kb_sub(x.v[i], y.v[i]);
  return r;
}

LG_HD E6 e6_neg(const E6& x) {
  E6 r;
  for (int i = 0; i < 6; ++i) r.v[i] = kb_neg(x.v[i]);
  return r;
}

LG_HD E6 e6_scale(const E6& x, uint32_t s) {
  E6 r;
  for (int i = 0; i < 6; ++i) r.v[i] = kb_mul(x.v[i], s);
  return r;
}

LG_HD E6 e6_add_base(const E6& x, uint32_t b) {
  E6 r = x;
  r.v[0] = kb_add(r.v[0], b);
  return r;
}

// (x0 + x1 v + x2 v^2)(y0 + y1 v + y2 v^2) with v^3 = xi: the schoolbook
// coefficients z0 = x0y0 + xi (x1y2 + x2y1), z1 = x0y1 + x1y0 + xi x2y2,
// z2 = x0y2 + x1y1 + x2y0, with the cross sums from (xi + xj)(yi + yj) - xiyi
// - xjyj (six E2 products instead of nine).
LG_HD E6 e6_mul(const E6& x, const E6& y) {
  const E2 x0 = e6_c(x, 0), x1 = e6_c(x, 1), x2 = e6_c(x, 2);
  const E2 y0 = e6_c(y, 0), y1 = e6_c(y, 1), y2 = e6_c(y, 2);
  const E2 v0 = e2_mul(x0, y0), v1 = e2_mul(x1, y1), v2 = e2_mul(x2, y2);
  const E2 s12 = e2_sub(e2_mul(e2_add(x1, x2), e2_add(y1, y2)), e2_add(v1, v2));
  const E2 s01 = e2_sub(e2_mul(e2_add(x0, x1), e2_add(y0, y1)), e2_add(v0, v1));
  const E2 s02 = e2_sub(e2_mul(e2_add(x0, x2), e2_add(y0, y2)), e2_add(v0, v2));
  return e6_make(e2_add(v0, e2_mul_xi(s12)), e2_add(s01, e2_mul_xi(v2)), e2_add(s02, v1));
}

LG_HD E2 e2_inv(E2 x) {
  uint32_t n = kb_sub(kb_mul(x.a0, x.a0), kb_mul3(kb_mul(x.a1, x.a1)));
  uint32_t ni = kb_inv(n);
  return {kb_mul(x.a0, ni), kb_mul(kb_neg(x.a1), ni)};
}

LG_HD E6 e6_inv(const E6& x) {
  E2 a0 = e6_c(x, 0), a1 = e6_c(x, 1), a2 = e6_c(x, 2);
  E2 c0 = e2_sub(e2_mul(a0, a0), e2_mul_xi(e2_mul(a1, a2)));
  E2 c1 = e2_sub(e2_mul_xi(e2_mul(a2, a2)), e2_mul(a0, a1));
  E2 c2 = e2_sub(e2_mul(a1, a1), e2_mul(a0, a2));
  E2 t = e2_add(e2_mul(a0, c0), e2_mul_xi(e2_add(e2_mul(a2, c1), e2_mul(a1, c2))));
  E2 ti = e2_inv(t);
  return e6_make(e2_mul(c0, ti), e2_mul(c1, ti), e2_mul(c2, ti));
}

LG_HD bool e6_is_zero(const E6& x) {
  uint32_t acc = 0;
  for (int i = 0; i < 6; ++i) acc |= x.v[i];
  return acc == 0;
}

LG_HD bool e6_eq(const E6& x, const E6& y) {
  uint32_t acc = 0;
  for (int i = 0; i < 6; ++i) acc |= x.v[i] ^ y.v[i];
  return acc == 0;
}

LG_HD E6 e6_pow(E6 a, uint64_t e) {
  E6 r = e6_one();
  while (e) {
    if (e & 1) r = e6_mul(r, a);
    a = e6_mul(a, a);
    e >>= 1;
  }
  return r;
}

}  // namespace lg
