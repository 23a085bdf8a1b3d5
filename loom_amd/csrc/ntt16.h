// Radix-16 building blocks of the LDS-resident NTT (sizes 2^8 .. 2^14, one
// workgroup of n/16 threads per transform, 16 values per thread).
//
// Forward transforms are decimation-in-frequency (natural input, output in
// bit-reversed order); inverse transforms are decimation-in-time
// (bit-reversed input, natural output).  log2 n = 4P + r: P passes of four
// radix-2 stages on values held in registers, and one pass of r stages.  A
// full pass on sub-transforms of size m (q = m/16) gives thread t the 16
// positions blk*m + b + k*q (blk = t / q, b = t % q, k < 16); its twiddles are
// W = w_m^b (one table load) times constant roots of order 16, 8, 4, 2.  The
// pass of r < 4 stages works on contiguous blocks of 2^r values, 16 per
// thread, with constant twiddles.  The functions are host-callable so that a
// CPU build can run every thread of a pass in turn.
#pragma once
#include "kb.h"

namespace lg {

// Roots of unity of order 16, 8, 4 (Montgomery), in the transform's
// direction: r16[k] = w16^k (k < 8), r8[k] = w8^k (k < 4), r4 = w4.
struct NttRoots {
  uint32_t r16[8];
  uint32_t r8[4];
  uint32_t r4;
};

// Position of value k of thread t in a full pass on sub-size m = 16 q.
LG_HD uint32_t ntt16_pos(uint32_t t, uint32_t q, int k) { return (t / q) * 16u * q + (t % q) + (uint32_t)k * q; }

// LDS word index of element a in the LDS-resident NTT: a ^ f(a >> 6) with
// f(h) = ((h << 2) ^ h ^ (h >> 5)) mod 64, a bijection on [0, n) for n >= 64
// (only the bank bits change, as a function of the others).  The kernel's
// access patterns (bit-reversed scatter, 16 consecutive words per thread,
// pass strides q = n/16 .. 1) then meet at most 2 words per bank where the
// plain index meets up to 64 (found by exhaustive search over XOR forms for
// 2^8..2^14 with 32 and 64 banks).
//
// The map is linear over GF(2) (XORs of shifted copies, masked), so for a
// and b with disjoint bits lds_swz(a | b) = lds_swz(a) ^ lds_swz(b).  Every
// access pattern of a pass is a per-thread base OR-ed with k << lq (value
// index k < 16, fixed lq), so its LDS index is lds_swz(base) ^ lds_swz(k <<
// lq): one XOR with a compile-time constant once the pass is unrolled.
LG_HD constexpr uint32_t lds_swz(uint32_t a) {
  return a ^ ((((a >> 6) << 2) ^ (a >> 6) ^ (a >> 11)) & 63u);
}

// The base of thread t in a full pass with q = 2^lq: ntt16_pos(t, q, k) =
// ntt16_base(t, lq) | (k << lq), the two parts with disjoint bits.
LG_HD uint32_t ntt16_base(uint32_t t, int lq) {
  return ((t >> lq) << (lq + 4)) | (t & ((1u << lq) - 1u));
}

// Twiddle ladder of a full pass from W = w_m^b.
struct Ladder {
  uint32_t a[8];  // W * w16^k  (size-m stage)
  uint32_t b[4];  // W^2 * w8^k (size m/2)
  uint32_t c[2];  // W^4 * w4^k (size m/4)
  uint32_t d;     // W^8        (size m/8)
};

LG_HD Ladder ntt16_ladder(uint32_t W, const NttRoots& R) {
  Ladder L;
  const uint32_t W2 = kb_mul(W, W), W4 = kb_mul(W2, W2);
  L.d = kb_mul(W4, W4);
  L.a[0] = W;
#pragma unroll
  for (int k = 1; k < 8; ++k) L.a[k] = kb_mul(W, R.r16[k]);
  L.b[0] = W2;
#pragma unroll
  for (int k = 1; k < 4; ++k) L.b[k] = kb_mul(W2, R.r8[k]);
  L.c[0] = W4;
  L.c[1] = kb_mul(W4, R.r4);
  return L;
}

// DIF butterfly: (u, v) -> (u + v, (u - v) w).  The difference enters the
// product unreduced as u + (p - v) in [1, 2p): kb_mul's result is reduced for
// a first operand below 2p (twiddle w < p), which saves the difference's
// correction.
LG_HD void dif_bf(uint32_t& u, uint32_t& v, uint32_t w) {
  const uint32_t s = kb_add(u, v);
  v = kb_mul(u + (KB_P - v), w);
  u = s;
}
LG_HD void dif_bf1(uint32_t& u, uint32_t& v) {
  const uint32_t s = kb_add(u, v);
  v = kb_sub(u, v);
  u = s;
}
// DIT butterfly: (u, v) -> (u + v w, u - v w).
LG_HD void dit_bf(uint32_t& u, uint32_t& v, uint32_t w) {
  const uint32_t t = kb_mul(v, w);
  v = kb_sub(u, t);
  u = kb_add(u, t);
}
LG_HD void dit_bf1(uint32_t& u, uint32_t& v) {
  const uint32_t t = v;
  v = kb_sub(u, t);
  u = kb_add(u, t);
}

// Four DIF stages (sizes m, m/2, m/4, m/8) on the 16 values of a full pass.
LG_HD void dif16(uint32_t* x, const Ladder& L) {
#pragma unroll
  for (int k = 0; k < 8; ++k) dif_bf(x[k], x[k + 8], L.a[k]);
#pragma unroll
  for (int h = 0; h < 16; h += 8)
#pragma unroll
    for (int k = 0; k < 4; ++k) dif_bf(x[h + k], x[h + k + 4], L.b[k]);
#pragma unroll
  for (int h = 0; h < 16; h += 4)
#pragma unroll
    for (int k = 0; k < 2; ++k) dif_bf(x[h + k], x[h + k + 2], L.c[k]);
#pragma unroll
  for (int h = 0; h < 16; h += 2) dif_bf(x[h], x[h + 1], L.d);
}

// Four DIT stages (sizes m/8, m/4, m/2, m), the mirror of dif16.
LG_HD void dit16(uint32_t* x, const Ladder& L) {
#pragma unroll
  for (int h = 0; h < 16; h += 2) dit_bf(x[h], x[h + 1], L.d);
#pragma unroll
  for (int h = 0; h < 16; h += 4)
#pragma unroll
    for (int k = 0; k < 2; ++k) dit_bf(x[h + k], x[h + k + 2], L.c[k]);
#pragma unroll
  for (int h = 0; h < 16; h += 8)
#pragma unroll
    for (int k = 0; k < 4; ++k) dit_bf(x[h + k], x[h + k + 4], L.b[k]);
#pragma unroll
  for (int k = 0; k < 8; ++k) dit_bf(x[k], x[k + 8], L.a[k]);
}

// The r-stage pass (r = 1, 2, 3) on 16 contiguous values: 16 / 2^r blocks.
template <int R>
LG_HD void dif_small(uint32_t* x, const NttRoots& Rt) {
  constexpr int m = 1 << R;
#pragma unroll
  for (int g = 0; g < 16; g += m) {
    uint32_t* y = x + g;
    if constexpr (R == 3) {
#pragma unroll
      for (int k = 0; k < 4; ++k) dif_bf(y[k], y[k + 4], Rt.r8[k]);
#pragma unroll
      for (int h = 0; h < 8; h += 4) {
        dif_bf1(y[h], y[h + 2]);
        dif_bf(y[h + 1], y[h + 3], Rt.r4);
      }
#pragma unroll
      for (int h = 0; h < 8; h += 2) dif_bf1(y[h], y[h + 1]);
    } else if constexpr (R == 2) {
      dif_bf1(y[0], y[2]);
      dif_bf(y[1], y[3], Rt.r4);
      dif_bf1(y[0], y[1]);
      dif_bf1(y[2], y[3]);
    } else {
      dif_bf1(y[0], y[1]);
    }
  }
}

template <int R>
LG_HD void dit_small(uint32_t* x, const NttRoots& Rt) {
  constexpr int m = 1 << R;
#pragma unroll
  for (int g = 0; g < 16; g += m) {
    uint32_t* y = x + g;
    if constexpr (R == 3) {
#pragma unroll
      for (int h = 0; h < 8; h += 2) dit_bf1(y[h], y[h + 1]);
#pragma unroll
      for (int h = 0; h < 8; h += 4) {
        dit_bf1(y[h], y[h + 2]);
        dit_bf(y[h + 1], y[h + 3], Rt.r4);
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) dit_bf(y[k], y[k + 4], Rt.r8[k]);
    } else if constexpr (R == 2) {
      dit_bf1(y[0], y[1]);
      dit_bf1(y[2], y[3]);
      dit_bf1(y[0], y[2]);
      dit_bf(y[1], y[3], Rt.r4);
    } else {
      dit_bf1(y[0], y[1]);
    }
  }
}

LG_HD void dif_small_r(uint32_t* x, int r, const NttRoots& Rt) {
  if (r == 3) dif_small<3>(x, Rt);
  else if (r == 2) dif_small<2>(x, Rt);
  else if (r == 1) dif_small<1>(x, Rt);
}

LG_HD void dit_small_r(uint32_t* x, int r, const NttRoots& Rt) {
  if (r == 3) dit_small<3>(x, Rt);
  else if (r == 2) dit_small<2>(x, Rt);
  else if (r == 1) dit_small<1>(x, Rt);
}

// Roots from a size-n twiddle table tw[e] = w_n^e (e < n/2).
LG_HD NttRoots ntt16_roots(const uint32_t* tw, int logn) {
  NttRoots R;
  const uint32_t s16 = 1u << (logn - 4), s8 = 1u << (logn - 3), s4 = 1u << (logn - 2);
  for (int k = 0; k < 8; ++k) R.r16[k] = tw[k * s16];
  for (int k = 0; k < 4; ++k) R.r8[k] = tw[k * s8];
  R.r4 = tw[s4];
  return R;
}

}  // namespace lg
