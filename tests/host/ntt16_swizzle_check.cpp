// Host check of the LDS index decomposition the radix-16 NTT kernel relies
// on: for every size 2^8..2^14, every pass stride q = 2^lq, thread t and
// value k, lds_swz(ntt16_pos(t, q, k)) = lds_swz(ntt16_base(t, lq)) ^
// lds_swz(k << lq); the bit-reversed scatter index lds_swz(brev_L(t + k n/16))
// = lds_swz(brev_{L-4}(t) << 4) ^ brev4(k); the 16-value row index
// lds_swz(16 t + k) = lds_swz(16 t) ^ k; and lds_swz is a bijection on
// [0, n).  Prints the number of mismatches and exits non-zero on any.
#include <cstdio>
#include <vector>

#include "ntt16.h"

using namespace lg;

static uint32_t brev_bits(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

int main() {
  long bad = 0, checked = 0;
  for (int L = 8; L <= 14; ++L) {
    const uint32_t n = 1u << L, T = n / 16, q0 = n / 16;
    std::vector<int> seen(n, 0);
    for (uint32_t a = 0; a < n; ++a) {
      const uint32_t s = lds_swz(a);
      if (s >= n || seen[s]++) ++bad;
    }
    for (int lq = 0; lq + 4 <= L; ++lq) {
      const uint32_t q = 1u << lq;
      for (uint32_t t = 0; t < T; ++t) {
        const uint32_t sb = lds_swz(ntt16_base(t, lq));
        for (int k = 0; k < 16; ++k) {
          ++checked;
          if (lds_swz(ntt16_pos(t, q, k)) != (sb ^ lds_swz((uint32_t)k << lq))) ++bad;
        }
      }
    }
    for (uint32_t t = 0; t < T; ++t) {
      const uint32_t s_br = lds_swz(brev_bits(t, L - 4) << 4), s_row = lds_swz(16 * t);
      for (int k = 0; k < 16; ++k) {
        checked += 2;
        if (lds_swz(brev_bits(t + (uint32_t)k * q0, L)) != (s_br ^ brev_bits((uint32_t)k, 4))) ++bad;
        if (lds_swz(16 * t + (uint32_t)k) != (s_row ^ (uint32_t)k)) ++bad;
      }
    }
  }
  printf("checked %ld mismatches %ld\n", checked, bad);
  return bad ? 1 : 0;
}
