// Host build of the device field and Poseidon2 code (kb.h, p2.h): reads
// states from stdin, prints each permuted state twice (the host form of the
// partial rounds, then their device form), and checks the specialised
// divisions by powers of two against Montgomery products.
#include <cstdio>
#include <cstdlib>

#include "p2.h"

using namespace lg;

int main() {
  // kb_div_pow2<K>(a) == a * (2^-K) for a spread of a
  uint32_t inv2k[25];
  for (int k = 1; k <= 24; ++k) inv2k[k] = kb_to_mont(p2_pow_canon(p2_pow_canon(2, k), KB_P - 2));
  uint64_t x = 0x9E3779B97F4A7C15ull;
  int bad = 0;
  for (int it = 0; it < 200000; ++it) {
    x ^= x << 13; x ^= x >> 7; x ^= x << 17;
    uint32_t a = (uint32_t)(x % KB_P);
    if (it < 4) a = it == 0 ? 0u : (it == 1 ? KB_P - 1 : (it == 2 ? 1u : (1u << 24) - 1));
#define CHK(K) bad += kb_div_pow2<K>(a) != kb_mul(a, inv2k[K]);
    CHK(1) CHK(2) CHK(3) CHK(4) CHK(5) CHK(6) CHK(7) CHK(8) CHK(9) CHK(10) CHK(11) CHK(12)
    CHK(13) CHK(14) CHK(15) CHK(16) CHK(17) CHK(18) CHK(19) CHK(20) CHK(21) CHK(22) CHK(23) CHK(24)
#undef CHK
  }
  printf("div_pow2_mismatches %d\n", bad);
  // permutations: n states of 24 canonical values in, permuted canonical out
  int n = 0;
  if (scanf("%d", &n) != 1) return 1;
  for (int t = 0; t < n; ++t) {
    uint32_t s[24];
    for (int i = 0; i < 24; ++i) {
      unsigned v;
      if (scanf("%u", &v) != 1) return 1;
      s[i] = kb_to_mont(v);
    }
    uint32_t d[24];
    for (int i = 0; i < 24; ++i) d[i] = s[i];
    p2_permute<24>(s);
    p2_permute_form<24, true>(d);  // the device form of the partial rounds
    for (int i = 0; i < 24; ++i) printf("%u%c", kb_from_mont(s[i]), i == 23 ? '\n' : ' ');
    for (int i = 0; i < 24; ++i) printf("%u%c", kb_from_mont(d[i]), i == 23 ? '\n' : ' ');
  }
  return bad ? 2 : 0;
}
