// E6 and E2 products of kb.h on random operands read from stdin (canonical
// values, one operand pair per line: 12 numbers); prints the canonical
// product limbs, for comparison with the oracle's schoolbook product.
#include <cstdio>

#include "kb.h"

int main() {
  using namespace lg;
  int n = 0;
  if (scanf("%d", &n) != 1) return 1;
  for (int k = 0; k < n; ++k) {
    E6 x, y;
    for (int i = 0; i < 6; ++i) {
      unsigned v;
      if (scanf("%u", &v) != 1) return 1;
      x.v[i] = kb_to_mont(v);
    }
    for (int i = 0; i < 6; ++i) {
      unsigned v;
      if (scanf("%u", &v) != 1) return 1;
      y.v[i] = kb_to_mont(v);
    }
    const E6 z = e6_mul(x, y);
    for (int i = 0; i < 6; ++i) printf("%u%c", kb_from_mont(z.v[i]), i == 5 ? '\n' : ' ');
  }
  return 0;
}
