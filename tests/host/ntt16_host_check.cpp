// Host run of the radix-16 NTT building blocks (ntt16.h) with the pass
// schedule of the LDS kernel: every pass visits the threads in turn.  Reads
// log2 n and n canonical values; prints the forward transform (bit-reversed
// order, natural input) and the inverse transform of the bit-reversed
// forward output (natural order, unscaled).
#include <cstdio>
#include <vector>

#include "ntt16.h"

using namespace lg;

static uint32_t root(int logn, bool inv) {
  // w_{2^24} = 3^127 (canonical), w_n = w_{2^24}^(2^24 / n)
  uint64_t w = 1, b = 3;
  for (uint32_t e = 127; e; e >>= 1, b = b * b % KB_P)
    if (e & 1) w = w * b % KB_P;
  for (int i = logn; i < 24; ++i) w = w * w % KB_P;
  uint32_t m = kb_to_mont((uint32_t)w);
  return inv ? kb_inv(m) : m;
}

int main() {
  int L;
  if (scanf("%d", &L) != 1) return 1;
  const uint32_t n = 1u << L, T = n / 16, q0 = n / 16;
  const int P = L / 4, r = L % 4;
  std::vector<uint32_t> in(n), sh(n), twf(n / 2), twi(n / 2);
  for (uint32_t i = 0; i < n; ++i) {
    unsigned v;
    if (scanf("%u", &v) != 1) return 1;
    in[i] = kb_to_mont(v);
  }
  const uint32_t wf = root(L, false), wi = root(L, true);
  twf[0] = twi[0] = KB_R;
  for (uint32_t e = 1; e < n / 2; ++e) {
    twf[e] = kb_mul(twf[e - 1], wf);
    twi[e] = kb_mul(twi[e - 1], wi);
  }
  std::vector<std::vector<uint32_t>> reg(T, std::vector<uint32_t>(16));
  // ---- forward: natural input in registers at t + k q0
  const NttRoots Rf = ntt16_roots(twf.data(), L);
  for (uint32_t t = 0; t < T; ++t)
    for (int k = 0; k < 16; ++k) reg[t][k] = in[t + k * q0];
  for (int p = 0; p < P; ++p) {
    const int logm = L - 4 * p;
    const uint32_t q = 1u << (logm - 4);
    for (uint32_t t = 0; t < T; ++t) {
      uint32_t* y = reg[t].data();
      if (p > 0)
        for (int k = 0; k < 16; ++k) y[k] = sh[ntt16_pos(t, q, k)];
      const Ladder Ld = ntt16_ladder(twf[(t % q) << (L - logm)], Rf);
      dif16(y, Ld);
      if (p < P - 1 || r)
        for (int k = 0; k < 16; ++k) sh[ntt16_pos(t, q, k)] = y[k];
    }
  }
  if (r)
    for (uint32_t t = 0; t < T; ++t) {
      uint32_t* y = reg[t].data();
      for (int k = 0; k < 16; ++k) y[k] = sh[16 * t + k];
      dif_small_r(y, r, Rf);
    }
  std::vector<uint32_t> fwd(n);
  for (uint32_t t = 0; t < T; ++t)
    for (int k = 0; k < 16; ++k) fwd[16 * t + k] = reg[t][k];
  for (uint32_t i = 0; i < n; ++i) printf("%u%c", kb_from_mont(fwd[i]), i + 1 == n ? '\n' : ' ');
  // ---- inverse of fwd (bit-reversed order): DIT, natural output
  const NttRoots Ri = ntt16_roots(twi.data(), L);
  for (uint32_t i = 0; i < n; ++i) sh[i] = fwd[i];  // already bit-reversed: X[brev(i)] at i
  if (r)
    for (uint32_t t = 0; t < T; ++t) {
      uint32_t y[16];
      for (int k = 0; k < 16; ++k) y[k] = sh[16 * t + k];
      dit_small_r(y, r, Ri);
      for (int k = 0; k < 16; ++k) sh[16 * t + k] = y[k];
    }
  for (int p = 0; p < P; ++p) {
    const int logm = r + 4 * (p + 1);
    const uint32_t q = 1u << (logm - 4);
    for (uint32_t t = 0; t < T; ++t) {
      uint32_t* x = reg[t].data();
      for (int k = 0; k < 16; ++k) x[k] = sh[ntt16_pos(t, q, k)];
      const Ladder Ld = ntt16_ladder(twi[(t % q) << (L - logm)], Ri);
      dit16(x, Ld);
      if (p < P - 1)
        for (int k = 0; k < 16; ++k) sh[ntt16_pos(t, q, k)] = x[k];
    }
  }
  std::vector<uint32_t> inv(n);
  for (uint32_t t = 0; t < T; ++t)
    for (int k = 0; k < 16; ++k) inv[t + k * q0] = reg[t][k];
  for (uint32_t i = 0; i < n; ++i) printf("%u%c", kb_from_mont(inv[i]), i + 1 == n ? '\n' : ' ');
  return 0;
}
