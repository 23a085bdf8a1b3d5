// Montgomery product of kb.h against a 128-bit closed form: a*b*2^-32 mod p,
// with 2^-32 mod p from Fermat, on edge values and random pairs; and the
// S-box cube of p2.h (host form: one reduction by 2^64) against
// a^3*2^-64 mod p.  Prints the mismatch counts and exits non-zero if there is
// one.
#include <cstdio>
#include <random>

#include "p2.h"

int main() {
  using namespace lg;
  uint64_t base = (uint64_t)(((unsigned __int128)1 << 32) % KB_P), e = KB_P - 2, rinv = 1;
  while (e) {
    if (e & 1) rinv = rinv * base % KB_P;
    base = base * base % KB_P;
    e >>= 1;
  }
  long bad = 0;
  auto chk = [&](uint32_t a, uint32_t b) {
    const uint64_t want = (uint64_t)((unsigned __int128)a * b % KB_P * rinv % KB_P);
    if (kb_mul(a, b) != want) ++bad;
  };
  const uint32_t edge[] = {0u, 1u, 2u, KB_P - 1, KB_P - 2, 0x7effffffu, 0x10000000u, KB_R, KB_R2};
  for (uint32_t a : edge)
    for (uint32_t b : edge) chk(a, b);
  std::mt19937_64 g(1);
  for (int i = 0; i < 2000000; ++i) chk((uint32_t)(g() % KB_P), (uint32_t)(g() % KB_P));
  printf("kb_mul_mismatches %ld\n", bad);
  long bad3 = 0;
  auto chk3 = [&](uint32_t a) {
    const uint64_t a3 = (uint64_t)((unsigned __int128)a * a % KB_P * a % KB_P);
    const uint64_t want = (uint64_t)((unsigned __int128)a3 * rinv % KB_P * rinv % KB_P);
    if (kb_cube(a) != want) ++bad3;
  };
  for (uint32_t a : edge) chk3(a);
  for (uint32_t a = 0; a < 4096; ++a) chk3(a), chk3(KB_P - 1 - a);
  for (int i = 0; i < 2000000; ++i) chk3((uint32_t)(g() % KB_P));
  printf("kb_cube_mismatches %ld\n", bad3);
  return bad != 0 || bad3 != 0;
}
