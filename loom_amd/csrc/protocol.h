// Protocol constants and domain helpers shared by the prover and the
// verifier: the FRI domain (RATE * 2^MAX_LOG_MODULE_SIZE,
// internal/constants/const.go:24-32), the transcript tags, the subgroup
// generators and the round challenge names.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "kb.h"
#include "plan.h"
#include "transcript.h"

namespace lg {

constexpr int RATE_LOG = 2;
constexpr int FRI_LOG_N = 24;  // RATE * 2^MAX_LOG_MODULE_SIZE
constexpr int FRI_LOG_D = 22;
constexpr int FRI_ROUNDS = FRI_LOG_D;
constexpr uint32_t TAG_HASH = 0x48415348u;
constexpr uint32_t TAG_SIZE = 0x53495a45u;
constexpr uint32_t TAG_EXTP = 0x45585450u;
constexpr uint32_t TAG_PUBI = 0x50554249u;
constexpr uint32_t MULT_GEN = 3;  // fft.GeneratorFullMultiplicativeGroup (canonical)

inline uint32_t M(uint32_t x) { return kb_to_mont(x); }

// w_{2^24} = g^((p-1)/2^24)
inline uint32_t omega_max() {
  static const uint32_t w = kb_pow(M(MULT_GEN), (KB_P - 1) >> 24);
  return w;
}

// w_{2^logn}, Montgomery
inline uint32_t omega(int logn) { return kb_pow(omega_max(), 1ull << (24 - logn)); }

inline uint32_t brev_host(uint32_t x, int bits) {
  uint32_t r = 0;
  for (int i = 0; i < bits; ++i) r |= ((x >> i) & 1u) << (bits - 1 - i);
  return r;
}

// Name of round r's challenge (prover/prover.go:208-212).
inline std::string challenge_name(size_t r) { return "challenge@loom_" + std::to_string(r); }

// Transcript elements of one instance's public inputs, bound to the first
// challenge after the setup roots (public/public_inputs.go:50-74): PUBI, the
// input count, then per input (by name) its name and module name as
// StringToElements(PUBI, .), the entry count and per entry [row, field.Base,
// value].  vals[i]: the entry values of public input i (Montgomery).
inline std::vector<uint32_t> public_input_words(const PlanData& d, const std::vector<const uint32_t*>& vals) {
  std::vector<uint32_t> w;
  if (d.publics.empty()) return w;
  if (vals.size() != d.publics.size()) throw std::runtime_error("public inputs: wrong number of value arrays");
  w.push_back(M(TAG_PUBI));
  w.push_back(M((uint32_t)d.publics.size()));
  for (size_t i = 0; i < d.publics.size(); ++i) {
    const PublicCol& pc = d.publics[i];
    auto a = string_to_elements(TAG_PUBI, pc.name), m = string_to_elements(TAG_PUBI, pc.module_name);
    w.insert(w.end(), a.begin(), a.end());
    w.insert(w.end(), m.begin(), m.end());
    w.push_back(M((uint32_t)pc.rows.size()));
    size_t o = 0;  // entry e's value words in vals[i]
    for (size_t e = 0; e < pc.rows.size(); ++e) {
      w.push_back(M(pc.rows[e]));
      w.push_back(M(pc.fields[e]));  // field.Base 0, field.Ext 1
      const size_t k = pc.fields[e] ? 6 : 1;
      w.insert(w.end(), vals[i] + o, vals[i] + o + k);
      o += k;
    }
  }
  return w;
}

}  // namespace lg
