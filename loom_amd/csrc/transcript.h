// Fiat-Shamir transcript on the host (internal/fiat-shamir/transcript.go:102-315):
// challenges are registered in order; a challenge's digest is the Poseidon2
// sponge of StringToElements(FSID, name) || the previous challenge's digest
// (except for the first challenge) || its bindings.  Grinding is not used
// (the reference's default).  All elements are Montgomery-form.
#pragma once
#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "p2.h"

namespace lg {

constexpr uint32_t TAG_FSID = 0x46534944u;
constexpr uint32_t TAG_FSPW = 0x46535057u;  // "FSPW": before a proof of work's bit count and salt

// hash.StringToElements (internal/hash/hash.go:36-53), Montgomery form.
inline std::vector<uint32_t> string_to_elements(uint32_t tag, const std::string& s) {
  std::vector<uint32_t> r;
  r.push_back(kb_to_mont(tag));
  r.push_back(kb_to_mont((uint32_t)s.size()));
  for (size_t i = 0; i < s.size(); i += 3) {
    uint32_t limb = 0;
    for (size_t j = 0; j < 3 && i + j < s.size(); ++j) limb |= (uint32_t)(uint8_t)s[i + j] << (8 * j);
    r.push_back(kb_to_mont(limb));
  }
  return r;
}

struct Digest {
  uint32_t v[8];
};

// The bindings of one challenge: up to 8 words inline (a digest, the common
// case of the FRI query chain), longer runs on the heap.  A proof's hundreds
// of short bindings then cost no allocation to build or to free.
class Words {
 public:
  void append(const uint32_t* e, size_t n) {
    if (heap_.empty() && n_ + n <= kInline) {
      std::copy(e, e + n, inl_ + n_);
    } else {
      if (heap_.empty()) heap_.assign(inl_, inl_ + n_);
      heap_.insert(heap_.end(), e, e + n);
    }
    n_ += n;
  }
  size_t size() const { return n_; }
  const uint32_t* begin() const { return heap_.empty() ? inl_ : heap_.data(); }
  const uint32_t* end() const { return begin() + n_; }

 private:
  static constexpr size_t kInline = 8;
  size_t n_ = 0;
  uint32_t inl_[kInline] = {};
  std::vector<uint32_t> heap_;
};

// Challenge names in registration order with their message prefixes
// (string_to_elements(FSID, name)).  A prover builds the schedule of a proof
// shape once and shares it between the transcripts of all its proofs.
struct ChallengeSchedule {
  std::vector<std::string> names;
  std::map<std::string, size_t> pos;
  std::vector<std::vector<uint32_t>> prefix;
  void add(const std::string& name) {
    if (pos.count(name)) throw std::runtime_error("transcript: challenge already exists: " + name);
    pos[name] = names.size();
    names.push_back(name);
    prefix.push_back(string_to_elements(TAG_FSID, name));
  }
};

class Transcript {
 public:
  Transcript() = default;
  // A transcript with the schedule's challenges registered, in its order.
  explicit Transcript(std::shared_ptr<const ChallengeSchedule> s) : fixed_(std::move(s)) {
    const size_t n = fixed_->names.size();
    bindings_.resize(n);
    values_.resize(n);
    computed_.assign(n, false);
    pow_.assign(n, {0u, 0u});
  }
  void new_challenge(const std::string& name) {
    if (fixed_) throw std::runtime_error("transcript: the challenge schedule is fixed: " + name);
    own_.add(name);
    bindings_.emplace_back();
    values_.emplace_back();
    computed_.push_back(false);
    pow_.push_back({0u, 0u});
  }
  // A proof of work for the challenge at p (ComputeChallenge WithGrinding,
  // transcript.go:216-332): its message ends with [FSPW, bits, salt] (salt
  // canonical, below p) and its digest must have `bits` zero low bits
  // (hasZeroGrindingBits, :349-365); computing it otherwise is an error.
  void set_pow(size_t p, uint32_t bits, uint32_t salt) {
    if (p >= computed_.size() || computed_[p]) throw std::runtime_error("transcript: bad proof-of-work position");
    if (bits < 1 || bits > 31 || salt >= KB_P) throw std::runtime_error("transcript: invalid proof of work");
    pow_[p] = {bits, salt};
  }
  // prover.WithTranscript / verifier.WithTranscript with a running transcript
  // whose last challenge is computed: this transcript's challenges follow it,
  // so the first one's message carries that digest after its name, as every
  // later challenge carries its predecessor's (transcript.go:285-315).
  void chain_from(const uint32_t* digest) {
    if (!computed_.empty() && computed_[0]) throw std::runtime_error("transcript: chained after a computation");
    chained_ = true;
    std::copy(digest, digest + 8, prev0_.v);
  }
  static bool zero_grinding_bits(const uint32_t* digest, uint32_t bits) {
    for (int i = 0; i < 8 && bits > 0; ++i) {
      const uint32_t k = bits < 31 ? bits : 31;
      if (kb_from_mont(digest[i]) & ((1u << k) - 1u)) return false;
      bits -= k;
    }
    return true;
  }
  void bind(const std::string& name, const uint32_t* e, size_t n) {
    size_t p = at(name);
    if (computed_[p]) throw std::runtime_error("transcript: challenge already computed: " + name);
    bindings_[p].append(e, n);
  }
  void bind(const std::string& name, const std::vector<uint32_t>& e) { bind(name, e.data(), e.size()); }
  const Digest& compute(const std::string& name) {
    size_t p = at(name);
    if (computed_[p]) return values_[p];
    std::vector<uint32_t> msg = sch().prefix[p];
    if (p != 0) {
      if (!computed_[p - 1]) throw std::runtime_error("transcript: previous challenge not computed: " + name);
      msg.insert(msg.end(), values_[p - 1].v, values_[p - 1].v + 8);
    } else if (chained_) {
      msg.insert(msg.end(), prev0_.v, prev0_.v + 8);
    }
    msg.insert(msg.end(), bindings_[p].begin(), bindings_[p].end());
    if (pow_[p].first) {
      const uint32_t t[3] = {kb_to_mont(TAG_FSPW), kb_to_mont(pow_[p].first), kb_to_mont(pow_[p].second)};
      msg.insert(msg.end(), t, t + 3);
    }
    hash(msg.data(), msg.size(), values_[p].v);
    check_pow(p);
    computed_[p] = true;
    log_.push_back(p);
    return values_[p];
  }
  // The function a challenge's message is hashed with (the transcript
  // hasher, fiatshamir.NewTranscriptHasher, transcript.go:74): the Poseidon2
  // sponge unless one is set (a test records the messages with it).
  using Hasher = void (*)(const uint32_t* msg, size_t len, uint32_t out[8], void* user);
  void set_hasher(Hasher h, void* user) {
    hasher_ = h;
    hasher_user_ = user;
  }
  // Position-based forms for long runs of consecutive challenges (the FRI
  // queries).
  size_t index(const std::string& name) const { return at(name); }
  void bind_at(size_t p, const uint32_t* e, size_t n) {
    if (p >= computed_.size() || computed_[p]) throw std::runtime_error("transcript: bad bind position");
    bindings_[p].append(e, n);
  }
  const Digest& compute_at(size_t p) {
    if (p >= computed_.size()) throw std::runtime_error("transcript: bad challenge position");
    if (computed_[p]) return values_[p];
    const std::vector<uint32_t>& prefix = sch().prefix[p];
    if (p != 0 && !computed_[p - 1])
      throw std::runtime_error("transcript: previous challenge not computed: " + sch().names[p]);
    uint32_t small[64];
    const size_t len = prefix.size() + (p != 0 || chained_ ? 8 : 0) + bindings_[p].size() + (pow_[p].first ? 3 : 0);
    std::vector<uint32_t> big;
    uint32_t* msg = small;
    if (len > 64) {
      big.resize(len);
      msg = big.data();
    }
    size_t o = 0;
    for (uint32_t x : prefix) msg[o++] = x;
    if (p != 0)
      for (int i = 0; i < 8; ++i) msg[o++] = values_[p - 1].v[i];
    else if (chained_)
      for (int i = 0; i < 8; ++i) msg[o++] = prev0_.v[i];
    for (uint32_t x : bindings_[p]) msg[o++] = x;
    if (pow_[p].first) {
      msg[o++] = kb_to_mont(TAG_FSPW);
      msg[o++] = kb_to_mont(pow_[p].first);
      msg[o++] = kb_to_mont(pow_[p].second);
    }
    hash(msg, len, values_[p].v);
    check_pow(p);
    computed_[p] = true;
    log_.push_back(p);
    return values_[p];
  }
  // The digest of the challenge registered just before `name` (the "previous
  // value" its message starts with); it must have been computed.
  const Digest& previous_value(const std::string& name) const {
    const size_t p = at(name);
    if (p == 0 && chained_) return prev0_;
    if (p == 0 || !computed_[p - 1]) throw std::runtime_error("transcript: no computed predecessor of " + name);
    return values_[p - 1];
  }
  // Computes `name` and checks it against a digest computed elsewhere (the
  // device) from the same message; a mismatch is an error.
  const Digest& check(const std::string& name, const uint32_t* digest) {
    const Digest& d = compute(name);
    for (int i = 0; i < 8; ++i)
      if (d.v[i] != digest[i]) throw std::runtime_error("transcript: device digest differs for " + name);
    return d;
  }
  const Digest& check_at(size_t p, const uint32_t* digest) {
    const Digest& d = compute_at(p);
    for (int i = 0; i < 8; ++i)
      if (d.v[i] != digest[i]) throw std::runtime_error("transcript: device digest differs for " + sch().names[p]);
    return d;
  }
  // Records the digest of `name` as computed elsewhere (the device) from this
  // transcript's message for it (string_to_elements(FSID, name) || previous
  // digest || bindings); the same order checks as compute().
  void accept(const std::string& name, const uint32_t* digest) {
    size_t p = at(name);
    if (computed_[p]) throw std::runtime_error("transcript: challenge already computed: " + name);
    if (p != 0 && !computed_[p - 1]) throw std::runtime_error("transcript: previous challenge not computed: " + name);
    for (int i = 0; i < 8; ++i) values_[p].v[i] = digest[i];
    computed_[p] = true;
    log_.push_back(p);
  }
  // Digests in computation order.
  std::vector<uint32_t> log_words() const {
    std::vector<uint32_t> w;
    for (size_t p : log_) w.insert(w.end(), values_[p].v, values_[p].v + 8);
    return w;
  }
  std::vector<std::string> log_names() const {
    std::vector<std::string> r;
    for (size_t p : log_) r.push_back(sch().names[p]);
    return r;
  }

 private:
  const ChallengeSchedule& sch() const { return fixed_ ? *fixed_ : own_; }
  void hash(const uint32_t* msg, size_t len, uint32_t out[8]) {
    if (hasher_) hasher_(msg, len, out, hasher_user_);
    else p2_sponge_host(msg, len, out);
  }
  void check_pow(size_t p) const {
    if (pow_[p].first && !zero_grinding_bits(values_[p].v, pow_[p].first))
      throw std::runtime_error("transcript: invalid proof of work for " + sch().names[p]);
  }
  Hasher hasher_ = nullptr;
  void* hasher_user_ = nullptr;
  size_t at(const std::string& name) const {
    const auto& pos = sch().pos;
    auto it = pos.find(name);
    if (it == pos.end()) throw std::runtime_error("transcript: unknown challenge " + name);
    return it->second;
  }
  ChallengeSchedule own_;                           // challenges registered one by one
  std::shared_ptr<const ChallengeSchedule> fixed_;  // or a shared schedule
  std::vector<Words> bindings_;
  std::vector<Digest> values_;
  std::vector<bool> computed_;
  std::vector<std::pair<uint32_t, uint32_t>> pow_;  // (bits, salt) per position, bits 0: none
  bool chained_ = false;  // the first challenge follows a caller's computed one
  Digest prev0_{};
  std::vector<size_t> log_;
};

}  // namespace lg
