// The host transcript with a recording hasher (the recordingHasher of
// internal/fiat-shamir/transcript_test.go:13-41): every message it is asked
// to hash is printed (canonical values, one line per message) and the k-th
// digest returned is (100 k + i + 1) for i < 8.  Two transcripts run the
// scenario of TestComputeChallengeAbsorbsChallengeID (transcript_test.go:
// 43-78): one registering names one by one and computing by name, one over
// a shared schedule computing by position.
#include <cstdio>
#include <memory>
#include <vector>

#include "transcript.h"

using namespace lg;

namespace {
struct Recorder {
  std::vector<std::vector<uint32_t>> inputs;
};

void record(const uint32_t* msg, size_t len, uint32_t out[8], void* user) {
  auto* r = static_cast<Recorder*>(user);
  r->inputs.emplace_back(msg, msg + len);
  for (int i = 0; i < 8; ++i) out[i] = kb_to_mont((uint32_t)(100 * r->inputs.size() + i + 1));
}

void dump(const Recorder& r) {
  for (auto& in : r.inputs) {
    for (size_t i = 0; i < in.size(); ++i) printf("%u%c", kb_from_mont(in[i]), i + 1 == in.size() ? '\n' : ' ');
  }
}
}  // namespace

int main() {
  const uint32_t zeta_binding = kb_to_mont(11), alpha_binding = kb_to_mont(22);
  {
    Recorder rec;
    Transcript t;
    t.set_hasher(record, &rec);
    t.new_challenge("zeta");
    t.new_challenge("alpha_DEEP");
    t.bind("zeta", &zeta_binding, 1);
    t.compute("zeta");
    t.bind("alpha_DEEP", &alpha_binding, 1);
    t.compute("alpha_DEEP");
    dump(rec);
  }
  {
    Recorder rec;
    auto s = std::make_shared<ChallengeSchedule>();
    s->add("zeta");
    s->add("alpha_DEEP");
    Transcript t(s);
    t.set_hasher(record, &rec);
    t.bind_at(0, &zeta_binding, 1);
    t.compute_at(0);
    t.bind_at(1, &alpha_binding, 1);
    t.compute_at(1);
    dump(rec);
  }
  return 0;
}
