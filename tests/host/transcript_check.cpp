// Host transcript of transcript.h against the oracle: reads challenge names
// with canonical binding elements, computes every challenge three ways (a
// transcript over a shared ChallengeSchedule by name, the same by position,
// and one registering names one by one) and prints the canonical digests of
// the first; a fourth transcript accepts those digests and a fifth checks
// them against its own computation, and both must log the same words; a
// changed digest must be refused.  Exits non-zero when the forms disagree.
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "transcript.h"

using namespace lg;

int main() {
  int n = 0;
  if (scanf("%d", &n) != 1) return 1;
  std::vector<std::string> names(n);
  std::vector<std::vector<uint32_t>> binds(n);
  auto sched = std::make_shared<ChallengeSchedule>();
  for (int i = 0; i < n; ++i) {
    char buf[256];
    int k = 0;
    if (scanf("%255s %d", buf, &k) != 2) return 1;
    names[i] = buf;
    for (int j = 0; j < k; ++j) {
      unsigned v;
      if (scanf("%u", &v) != 1) return 1;
      binds[i].push_back(kb_to_mont(v));
    }
    sched->add(names[i]);
  }
  Transcript by_name(sched), by_pos(sched), dyn, acc(sched), chk(sched);
  for (int i = 0; i < n; ++i) dyn.new_challenge(names[i]);
  int bad = 0;
  for (int i = 0; i < n; ++i) {
    by_name.bind(names[i], binds[i]);
    by_pos.bind_at(by_pos.index(names[i]), binds[i].data(), binds[i].size());
    dyn.bind(names[i], binds[i]);
    acc.bind(names[i], binds[i]);
    const Digest a = by_name.compute(names[i]);
    const Digest b = by_pos.compute_at(by_pos.index(names[i]));
    const Digest c = dyn.compute(names[i]);
    acc.accept(names[i], a.v);
    chk.bind(names[i], binds[i]);
    if (i % 2) chk.check_at(chk.index(names[i]), a.v);
    else chk.check(names[i], a.v);
    for (int l = 0; l < 8; ++l) {
      bad += a.v[l] != b.v[l] || a.v[l] != c.v[l];
      printf("%u%c", kb_from_mont(a.v[l]), l == 7 ? '\n' : ' ');
    }
  }
  bad += acc.log_words() != by_name.log_words();
  bad += chk.log_words() != by_name.log_words();
  // a digest that differs from the host's computation is refused
  if (n > 0) {
    Transcript wrong(sched);
    wrong.bind(names[0], binds[0]);
    uint32_t d[8];
    const Digest a = by_name.compute(names[0]);
    for (int l = 0; l < 8; ++l) d[l] = a.v[l];
    d[3] ^= 1u;
    bool threw = false;
    try {
      wrong.check(names[0], d);
    } catch (const std::runtime_error&) {
      threw = true;
    }
    bad += !threw;
  }
  bad += dyn.log_names() != by_name.log_names();
  fprintf(stderr, "forms_disagree %d\n", bad);
  return bad != 0;
}
