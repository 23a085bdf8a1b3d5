// Proof checking for many proofs of one plan (verifier.Verify,
// verifier/verifier.go:479-533, and the PCS check it ends with,
// internal/fri/verify.go:60-157): the host decodes each proof, replays its
// transcript and checks the buses and the AIR identity at zeta; the device
// hashes every opened row and authenticates every Merkle path of the batch
// (trace, setup and AIR trees with their level injections, the FRI levels
// tree and the folded layers); the host checks the DEEP bridge and the FRI
// folds meanwhile.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace lg {

class Plan;

// Which check a proof failed (0: none).
enum VerifyStatus : int {
  VERIFY_OK = 0,
  VERIFY_FORMAT = 1,   // proof words do not parse, or their shape differs from the plan's
  VERIFY_BUS = 2,      // a logup bus does not sum to zero
  VERIFY_AIR = 3,      // a module's relation at zeta differs from (zeta^N - 1) Q(zeta)
  VERIFY_OPENING = 4,  // an opened row does not authenticate against its commitment
  VERIFY_DEEP = 5,     // a DEEP quotient recomputed at a query differs from the FRI level value
  VERIFY_FRI = 6,      // a FRI layer opening, fold or final-polynomial check fails
};

struct VerifyInput {
  const uint32_t* words = nullptr;  // canonical proof words
  size_t len = 0;
  std::vector<const uint32_t*> publics;  // per public-input column of the plan: its entry values (Montgomery)
  // verifier.WithTranscript with a running transcript: its last digest, which
  // the proof's first challenge follows (as in the prove)
  bool chained = false;
  uint32_t fs_prev[8] = {0};
};

struct VerifyResult {
  int status = VERIFY_OK;
  std::string message;
};

// setup_roots: the verification key, 8 Montgomery words per setup tree.
// skip_fri: stop after the AIR check (verifier.SkipFRI, verifier.go:52-57).
void verify_batch(Plan& plan, const std::vector<VerifyInput>& in, const uint32_t* setup_roots, size_t n_setup_roots,
                  bool skip_fri, std::vector<VerifyResult>& out);

}  // namespace lg
