// Prove phase on one HIP device for a batch of instances of one plan.
#pragma once
#include <hip/hip_runtime.h>

#include <atomic>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "dagjit.h"
#include "gpu.h"
#include "plan.h"
#include "kernels.h"

namespace lg {

// Per-device constant tables (twiddles, coset scalings), built on first use.
class Tables {
 public:
  ~Tables();
  const uint32_t* tw_fwd(int logn);
  const uint32_t* tw_inv(int logn);
  const uint32_t* wpow(int logn);                 // w_n^j, j < n
  const uint32_t* wpow_inv(int logn);             // w_n^-j, j < n
  const uint32_t* lde_pre(int logn, int b, int rate_log, bool from_values);
  const uint32_t* quot_pre(int logn, int i, int rho_log);   // (g w_{rho n}^i)^k / n
  const uint32_t* quot_post(int logn, int i, int rho_log);  // w_{rho n}^{-a i} / (rho n)
  const uint32_t* ginv_pow(int log_len);          // g^{-m}, m < 2^log_len
  const uint32_t* fri_xinv2(int log_m);           // (w_m^-1)^{bitrev_{m/2}(i)} / 2
  const uint32_t* gpow_small(int count);          // g^c, c < count
  // Radix-16 pass twiddles of the LDS NTT: for each sub-size m = 2^logm
  // (4 <= logm <= 14) and b < m/16, the 15 words of ntt16_ladder(w_m^b)
  // (w_m^-b when inverse) at 16 * (m/16 - 1 + b), padded to 16.
  const uint32_t* ntt_ladders(bool inverse);
  // The geometric scale tables (lde_pre, quot_pre: t[k] = sc s^k): their
  // ratio s, by device address; false for any other table.
  bool geometric(const uint32_t* t, uint32_t* s);
  // blk[p] = s^(2^b brev_a(p)), p < 2^a: the block factors of a geometric
  // scale table in the fused four-step LDE (run_ntt).
  const uint32_t* geo_blocks(uint32_t s, int a, int b);

 private:
  const uint32_t* get(const std::string& key, const std::function<std::vector<uint32_t>()>& make);
  const uint32_t* geo(const std::string& key, uint32_t s, size_t n, uint32_t sc);
  std::mutex mu_;
  std::map<std::string, DevBuf> bufs_;
  std::mutex geo_mu_;
  std::map<const uint32_t*, uint32_t> geo_;
  hipStream_t stream_ = nullptr;
};

// Transforms of NttCfg semantics (kernels.h) on every job, at any size up to
// 2^22: one LDS-resident launch up to 2^14, else the four-step two-pass form
// with scratch(words) device scratch.  stage(p, bytes) returns a device copy
// of a small host table valid for the launches that follow on s.
using StageFn = std::function<const void*(const void*, size_t)>;
using ScratchFn = std::function<uint32_t*(size_t)>;
void run_ntt(Tables& tb, const NttCfg& cfg, const std::vector<NttJob>& jobs, hipStream_t s, const StageFn& stage,
             const ScratchFn& scratch);

// Reusable device buffers (by exact size) and pinned staging areas, so that
// repeated proofs of one shape allocate nothing after the first.
class Pool {
 public:
  DevBuf get(size_t bytes);
  void put(DevBuf&& b);
  // Frees cached buffers, largest first, until at most keep_bytes remain.
  void trim(size_t keep_bytes);
  ~Pool();

 private:
  std::mutex mu_;
  std::multimap<size_t, DevBuf> free_;
  size_t cached_ = 0;
};

// Pinned host staging + matching device area: small tables (pointer arrays,
// challenges, positions) go host -> device with asynchronous copies.
class Staging {
 public:
  explicit Staging(size_t bytes);
  ~Staging();
  Staging(const Staging&) = delete;
  Staging& operator=(const Staging&) = delete;
  void reset() { off_ = flushed_ = 0; }
  // Appends n bytes to the pinned area and returns the device address they
  // will have after the next flush.
  void* put(const void* src, size_t n, hipStream_t s);
  // Copies everything appended since the last flush to the device (one copy).
  void flush(hipStream_t s);
  // Marks everything appended as copied (a replayed graph holds the copies).
  void assume_flushed();
  // flush() on `side` (a stream no graph records), waited for.
  void flush_sync(hipStream_t side);
  // While verifying, put() notes whether any appended bytes differ from what
  // the area already held at that offset (a replayed graph reads the bytes
  // recorded with it, so they must not change).
  void set_verify(bool on) {
    verify_ = on;
    changed_ = false;
  }
  bool changed() const { return changed_; }
  template <class T>
  const T* put(const std::vector<T>& v, hipStream_t s) {
    return static_cast<const T*>(put(v.data(), v.size() * sizeof(T), s));
  }

 private:
  uint8_t* host_ = nullptr;
  DevBuf dev_;
  size_t cap_ = 0, off_ = 0, flushed_ = 0;
  bool verify_ = false, changed_ = false;
};

// Optional per-kernel timing with HIP events on the launching stream: every
// launch the prover wraps is bracketed by two events; totals per kernel name
// (launch count, device milliseconds, algorithmic bytes, Poseidon2
// permutations) are accumulated once the stream has been synchronised.
class KernelTimer {
 public:
  bool on = false;
  struct Stat {
    uint64_t count = 0;
    double ms = 0;
    double bytes = 0;
    double perms = 0;
  };
  // begin() opens a record whose start event is recorded at the next launch
  // point (after that launch's staging copy), so a record times kernels only.
  void begin(const char* name, hipStream_t s, double bytes, double perms);
  void launch_point(hipStream_t s);
  void end(hipStream_t s);
  void resolve();  // call after the stream is synchronised
  std::map<std::string, Stat> stats;
  ~KernelTimer();

 private:
  struct Rec {
    std::string name;
    hipEvent_t a, b;
    double bytes, perms;
  };
  hipEvent_t ev();
  std::vector<Rec> pending_;
  std::vector<hipEvent_t> free_;
  bool start_pending_ = false;
};

struct Context {
  KernelTimer timer;
  int device = 0;
  hipStream_t stream = nullptr;
  Tables tables;
  Pool pool;
  std::mutex staging_mu;
  std::vector<std::unique_ptr<Staging>> staging_free;
  std::unique_ptr<Staging> take_staging();
  void give_staging(std::unique_ptr<Staging> s);
  // A batch is split over `lanes` sub-batches proved by concurrent host
  // threads, each on its own stream (lane 0 uses `stream`), so that one
  // lane's host work (transcripts, query derivation, proof assembly)
  // overlaps another lane's kernels.  LOOMGPU_LANES overrides the default.
  int lanes = 2;
  bool lanes_forced = false;  // LOOMGPU_LANES set: overrides the tuned lane count
  hipStream_t lane_stream(int lane);
  // a non-blocking stream for one-off copies beside the lanes' work
  hipStream_t side_stream();
  // Pinned host buffers per lane, reused across proves.
  uint32_t* pinned(size_t bytes, int lane);     // device -> host (query openings)
  uint32_t* pinned_in(size_t bytes, int lane);  // host -> device (trace columns)
  uint32_t* pinned_aux(size_t bytes, int lane);  // small device -> host results (claimed values, digests)
  uint32_t* pinned_fs(size_t bytes, int lane);   // transcript data (round digests, roots, logup totals)
  uint32_t* pinned_ext(size_t bytes, int lane, size_t idx);  // device-phase results, idx-th buffer of the lane
  // Record proves into HIP graphs per shape bucket (LOOMGPU_GRAPHS=0: off).
  bool graphs = true;
  Context();
  ~Context();

 private:
  struct LaneBufs {
    hipStream_t stream = nullptr;  // owned for lanes >= 1
    uint32_t* pinned = nullptr;
    size_t pinned_cap = 0;
    uint32_t* pinned_in = nullptr;
    size_t pinned_in_cap = 0;
    uint32_t* pinned_aux = nullptr;
    size_t pinned_aux_cap = 0;
    uint32_t* pinned_fs = nullptr;
    size_t pinned_fs_cap = 0;
    std::vector<std::pair<uint32_t*, size_t>> ext;
  };
  std::mutex lane_mu_;
  std::vector<LaneBufs> lane_;
  hipStream_t side_ = nullptr;
};

// Prove options: SkipFRI (prover.SkipFRI, prover/prover.go:78-83): the
// proof carries the claimed values and no opening.
constexpr uint32_t PROVE_SKIP_FRI = 1u;

// Per-instance caller data.
struct InstanceInput {
  std::vector<const uint32_t*> columns;  // plan input order, Montgomery, N_module words each
  std::vector<const uint32_t*> perms;    // plan permutation order, len words each
  bool device = false;                   // pointers are device memory of the context's device (else host)
  uint32_t* proof_dst = nullptr;         // optional caller buffer for the proof words
  size_t proof_cap = 0;                  // its capacity in words
  // prover.WithTranscript with a running transcript: the digest of its last
  // (computed) challenge, which the prove's first challenge follows
  bool chained = false;
  uint32_t fs_prev[8] = {0};
};

struct ProveOutput {
  std::vector<uint32_t> proof;       // canonical word stream (when no caller buffer was given)
  size_t proof_len = 0;              // proof length in words
  bool direct = false;               // written to the caller's buffer (complete if proof_len <= cap)
  std::vector<uint32_t> transcript;  // challenge digests in computation order
  // Non-empty when this instance failed (a lookup value outside its table, a
  // permutation entry outside its cells): no proof was written.  The other
  // instances of the batch are unaffected.
  std::string error;
};

// Prover-side timing of each phase (wall clock, host), for the phase callback.
using PhaseCallback = std::function<void(const char*, double)>;

struct ChallengeSchedule;
struct SetupData;

// One shape bucket of a plan (batch size, lane and whether the inputs are
// host or device memory): the device buffers,
// staging area and pinned buffers of its prove in allocation order, so that
// every prove of the bucket uses the same addresses, and the instantiated
// HIP graph of the prove's device phase (generated columns through the FRI
// commit phase).
struct GraphBucket {
  int batch = 0, lane = 0;
  std::vector<DevBuf> bufs;
  std::vector<std::pair<uint32_t*, size_t>> pins;
  std::unique_ptr<Staging> stg;
  hipGraphExec_t exec = nullptr;
  int proves = 0;       // proves of this shape seen (the graph is recorded on the second)
  uint64_t last = 0;    // use stamp (least recently used buckets are dropped)
  size_t bytes = 0;     // device bytes held
  bool broken = false;  // a recording or replay failed: prove this shape eagerly
  bool busy = false;
  ~GraphBucket();
};

class Plan {
 public:
  Plan(Context* ctx, const uint32_t* words, size_t n);
  ~Plan();
  // setup.Setup (setup/setup.go:68-124) on the device: commits the setup
  // columns (host, N words each, in setup order) into the setup tree once and
  // keeps columns, LDE and tree resident for every later prove of the plan.
  // root: 8 words (Montgomery), the verification key.
  void set_setup(const uint32_t* const* cols, size_t n_cols, uint32_t* root);
  const SetupData* setup() const { return setup_.get(); }
  const PlanData& data() const { return d_; }
  Context* ctx() const { return ctx_; }
  // Runs the prove phase for every instance; outputs are resized to batch.
  // An instance whose witness is invalid fails alone (ProveOutput::error);
  // a failure of the plan or the device throws for the whole batch.
  void prove(const std::vector<InstanceInput>& inputs, std::vector<ProveOutput>& outputs,
             const PhaseCallback& cb = nullptr);
  // input arrays per instance: the input columns, then each public input's
  // entry values
  size_t num_inputs() const { return d_.inputs.size() + d_.publics.size(); }
  size_t num_perms() const { return d_.perms.size(); }

  std::shared_ptr<DagKernel> relation_kernel(size_t m) const { return rel_[m]; }
  std::shared_ptr<DagKernel> step_kernel(size_t r, size_t s, size_t k) const { return steps_.at({r, s, k}); }
  // the Fiat-Shamir challenges of one proof of this plan, in order
  std::shared_ptr<const ChallengeSchedule> schedule() const { return sched_; }

  // Launch choices for proves of `batch` instances (defaults until tuned) and
  // the batch of the prove in progress.
  Tuning tuning(int batch) const;
  int tuning_batch() const { return cur_batch_; }
  void set_tuning(int batch, const Tuning& t);
  // The autotuner: times whole proves of these instances (as they run in
  // production: graph replays when graphs are on) over a
  // coordinate descent of the Tuning fields, starting from the current
  // choices, and keeps the fastest for this batch size.  With
  // LOOMGPU_TUNE_FILE set the result is appended to that table, which plan
  // creation reads back.  Returns the choice; ms[0] = the starting
  // configuration's median prove time, ms[1] = the chosen one's.
  Tuning autotune(const std::vector<InstanceInput>& inputs, int reps, double* ms, int* trials);

 private:
  Context* ctx_;
  PlanData d_;
  std::vector<std::shared_ptr<DagKernel>> rel_;
  std::map<std::tuple<size_t, size_t, size_t>, std::shared_ptr<DagKernel>> steps_;
  std::shared_ptr<const ChallengeSchedule> sched_;
  std::unique_ptr<SetupData> setup_;
  std::map<int, Tuning> tuned_;
  int cur_batch_ = 0;
  uint64_t hash_ = 0;  // of the plan words (tuning table key)
  // Graph buckets by (batch, lane); a plan whose modules form one DEEP class
  // (its device phase needs no host round trip) is eligible.
  bool graph_eligible_ = false;
  std::mutex gmu_;
  std::map<std::tuple<int, int, bool>, std::unique_ptr<GraphBucket>> buckets_;  // (batch, lane, device inputs)
  uint64_t tick_ = 0;
  uint32_t options_ = 0;
  uint32_t grinding_ = 0;
  GraphBucket* take_bucket(int batch, int lane, bool device_inputs, int* mode);
  void give_bucket(GraphBucket* b);

 public:
  static constexpr int MAX_GRAPH_BUCKETS = 4;
  // Device bytes held by this plan's graph buckets.
  size_t graph_bytes();
  // Drops every graph bucket (frees its buffers and graph).
  void drop_graphs();
  // Prove options (PROVE_* flags), for the proves that follow.
  uint32_t options() const { return options_; }
  void set_options(uint32_t flags) { options_ = flags; }
  // Proof-of-work bits of every FRI fold challenge (fri.WithGrinding,
  // internal/fri/fri.go:71-78); 0: off, the default.
  uint32_t grinding() const { return grinding_; }
  void set_grinding(uint32_t bits);
  // Proves of this plan that recorded a graph, replayed one, or ran again
  // eagerly after a failed recording or replay.
  std::atomic<uint64_t> graph_records{0}, graph_replays{0}, graph_fallbacks{0};
};

}  // namespace lg
