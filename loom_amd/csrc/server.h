// Serving loop on one device: proof requests of any plan and batch size are
// queued, admitted onto proving slots (one context and stream each) when a
// slot is free and the request's HBM footprint fits beside the requests
// already running, and proved concurrently.  Queued requests of the same plan
// are merged into one batched prove when a slot takes them (continuous
// batching): instances are independent, so each proof is the one it would be
// alone.
#pragma once
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "prover.h"

namespace lg {

// Device bytes one instance of the plan needs while it is proved: committed
// columns (Lagrange + rate-4 LDE), other columns, Merkle trees, the largest
// module's quotient buffers, DEEP codewords and FRI layers.  An estimate for
// admission, not an allocator bound.
size_t plan_bytes_per_instance(const PlanData& d);

class Server {
 public:
  struct Stats {
    uint64_t submitted = 0, completed = 0, failed = 0, queued = 0, running = 0;
    uint64_t peak_running = 0, peak_bytes = 0, proves = 0, merged = 0;
  };

  Server(int device, int slots, size_t budget_bytes, int max_batch);
  ~Server();

  // Copies the plan words and the inputs; returns the request's ticket.
  uint64_t submit(const uint32_t* words, size_t n_words, int batch, const uint32_t* const* columns,
                  const uint32_t* const* perms);
  // Waits up to timeout_ms (< 0: no limit).  Returns false on timeout.  On
  // completion out receives the proofs (or err the failure) and the request
  // is forgotten unless keep is set.
  bool wait(uint64_t ticket, int timeout_ms, std::vector<ProveOutput>* out, std::string* err, bool keep);
  Stats stats();
  size_t budget() const { return budget_; }

 private:
  struct PlanEntry {
    std::vector<uint32_t> words;
    PlanData data;
    size_t bytes_per_instance = 0;
  };
  struct Request {
    uint64_t ticket = 0;
    std::shared_ptr<PlanEntry> plan;
    int batch = 0;
    std::vector<uint32_t> data;  // inputs, copied at submit
    std::vector<InstanceInput> inputs;
    size_t bytes = 0;
    bool done = false;
    std::string error;
    std::vector<ProveOutput> out;
  };
  struct Slot {
    std::unique_ptr<Context> ctx;
    std::map<const PlanEntry*, std::unique_ptr<Plan>> plans;
  };

  void worker(int slot);
  std::shared_ptr<PlanEntry> plan_entry(const uint32_t* words, size_t n);

  int device_;
  size_t budget_;
  int max_batch_;
  std::vector<Slot> slots_;
  std::vector<std::thread> threads_;
  std::mutex mu_;
  std::condition_variable work_cv_, done_cv_;
  std::deque<std::shared_ptr<Request>> queue_;
  std::map<uint64_t, std::shared_ptr<Request>> requests_;
  std::map<std::string, std::shared_ptr<PlanEntry>> plans_;
  uint64_t next_ticket_ = 1;
  size_t in_flight_ = 0;
  bool stop_ = false;
  Stats st_;
};

}  // namespace lg
