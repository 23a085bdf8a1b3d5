// Serving loop: admission by HBM footprint, one proving slot per worker
// thread (own context and stream), continuous batching of same-plan requests.
#include "server.h"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>

#include "gpu.h"

namespace lg {

size_t plan_bytes_per_instance(const PlanData& d) {
  std::vector<char> committed(d.cols.size(), 0), is_total(d.cols.size(), 0);
  for (auto& t : d.trees)
    for (auto& g : t.groups) {
      for (auto c : g.base) committed[c] = 1;
      for (auto c : g.ext) committed[c] = 1;
    }
  for (auto& x : d.exposed) is_total[x.col] = 1;
  size_t words = 0;
  for (size_t c = 0; c < d.cols.size(); ++c) {
    const ColInfo& ci = d.cols[c];
    if (ci.src == SRC_CHALLENGE || ci.src == SRC_SETUP || is_total[c]) continue;  // setup: shared by the plan
    words += (size_t)ci.limbs() * d.modules[ci.module].N * (committed[c] ? 5 : 1);
  }
  size_t nmax = 0;
  for (auto& m : d.modules) nmax = std::max<size_t>(nmax, m.N);
  for (auto& t : d.trees) {
    size_t top = 0, inj = 0;
    for (auto& g : t.groups) top = std::max<size_t>(top, g.N);
    for (auto& g : t.groups) inj += 2 * (size_t)g.N * 8;
    words += 2 * (2 * top) * 8 + inj;  // node array of the largest group + injection digests
  }
  size_t quot = 0;
  for (size_t mi = 0; mi < d.modules.size(); ++mi) {
    const ModuleInfo& m = d.modules[mi];
    size_t limbs = 0;
    for (auto& ci : d.cols)
      if (ci.module == mi && ci.src != SRC_CHALLENGE) limbs += ci.limbs();
    const size_t rn = (size_t)m.n_chunks * m.N;
    quot = std::max(quot, limbs * rn + 3 * 6 * rn);
  }
  words += quot;
  words += (size_t)d.modules.size() * 0 + 6 * 5 * nmax;           // DEEP codeword (trace domain + LDE)
  words += 2 * 6 * 4 * nmax + 2 * (2 * 2 * nmax) * 8;             // FRI running layers + layer tree
  for (auto c : d.inputs) words += d.modules[d.cols[c].module].N;  // uploaded inputs
  for (auto& p : d.perms) words += p.len;
  return words * 4;
}

Server::Server(int device, int slots, size_t budget_bytes, int max_batch)
    : device_(device), budget_(budget_bytes), max_batch_(std::max(1, max_batch)) {
  if (slots < 1 || slots > 16) throw std::runtime_error("server: slots must be in [1, 16]");
  int n = 0;
  LG_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) throw std::runtime_error("server: no such HIP device");
  LG_HIP(hipSetDevice(device));
  if (budget_ == 0) {
    size_t fr = 0, tot = 0;
    LG_HIP(hipMemGetInfo(&fr, &tot));
    budget_ = fr / 10 * 8;
  }
  slots_.resize(slots);
  for (auto& s : slots_) {
    s.ctx = std::make_unique<Context>();
    s.ctx->device = device;
    LG_HIP(hipStreamCreateWithFlags(&s.ctx->stream, hipStreamNonBlocking));
  }
  for (int i = 0; i < slots; ++i) threads_.emplace_back([this, i] { worker(i); });
}

Server::~Server() {
  {
    std::lock_guard<std::mutex> lk(mu_);
    stop_ = true;
  }
  work_cv_.notify_all();
  for (auto& t : threads_) t.join();
  {
    std::lock_guard<std::mutex> lk(mu_);
    for (auto& r : queue_) {
      r->error = "server stopped";
      r->done = true;
    }
    queue_.clear();
  }
  done_cv_.notify_all();
  (void)hipSetDevice(device_);
  for (auto& s : slots_) {
    s.plans.clear();
    if (s.ctx && s.ctx->stream) {
      (void)hipStreamSynchronize(s.ctx->stream);
      (void)hipStreamDestroy(s.ctx->stream);
    }
    s.ctx.reset();
  }
}

std::shared_ptr<Server::PlanEntry> Server::plan_entry(const uint32_t* words, size_t n) {
  std::string key(reinterpret_cast<const char*>(words), n * 4);
  auto it = plans_.find(key);
  if (it != plans_.end()) return it->second;
  auto e = std::make_shared<PlanEntry>();
  e->words.assign(words, words + n);
  e->data = parse_plan(words, n);
  e->bytes_per_instance = plan_bytes_per_instance(e->data);
  plans_[key] = e;
  return e;
}

uint64_t Server::submit(const uint32_t* words, size_t n_words, int batch, const uint32_t* const* columns,
                        const uint32_t* const* perms) {
  if (!words || batch <= 0) throw std::runtime_error("server submit: bad arguments");
  auto r = std::make_shared<Request>();
  {
    std::lock_guard<std::mutex> lk(mu_);
    r->plan = plan_entry(words, n_words);
  }
  const PlanData& d = r->plan->data;
  const size_t nc = d.inputs.size(), np = d.perms.size();
  if ((nc && !columns) || (np && !perms)) throw std::runtime_error("server submit: missing inputs");
  size_t total = 0;
  for (size_t k = 0; k < nc; ++k) total += d.modules[d.cols[d.inputs[k]].module].N;
  for (auto& p : d.perms) total += p.len;
  r->batch = batch;
  r->data.resize(total * (size_t)batch);
  r->inputs.resize(batch);
  size_t off = 0;
  for (int b = 0; b < batch; ++b) {
    for (size_t k = 0; k < nc; ++k) {
      const size_t n = d.modules[d.cols[d.inputs[k]].module].N;
      const uint32_t* src = columns[(size_t)b * nc + k];
      if (!src) throw std::runtime_error("server submit: null column pointer");
      std::memcpy(r->data.data() + off, src, n * 4);
      r->inputs[b].columns.push_back(r->data.data() + off);
      off += n;
    }
    for (size_t k = 0; k < np; ++k) {
      const uint32_t* src = perms[(size_t)b * np + k];
      if (!src) throw std::runtime_error("server submit: null permutation pointer");
      std::memcpy(r->data.data() + off, src, (size_t)d.perms[k].len * 4);
      r->inputs[b].perms.push_back(r->data.data() + off);
      off += d.perms[k].len;
    }
  }
  r->bytes = r->plan->bytes_per_instance * (size_t)batch;
  {
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw std::runtime_error("server stopped");
    r->ticket = next_ticket_++;
    requests_[r->ticket] = r;
    queue_.push_back(r);
    ++st_.submitted;
  }
  work_cv_.notify_one();
  return r->ticket;
}

void Server::worker(int si) {
  Slot& sl = slots_[si];
  (void)hipSetDevice(device_);
  for (;;) {
    std::vector<std::shared_ptr<Request>> take;
    size_t bytes = 0;
    {
      std::unique_lock<std::mutex> lk(mu_);
      // Admission: the oldest request goes first (no overtaking, so a large
      // request is not starved); it is admitted when its footprint fits the
      // budget beside the running ones, or when nothing else runs.
      work_cv_.wait(lk, [&] {
        return stop_ || (!queue_.empty() && (st_.running == 0 || in_flight_ + queue_.front()->bytes <= budget_));
      });
      if (stop_) return;
      auto head = queue_.front();
      queue_.pop_front();
      take.push_back(head);
      bytes = head->bytes;
      int inst = head->batch;
      // continuous batching: queued requests of the same plan join this prove
      for (auto it = queue_.begin(); it != queue_.end();) {
        const auto& r = *it;
        if (r->plan == head->plan && inst + r->batch <= max_batch_ && in_flight_ + bytes + r->bytes <= budget_) {
          take.push_back(r);
          bytes += r->bytes;
          inst += r->batch;
          it = queue_.erase(it);
        } else {
          ++it;
        }
      }
      in_flight_ += bytes;
      ++st_.running;
      ++st_.proves;
      st_.merged += take.size() - 1;
      st_.peak_running = std::max<uint64_t>(st_.peak_running, st_.running);
      st_.peak_bytes = std::max<uint64_t>(st_.peak_bytes, in_flight_);
    }
    std::vector<InstanceInput> in;
    for (auto& r : take) in.insert(in.end(), r->inputs.begin(), r->inputs.end());
    std::vector<ProveOutput> out;
    std::string err;
    try {
      auto& pl = sl.plans[take[0]->plan.get()];
      if (!pl) pl = std::make_unique<Plan>(sl.ctx.get(), take[0]->plan->words.data(), take[0]->plan->words.size());
      pl->prove(in, out, nullptr);
      // keep this slot's cached device buffers within its share of the budget:
      // graph buckets (buffers kept per recorded shape) get half of it
      const size_t share = budget_ / slots_.size();
      size_t held = 0;
      for (auto& kv : sl.plans) held += kv.second->graph_bytes();
      if (held > share / 2)
        for (auto& kv : sl.plans) kv.second->drop_graphs();
      sl.ctx->pool.trim(share);
    } catch (const std::exception& e) {
      err = e.what();
    }
    {
      std::lock_guard<std::mutex> lk(mu_);
      size_t off = 0;
      // Errors are per request: an instance that failed (invalid witness)
      // fails its own request only; its merge-mates get their proofs.  A
      // failure of the whole prove (err) fails every request in it.
      for (auto& r : take) {
        std::string rerr = err;
        if (rerr.empty())
          for (int b = 0; b < r->batch; ++b)
            if (!out[off + b].error.empty()) {
              rerr = "instance " + std::to_string(b) + ": " + out[off + b].error;
              break;
            }
        if (rerr.empty()) {
          r->out.assign(std::make_move_iterator(out.begin() + off), std::make_move_iterator(out.begin() + off + r->batch));
          ++st_.completed;
        } else {
          r->error = rerr;
          ++st_.failed;
        }
        off += r->batch;
        r->done = true;
      }
      in_flight_ -= bytes;
      --st_.running;
    }
    done_cv_.notify_all();
    work_cv_.notify_all();
  }
}

bool Server::wait(uint64_t ticket, int timeout_ms, std::vector<ProveOutput>* out, std::string* err, bool keep) {
  std::unique_lock<std::mutex> lk(mu_);
  auto it = requests_.find(ticket);
  if (it == requests_.end()) throw std::runtime_error("server wait: unknown ticket");
  auto r = it->second;
  auto pred = [&] { return r->done; };
  if (timeout_ms < 0) {
    done_cv_.wait(lk, pred);
  } else if (!done_cv_.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred)) {
    return false;
  }
  if (err) *err = r->error;
  if (out) *out = r->out;
  if (!keep) requests_.erase(ticket);
  return true;
}

Server::Stats Server::stats() {
  std::lock_guard<std::mutex> lk(mu_);
  Stats s = st_;
  s.queued = queue_.size();
  return s;
}

}  // namespace lg
