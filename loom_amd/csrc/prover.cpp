// Host orchestration of the prove phase (prover/prover.go:655-719 and the PCS
// it drives), for B instances of one plan in lockstep: every phase is one
// launch sequence over all instances, the Fiat-Shamir transcripts (one per
// instance) run on the host between phases, and only roots, claimed values,
// the final polynomial and the query openings come back from the device.
#include "prover.h"

#include <thread>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>

#include "kernels.h"
#include "ntt16.h"
#include "pool.h"
#include "protocol.h"
#include "transcript.h"

namespace lg {

namespace {

std::vector<uint32_t> powers(uint32_t g, size_t n, uint32_t scale = KB_R) {
  std::vector<uint32_t> v(n);
  uint32_t a = scale;
  for (size_t i = 0; i < n; ++i) {
    v[i] = a;
    a = kb_mul(a, g);
  }
  return v;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ------------------------------------------------------------------ tables
const uint32_t* Tables::get(const std::string& key, const std::function<std::vector<uint32_t>()>& make) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = bufs_.find(key);
  if (it != bufs_.end()) return it->second.as<uint32_t>();
  std::vector<uint32_t> v = make();
  // copied on a stream of its own, not the legacy stream: another thread may
  // be recording a graph (a legacy-stream copy would invalidate it)
  if (!stream_) LG_HIP(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  DevBuf b(v.size() * 4);
  LG_HIP(hipMemcpyAsync(b.as<void>(), v.data(), v.size() * 4, hipMemcpyHostToDevice, stream_));
  LG_HIP(hipStreamSynchronize(stream_));
  const uint32_t* p = b.as<uint32_t>();
  bufs_.emplace(key, std::move(b));
  return p;
}

Tables::~Tables() {
  if (stream_) (void)hipStreamDestroy(stream_);
}

const uint32_t* Tables::tw_fwd(int logn) {
  return get("twf" + std::to_string(logn), [&] { return powers(omega(logn), std::max<size_t>(1, (1u << logn) / 2)); });
}

const uint32_t* Tables::tw_inv(int logn) {
  return get("twi" + std::to_string(logn),
             [&] { return powers(kb_inv(omega(logn)), std::max<size_t>(1, (1u << logn) / 2)); });
}

const uint32_t* Tables::ntt_ladders(bool inverse) {
  return get(inverse ? "ladi" : "ladf", [&] {
    auto root = [&](int logn) { return inverse ? kb_inv(omega(logn)) : omega(logn); };
    NttRoots R;
    for (int k = 0; k < 8; ++k) R.r16[k] = kb_pow(root(4), k);
    for (int k = 0; k < 4; ++k) R.r8[k] = kb_pow(root(3), k);
    R.r4 = root(2);
    std::vector<uint32_t> t((size_t)16 * ((1u << (NTT_MAX_LOG_LDS - 3)) - 1), 0);
    for (int logm = 4; logm <= NTT_MAX_LOG_LDS; ++logm) {
      const uint32_t q = 1u << (logm - 4), w = root(logm);
      uint32_t W = KB_R;
      for (uint32_t b = 0; b < q; ++b, W = kb_mul(W, w)) {
        const Ladder L = ntt16_ladder(W, R);
        uint32_t* e = &t[(size_t)16 * (q - 1 + b)];
        for (int k = 0; k < 8; ++k) e[k] = L.a[k];
        for (int k = 0; k < 4; ++k) e[8 + k] = L.b[k];
        e[12] = L.c[0];
        e[13] = L.c[1];
        e[14] = L.d;
      }
    }
    return t;
  });
}

const uint32_t* Tables::wpow(int logn) {
  return get("wp" + std::to_string(logn), [&] { return powers(omega(logn), 1u << logn); });
}

const uint32_t* Tables::wpow_inv(int logn) {
  return get("wpi" + std::to_string(logn), [&] { return powers(kb_inv(omega(logn)), 1u << logn); });
}

const uint32_t* Tables::geo(const std::string& key, uint32_t s, size_t n, uint32_t sc) {
  const uint32_t* t = get(key, [&] { return powers(s, n, sc); });
  std::lock_guard<std::mutex> lock(geo_mu_);
  geo_[t] = s;
  return t;
}

bool Tables::geometric(const uint32_t* t, uint32_t* s) {
  std::lock_guard<std::mutex> lock(geo_mu_);
  auto it = geo_.find(t);
  if (it == geo_.end()) return false;
  *s = it->second;
  return true;
}

const uint32_t* Tables::geo_blocks(uint32_t s, int a, int b) {
  return get("gblk" + std::to_string(s) + "_" + std::to_string(a) + "_" + std::to_string(b), [&] {
    const uint32_t sb = kb_pow(s, 1ull << b);  // s^(2^b)
    const std::vector<uint32_t> pw = powers(sb, 1u << a);
    std::vector<uint32_t> v(1u << a);
    for (uint32_t p = 0; p < (1u << a); ++p) v[p] = pw[brev_host(p, a)];
    return v;
  });
}

const uint32_t* Tables::lde_pre(int logn, int b, int rate_log, bool from_values) {
  const uint32_t s = kb_pow(omega(logn + rate_log), brev_host(b, rate_log));
  const uint32_t sc = from_values ? kb_inv(M(1u << logn)) : KB_R;
  return geo("lde" + std::to_string(logn) + "_" + std::to_string(b) + "_" + std::to_string(from_values), s,
             (size_t)1 << logn, sc);
}

const uint32_t* Tables::quot_pre(int logn, int i, int rho_log) {
  const uint32_t s = kb_mul(M(MULT_GEN), kb_pow(omega(logn + rho_log), i));
  return geo("qpre" + std::to_string(logn) + "_" + std::to_string(i) + "_" + std::to_string(rho_log), s,
             (size_t)1 << logn, kb_inv(M(1u << logn)));
}

const uint32_t* Tables::quot_post(int logn, int i, int rho_log) {
  return get("qpost" + std::to_string(logn) + "_" + std::to_string(i) + "_" + std::to_string(rho_log), [&] {
    uint32_t s = kb_inv(kb_pow(omega(logn + rho_log), i));
    return powers(s, 1u << logn, kb_inv(M(1u << (logn + rho_log))));
  });
}

const uint32_t* Tables::ginv_pow(int log_len) {
  return get("ginv" + std::to_string(log_len), [&] { return powers(kb_inv(M(MULT_GEN)), 1u << log_len); });
}

const uint32_t* Tables::fri_xinv2(int log_m) {
  return get("fx" + std::to_string(log_m), [&] {
    const uint32_t half = 1u << (log_m - 1);
    const int hb = log_m - 1;
    std::vector<uint32_t> pw = powers(kb_inv(omega(log_m)), half);
    std::vector<uint32_t> v(half);
    const uint32_t inv2 = kb_inv(M(2));
    for (uint32_t i = 0; i < half; ++i) v[i] = kb_mul(pw[brev_host(i, hb)], inv2);
    return v;
  });
}

const uint32_t* Tables::gpow_small(int count) {
  return get("gs" + std::to_string(count), [&] { return powers(M(MULT_GEN), count); });
}

// ------------------------------------------------------------------ kernel timer
hipEvent_t KernelTimer::ev() {
  if (!free_.empty()) {
    hipEvent_t e = free_.back();
    free_.pop_back();
    return e;
  }
  hipEvent_t e;
  LG_HIP(hipEventCreate(&e));
  return e;
}

void KernelTimer::begin(const char* name, hipStream_t s, double bytes, double perms) {
  if (!on) return;
  (void)s;
  Rec r{name, ev(), ev(), bytes, perms};
  pending_.push_back(r);
  start_pending_ = true;
}

void KernelTimer::launch_point(hipStream_t s) {
  if (!on || !start_pending_) return;
  LG_HIP(hipEventRecord(pending_.back().a, s));
  start_pending_ = false;
}

void KernelTimer::end(hipStream_t s) {
  if (!on || pending_.empty()) return;
  launch_point(s);  // a record without a launch times nothing
  LG_HIP(hipEventRecord(pending_.back().b, s));
}

void KernelTimer::resolve() {
  for (auto& r : pending_) {
    float ms = 0;
    LG_HIP(hipEventSynchronize(r.b));
    LG_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    Stat& st = stats[r.name];
    st.count += 1;
    st.ms += ms;
    st.bytes += r.bytes;
    st.perms += r.perms;
    free_.push_back(r.a);
    free_.push_back(r.b);
  }
  pending_.clear();
}

KernelTimer::~KernelTimer() {
  for (auto& r : pending_) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : free_) (void)hipEventDestroy(e);
}

// ------------------------------------------------------------------ pool / staging
DevBuf Pool::get(size_t bytes) {
  std::lock_guard<std::mutex> lock(mu_);
  auto it = free_.find(bytes);
  if (it != free_.end()) {
    DevBuf b = std::move(it->second);
    free_.erase(it);
    cached_ -= bytes;
    return b;
  }
  return DevBuf(bytes);
}

void Pool::trim(size_t keep_bytes) {
  std::lock_guard<std::mutex> lock(mu_);
  while (cached_ > keep_bytes && !free_.empty()) {
    auto it = std::prev(free_.end());
    cached_ -= it->first;
    free_.erase(it);
  }
}

void Pool::put(DevBuf&& b) {
  if (!b.bytes() || !b.owned()) return;  // views of arena memory go back to their arena
  std::lock_guard<std::mutex> lock(mu_);
  size_t n = b.bytes();
  cached_ += n;
  free_.emplace(n, std::move(b));
}

Pool::~Pool() = default;

Staging::Staging(size_t bytes) : cap_(bytes) {
  LG_HIP(hipHostMalloc(reinterpret_cast<void**>(&host_), bytes, hipHostMallocDefault));
  dev_.alloc(bytes);
}

Staging::~Staging() {
  if (host_) (void)hipHostFree(host_);
}

void* Staging::put(const void* src, size_t n, hipStream_t) {
  size_t off = (off_ + 255) & ~size_t(255);
  if (off + n > cap_) throw std::runtime_error("staging area exhausted");
  if (n && verify_ && std::memcmp(host_ + off, src, n) != 0) changed_ = true;
  if (n) std::memcpy(host_ + off, src, n);
  off_ = off + n;
  return dev_.as<uint8_t>() + off;
}

void Staging::assume_flushed() { flushed_ = off_; }

void Staging::flush_sync(hipStream_t side) {
  if (off_ > flushed_) {
    LG_HIP(hipMemcpyAsync(dev_.as<uint8_t>() + flushed_, host_ + flushed_, off_ - flushed_, hipMemcpyHostToDevice,
                          side));
    LG_HIP(hipStreamSynchronize(side));
    flushed_ = off_;
  }
}

void Staging::flush(hipStream_t s) {
  if (off_ > flushed_) {
    LG_HIP(hipMemcpyAsync(dev_.as<uint8_t>() + flushed_, host_ + flushed_, off_ - flushed_, hipMemcpyHostToDevice, s));
    flushed_ = off_;
  }
}

Context::Context() {
  if (const char* e = getenv("LOOMGPU_LANES")) {
    lanes = std::max(1, std::min(8, atoi(e)));
    lanes_forced = true;
  }
  if (const char* e = getenv("LOOMGPU_GRAPHS")) graphs = atoi(e) != 0;
  lane_.resize(8);
}

hipStream_t Context::side_stream() {
  std::lock_guard<std::mutex> lock(lane_mu_);
  if (!side_) LG_HIP(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
  return side_;
}

hipStream_t Context::lane_stream(int lane) {
  if (lane == 0) return stream;
  std::lock_guard<std::mutex> lock(lane_mu_);
  LaneBufs& L = lane_.at(lane);
  if (!L.stream) LG_HIP(hipStreamCreateWithFlags(&L.stream, hipStreamNonBlocking));
  return L.stream;
}

static uint32_t* grow_pinned(uint32_t*& p, size_t& cap, size_t bytes) {
  if (bytes > cap) {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t c = std::max<size_t>(bytes, 1u << 20);
    LG_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), c, hipHostMallocDefault));
    cap = c;
  }
  return p;
}

uint32_t* Context::pinned(size_t bytes, int lane) {
  LaneBufs& L = lane_.at(lane);
  return grow_pinned(L.pinned, L.pinned_cap, bytes);
}

uint32_t* Context::pinned_in(size_t bytes, int lane) {
  LaneBufs& L = lane_.at(lane);
  return grow_pinned(L.pinned_in, L.pinned_in_cap, bytes);
}

uint32_t* Context::pinned_aux(size_t bytes, int lane) {
  LaneBufs& L = lane_.at(lane);
  return grow_pinned(L.pinned_aux, L.pinned_aux_cap, bytes);
}

uint32_t* Context::pinned_fs(size_t bytes, int lane) {
  LaneBufs& L = lane_.at(lane);
  return grow_pinned(L.pinned_fs, L.pinned_fs_cap, bytes);
}

uint32_t* Context::pinned_ext(size_t bytes, int lane, size_t idx) {
  LaneBufs& L = lane_.at(lane);
  if (L.ext.size() <= idx) L.ext.resize(idx + 1, {nullptr, 0});
  return grow_pinned(L.ext[idx].first, L.ext[idx].second, bytes);
}

Context::~Context() {
  if (side_) {
    (void)hipStreamSynchronize(side_);
    (void)hipStreamDestroy(side_);
  }
  for (auto& L : lane_) {
    if (L.pinned) (void)hipHostFree(L.pinned);
    if (L.pinned_in) (void)hipHostFree(L.pinned_in);
    if (L.pinned_aux) (void)hipHostFree(L.pinned_aux);
    if (L.pinned_fs) (void)hipHostFree(L.pinned_fs);
    for (auto& e : L.ext)
      if (e.first) (void)hipHostFree(e.first);
    if (L.stream) {
      (void)hipStreamSynchronize(L.stream);
      (void)hipStreamDestroy(L.stream);
    }
  }
}

std::unique_ptr<Staging> Context::take_staging() {
  std::lock_guard<std::mutex> lock(staging_mu);
  if (!staging_free.empty()) {
    auto s = std::move(staging_free.back());
    staging_free.pop_back();
    s->reset();
    return s;
  }
  return std::make_unique<Staging>(64u << 20);
}

void Context::give_staging(std::unique_ptr<Staging> s) {
  std::lock_guard<std::mutex> lock(staging_mu);
  staging_free.push_back(std::move(s));
}

// ------------------------------------------------------------------ plan
// The modules whose claimed values form a DEEP class (fri.go:427-437): those
// with at least one claimed value in some tree group, in module order.  Each
// class is one FRI level.
static std::vector<size_t> deep_classes(const PlanData& d) {
  std::vector<char> has(d.modules.size(), 0);
  for (auto& t : d.trees)
    for (auto& g : t.groups) {
      bool any = false;
      for (auto& sh : g.base_shifts) any = any || !sh.empty();
      for (auto& sh : g.ext_shifts) any = any || !sh.empty();
      if (any) has.at(g.module) = 1;
    }
  std::vector<size_t> r;
  for (size_t m = 0; m < has.size(); ++m)
    if (has[m]) r.push_back(m);
  return r;
}

// Registration order of the challenges: the round challenges and __zeta
// (prover/prover.go:208-212), alpha_DEEP (internal/fri/open.go:151), the
// level gammas, the fold rounds and the queries (internal/fri/fri.go:425-436).
static std::shared_ptr<const ChallengeSchedule> challenge_schedule(const PlanData& d) {
  auto s = std::make_shared<ChallengeSchedule>();
  for (size_t r = 0; r < d.rounds.size(); ++r) s->add(challenge_name(r));
  s->add("__zeta");
  s->add("alpha_DEEP");
  const size_t L = deep_classes(d).size();
  for (size_t l = 1; l < L; ++l) s->add("fri_level_" + std::to_string(l) + "_gamma");
  for (int j = 0; j < FRI_ROUNDS; ++j) s->add("fri_fold_" + std::to_string(j));
  for (uint32_t k = 0; k < d.num_queries; ++k) s->add("fri_query_" + std::to_string(k));
  return s;
}

// FNV-1a over the plan words: the key of a plan in the tuning table.
static uint64_t plan_hash(const uint32_t* w, size_t n) {
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < n; ++i)
    for (int b = 0; b < 4; ++b) {
      h ^= (w[i] >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  return h;
}

static std::string device_arch(int dev) {
  hipDeviceProp_t p;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "unknown";
  std::string a = p.gcnArchName;
  const size_t c = a.find(':');
  return c == std::string::npos ? a : a.substr(0, c);
}

Tuning Plan::tuning(int batch) const {
  auto it = tuned_.find(batch);
  return it == tuned_.end() ? Tuning() : it->second;
}

void Plan::set_grinding(uint32_t bits) {
  if (bits > 31) throw std::runtime_error("grinding: at most 31 bits (one KoalaBear salt)");
  grinding_ = bits;
  drop_graphs();  // recorded graphs hold the fold-challenge launches
}

void Plan::set_tuning(int batch, const Tuning& t) {
  tuned_[batch] = t;
  drop_graphs();  // recorded graphs hold the previous launch choices
}

Tuning Plan::autotune(const std::vector<InstanceInput>& inputs, int reps, double* ms, int* trials) {
  if (inputs.empty()) throw std::runtime_error("autotune: no instances");
  const int B = (int)inputs.size();
  reps = std::max(1, std::min(reps, 9));
  std::vector<InstanceInput> in = inputs;
  for (auto& x : in) {  // the trial proofs are discarded
    x.proof_dst = nullptr;
    x.proof_cap = 0;
  }
  // Trials run as production proves do: with graphs on, a trial's first
  // prove runs eagerly and its second records the graph, the timed ones
  // replay it; the proofs go to caller-style buffers (allocated once, after
  // the first prove gives their length), as lg_prove writes them.
  drop_graphs();
  int n_trials = 0;
  std::vector<std::vector<uint32_t>> sink(B);
  auto time_it = [&](const Tuning& t) {
    tuned_[B] = t;
    drop_graphs();
    std::vector<ProveOutput> out;
    prove(in, out);  // warm-up: tables, buffers
    for (int b = 0; b < B; ++b) {
      if (in[b].proof_dst && in[b].proof_cap >= out[b].proof_len) continue;
      sink[b].assign(out[b].proof_len + 1024, 0u);
      in[b].proof_dst = sink[b].data();
      in[b].proof_cap = sink[b].size();
    }
    if (ctx_->graphs) prove(in, out);  // records the graph
    std::vector<double> v;
    for (int r = 0; r < reps; ++r) {
      const double a = now_s();
      prove(in, out);
      v.push_back((now_s() - a) * 1e3);
    }
    std::sort(v.begin(), v.end());
    ++n_trials;
    return v[v.size() / 2];
  };
  Tuning best = tuning(B);
  double best_ms = 0;
  try {
    best_ms = time_it(best);
    if (ms) ms[0] = best_ms;
    auto consider = [&](Tuning t) {
      const double m = time_it(t);
      if (m < 0.99 * best_ms) {
        best = t;
        best_ms = m;
      }
    };
    for (int v : {1, 2, 3})
      if (v != best.lanes && (v == 1 || B >= 8 * v)) {
        Tuning t = best;
        t.lanes = v;
        consider(t);
      }
    for (int v : {1 << 13, 1 << 14, 1 << 15, 1 << 16, 1 << 17})
      if (v != best.merkle_wide_min) {
        Tuning t = best;
        t.merkle_wide_min = v;
        consider(t);
      }
    for (int v : {64, 128, 256})
      if (v != best.hash_threads) {
        Tuning t = best;
        t.hash_threads = v;
        consider(t);
      }
    for (int v : {128, 512, 2048})
      if (v != best.fri_coop_pairs) {
        Tuning t = best;
        t.fri_coop_pairs = v;
        consider(t);
      }
    for (int v : {512, 1024, 4096})
      if (v != best.claims_rows) {
        Tuning t = best;
        t.claims_rows = v;
        consider(t);
      }
  } catch (...) {
    tuned_[B] = best;
    drop_graphs();
    throw;
  }
  tuned_[B] = best;
  drop_graphs();
  if (ms) ms[1] = best_ms;
  if (trials) *trials = n_trials;
  if (const char* f = getenv("LOOMGPU_TUNE_FILE")) {
    if (FILE* fp = fopen(f, "a")) {
      fprintf(fp, "LGTUNE1 %016llx %d %s %d %d %d %d %d %.4f\n", (unsigned long long)hash_, B,
              device_arch(ctx_->device).c_str(), best.lanes, best.merkle_wide_min, best.hash_threads,
              best.fri_coop_pairs, best.claims_rows, best_ms);
      fclose(fp);
    }
  }
  return best;
}

Plan::Plan(Context* ctx, const uint32_t* words, size_t n) : ctx_(ctx), d_(parse_plan(words, n)) {
  hash_ = plan_hash(words, n);
  if (const char* f = getenv("LOOMGPU_TUNE_FILE")) {  // earlier autotune results for this plan and device
    if (FILE* fp = fopen(f, "r")) {
      const std::string arch = device_arch(ctx->device);
      char tag[16], hs[32], ar[64];
      int batch;
      Tuning t;
      double m;
      while (fscanf(fp, "%15s %31s %d %63s %d %d %d %d %d %lf", tag, hs, &batch, ar, &t.lanes, &t.merkle_wide_min,
                    &t.hash_threads, &t.fri_coop_pairs, &t.claims_rows, &m) == 10) {
        if (std::string(tag) == "LGTUNE1" && strtoull(hs, nullptr, 16) == hash_ && arch == ar && t.lanes >= 1 &&
            t.lanes <= 8 && t.hash_threads >= 64 && t.hash_threads <= 256 && t.hash_threads % 64 == 0)
          tuned_[batch] = t;  // the last line of a key wins
      }
      fclose(fp);
    }
  }
  for (auto& m : d_.modules) {
    if (m.log_n > FRI_LOG_D) throw std::runtime_error("module size above 2^22 (MAX_LOG_MODULE_SIZE)");
    if (m.log_n < 1) throw std::runtime_error("module size below 2");
    if (m.n_chunks > 8) throw std::runtime_error("constraint degree above 9 is not supported");
  }
  for (auto& pm : d_.perms)
    if (pm.len / d_.modules[pm.module].N > 64) throw std::runtime_error("permutation with more than 64 chunks");
  sched_ = challenge_schedule(d_);
  LG_HIP(hipSetDevice(ctx->device));
  rel_.resize(d_.modules.size());
  for (size_t m = 0; m < d_.modules.size(); ++m)
    if (d_.modules[m].n_chunks) rel_[m] = dag_kernel_get(d_.modules[m].relation, d_.cols);
  std::vector<char> logup_total(d_.cols.size(), 0);
  for (size_t r = 0; r < d_.rounds.size(); ++r)
    for (size_t s = 0; s < d_.rounds[r].steps.size(); ++s) {
      const Step& st = d_.rounds[r].steps[s];
      size_t n_kernels = st.ins.size();
      switch (st.kind) {
        case STEP_GRAND_PRODUCT:
          if (st.ins.size() != 2 || st.outs.size() != 1) throw std::runtime_error("grand product step: bad shape");
          break;
        case STEP_LOOKUP_MULTIPLICITY: {
          // params: n_sources, n_targets, width, has_sel_s, has_sel_t, module
          // per source, module per target; ins: selS, selT, source tuples,
          // target tuples (board/step.go:257-295)
          if (st.params.size() < 5) throw std::runtime_error("lookup step: bad parameters");
          const uint32_t ns = st.params[0], nt = st.params[1], w = st.params[2];
          const bool hs = st.params[3] != 0, ht = st.params[4] != 0;
          if (ns < 1 || nt < 1 || w < 1 || st.params.size() != 5 + ns + nt)
            throw std::runtime_error("lookup step: bad shape");
          if (st.ins.size() != (hs ? ns : 0) + (ht ? nt : 0) + (size_t)(ns + nt) * w || st.outs.size() != nt)
            throw std::runtime_error("lookup step: bad inputs or outputs");
          for (uint32_t j = 0; j < ns + nt; ++j)
            if (st.params[5 + j] >= d_.modules.size()) throw std::runtime_error("lookup step: module out of range");
          for (uint32_t k = 0; k < nt; ++k)
            if (d_.cols[st.outs[k]].field != 0 || d_.cols[st.outs[k]].module != st.params[5 + ns + k])
              throw std::runtime_error("lookup step: bad multiplicity column");
          // tuples and selectors are base-field values (evalTable, poly/lookup_multiplicity.go:112-139)
          for (auto& d : st.ins)
            if (d.nodes.empty() || d.root >= d.nodes.size() || d.nodes[d.root].field != 0)
              throw std::runtime_error("lookup step: tuple entries and selectors must be base-field expressions");
          break;
        }
        case STEP_CYCLIC_LOGUP:
          if (st.params.size() != 1 || st.params[0] == 0 || st.ins.size() != 2 * st.params[0] || st.outs.size() != 2)
            throw std::runtime_error("cyclic logup step: bad shape");
          if (d_.cols[st.outs[0]].field != 1 || d_.cols[st.outs[1]].field != 1)
            throw std::runtime_error("cyclic logup step: outputs must be extension columns");
          logup_total[st.outs[1]] = 1;
          break;
        default:
          throw std::runtime_error("unsupported prover step");
      }
      for (size_t k = 0; k < n_kernels; ++k) steps_[{r, s, k}] = dag_kernel_get(st.ins[k], d_.cols);
    }
  for (auto& x : d_.exposed)
    if (!logup_total[x.col]) throw std::runtime_error("plan: exposed value is not a logup total");
  graph_eligible_ = deep_classes(d_).size() == 1;
}

GraphBucket::~GraphBucket() {
  if (exec) (void)hipGraphExecDestroy(exec);
  for (auto& p : pins) (void)hipHostFree(p.first);
}

// The bucket a lane's Job of `batch` instances uses, and its mode (Job::Mode):
// the first prove of a shape runs eagerly (and creates every lazily built
// table), the second records the graph, later ones replay it.  Without a
// bucket (ineligible plan, graphs off, per-kernel timing or diagnostics on,
// a failed recording) the prove is eager.
GraphBucket* Plan::take_bucket(int batch, int lane, bool device_inputs, int* mode) {
  *mode = 0;
  static const bool diag = getenv("LOOMGPU_TRACE") || getenv("LOOMGPU_DEBUG_DIR");
  if (!graph_eligible_ || !ctx_->graphs || ctx_->timer.on || diag || (options_ & PROVE_SKIP_FRI)) return nullptr;
  std::lock_guard<std::mutex> lk(gmu_);
  auto key = std::make_tuple(batch, lane, device_inputs);
  auto it = buckets_.find(key);
  if (it == buckets_.end()) {
    while ((int)buckets_.size() >= MAX_GRAPH_BUCKETS) {
      auto lru = buckets_.end();
      for (auto i = buckets_.begin(); i != buckets_.end(); ++i)
        if (!i->second->busy && (lru == buckets_.end() || i->second->last < lru->second->last)) lru = i;
      if (lru == buckets_.end()) return nullptr;
      buckets_.erase(lru);
    }
    auto b = std::make_unique<GraphBucket>();
    b->batch = batch;
    b->lane = lane;
    it = buckets_.emplace(key, std::move(b)).first;
  }
  GraphBucket* b = it->second.get();
  b->last = ++tick_;
  if (b->broken || b->busy) return nullptr;
  if (++b->proves == 1) return nullptr;
  b->busy = true;
  *mode = b->exec ? 2 : 1;
  return b;
}

void Plan::give_bucket(GraphBucket* b) {
  if (!b) return;
  std::lock_guard<std::mutex> lk(gmu_);
  b->busy = false;
  if (b->broken) {  // free what a failed recording or replay held
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    b->bufs.clear();
    b->bytes = 0;
  }
}

size_t Plan::graph_bytes() {
  std::lock_guard<std::mutex> lk(gmu_);
  size_t n = 0;
  for (auto& kv : buckets_) n += kv.second->bytes;
  return n;
}

void Plan::drop_graphs() {
  std::lock_guard<std::mutex> lk(gmu_);
  for (auto it = buckets_.begin(); it != buckets_.end();)
    if (it->second->busy) ++it;
    else it = buckets_.erase(it);
}

// ------------------------------------------------------------------ job
namespace {

// Brackets one launch with timer events when the context's timer is on.
struct TimeIt {
  Context& c;
  hipStream_t s;
  TimeIt(Context& ctx, hipStream_t st, const char* name, double bytes, double perms) : c(ctx), s(st) {
    c.timer.begin(name, s, bytes, perms);
  }
  ~TimeIt() { c.timer.end(s); }
};

// push_back / insert(end(), first, last) into a caller buffer of cap words;
// n counts every word, also those past the capacity (which are dropped).
struct SpanWriter {
  uint32_t* p;
  size_t cap;
  size_t n = 0;
  struct End {};
  End end() const { return {}; }
  void push_back(uint32_t x) {
    if (n < cap) p[n] = x;
    ++n;
  }
  void insert(End, const uint32_t* a, const uint32_t* b) {
    const size_t k = (size_t)(b - a);
    if (n + k <= cap) std::memcpy(p + n, a, k * 4);
    else
      for (size_t i = 0; i < k && n + i < cap; ++i) p[n + i] = a[i];
    n += k;
  }
};

struct ColBuf {
  DevBuf buf;
  bool committed = false;
  bool lde_ready = false;
  int limbs = 1;
  uint32_t n = 0;
  size_t blk = 0;  // words per instance
  uint32_t* lag(int b, int l) const { return buf.as<uint32_t>() + b * blk + (size_t)l * n; }
  uint32_t* lde(int b, int l) const { return buf.as<uint32_t>() + b * blk + (size_t)limbs * n + (size_t)l * 4 * n; }
};

struct TreeBuf {
  int nleaves = 0;
  DevBuf nodes;                       // B x 2 nleaves x 8
  std::vector<size_t> order;          // group indexes, decreasing N (stable)
  std::vector<DevBuf> inj;            // per order position >= 1: B x width x 8
  std::vector<int> inj_width;
  uint32_t* const* nodes_d = nullptr;  // staged node-array pointers of the batch (set by build_tree)
  bool shared = false;                 // one tree for every instance (the setup tree)
  uint32_t* node_ptr(int b) const { return nodes.as<uint32_t>() + (shared ? 0 : (size_t)b * 2 * nleaves * 8); }
  uint32_t* inj_ptr(size_t k, int b) const {
    return inj[k].as<uint32_t>() + (shared ? 0 : (size_t)b * inj_width[k] * 8);
  }
};

}  // namespace

// The resident proving key of a plan: every setup column (Lagrange values and
// rate-4 LDE, 5N words, in setup order) and the setup tree.
struct SetupData {
  std::vector<DevBuf> cols;
  TreeBuf tree;
  uint32_t root[8] = {0};  // Montgomery
};

namespace {

struct Layer {  // an E6 vector of len elements per instance, 6 limb arrays
  DevBuf buf;
  uint32_t len = 0;
  uint32_t* ptr(int b) const { return buf.as<uint32_t>() + (size_t)b * 6 * len; }
};

struct FsMsgBuilder;

class Job {
 public:
  // With a graph bucket the Job's device buffers, staging area and pinned
  // buffers are the bucket's (the same addresses on every prove of that
  // shape), and its device phase is recorded into the bucket's HIP graph
  // (mode CAPTURE) or replayed from it (mode REPLAY): in REPLAY the host
  // logic runs unchanged but launches nothing until region_end launches the
  // graph.
  enum Mode { EAGER, CAPTURE, REPLAY };
  Job(Plan& plan, const std::vector<InstanceInput>& in, std::vector<ProveOutput>& out, const PhaseCallback& cb,
      int lane, GraphBucket* gb = nullptr, Mode mode = EAGER)
      : plan_(plan), d_(plan.data()), ctx_(*plan.ctx()), tb_(ctx_.tables), in_(in), out_(out), cb_(cb),
        B_((int)in.size()), lane_(lane), s_(ctx_.lane_stream(lane)), gb_(gb), mode_(gb ? mode : EAGER),
        tune_(plan.tuning(plan.tuning_batch())), skip_fri_((plan.options() & PROVE_SKIP_FRI) != 0) {
    if (out.size() != in.size()) throw std::runtime_error("prove: output slots do not match the inputs");
    if (gb_) {
      if (!gb_->stg) gb_->stg = std::make_unique<Staging>(64u << 20);
      gb_->stg->reset();
    } else {
      stg_ = ctx_.take_staging();
    }
    set_prelaunch_hook([](void* u) { static_cast<Job*>(u)->flush_staging(); }, this);
  }
  Staging& stg() { return gb_ ? *gb_->stg : *stg_; }
  // While a graph is recorded, staged tables go to the device at once (a
  // synchronous copy outside the graph): they are the same on every prove of
  // the bucket and stay in its staging area, so a replay copies nothing.
  void flush_staging() {
    if (mode_ == CAPTURE && in_region_) stg().flush_sync(ctx_.side_stream());
    else stg().flush(s_);
    ctx_.timer.launch_point(s_);
  }
  ~Job() {
    if (in_region_) {  // an error inside the device phase: leave capture / suppression
      set_launch_suppression(false);
      if (mode_ == CAPTURE) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s_, &g);
        if (g) (void)hipGraphDestroy(g);
      }
      if (gb_) gb_->broken = true;
    }
    set_prelaunch_hook(nullptr, nullptr);
    const double t_in = marks_.empty() ? 0.0 : now_s();
    // after an error work may still be queued on the stream: drain it before
    // the buffers go back to the shared pool
    (void)hipStreamSynchronize(s_);
    for (hipEvent_t e : ph_ev_)
      if (e) (void)hipEventDestroy(e);
    const double t_sync = marks_.empty() ? 0.0 : now_s();
    ctx_.pool.put(std::move(chal_dev_));
    ctx_.pool.put(std::move(fs_dig_));
    for (auto& c : cols_) ctx_.pool.put(std::move(c.buf));
    for (auto& t : trees_) {
      ctx_.pool.put(std::move(t.nodes));
      for (auto& x : t.inj) ctx_.pool.put(std::move(x));
    }
    for (auto& x : keep_) ctx_.pool.put(std::move(x));
    ctx_.pool.put(std::move(in_arena_));
    ctx_.pool.put(std::move(pre_fold_.dev));
    ctx_.pool.put(std::move(fri_salts_));
    if (stg_) ctx_.give_staging(std::move(stg_));
    if (!marks_.empty())
      fprintf(stderr, "[loomgpu] lane %d job end (ms): entered %.3f, synced %.3f, buffers returned %.3f\n", lane_,
              (t_in - t0_run_) * 1e3, (t_sync - t0_run_) * 1e3, (now_s() - t0_run_) * 1e3);
  }

  void run();
  // The proof words up to the opening: hash id, exposed values, commitments
  // and claimed values (cv: instance b's values in canonical order).
  template <class W>
  void emit_head(W& w, int b, const E6* cv) const {
    const size_t nt = d_.trees.size();
    auto put_string = [&](const std::string& x) {
      w.push_back((uint32_t)x.size());
      for (size_t i = 0; i < x.size(); i += 4) {
        uint32_t v = 0;
        for (size_t k = 0; k < 4 && i + k < x.size(); ++k) v |= (uint32_t)(uint8_t)x[i + k] << (8 * k);
        w.push_back(v);
      }
    };
    put_string("poseidon2");
    w.push_back((uint32_t)d_.exposed.size());  // exposed values, sorted by name
    for (auto& x : d_.exposed) {
      put_string(x.name);
      w.push_back(1);  // entries
      w.push_back(0);  // index
      w.insert(w.end(), totals_[b][x.col].v, totals_[b][x.col].v + 6);
    }
    // commitments: the trace and AIR roots; setup roots are the verification key's
    w.push_back((uint32_t)(nt - (d_.n_setup ? 1 : 0)));
    for (size_t t = 0; t < nt; ++t)
      if (d_.trees[t].kind != TREE_SETUP) w.insert(w.end(), roots_[t][b].v, roots_[t][b].v + 8);
    // claimed values
    w.push_back((uint32_t)nt);
    size_t k = 0;
    for (size_t t = 0; t < nt; ++t) {
      w.push_back((uint32_t)d_.trees[t].groups.size());
      for (auto& g : d_.trees[t].groups) {
        for (int rail = 0; rail < 2; ++rail) {
          const auto& sh = rail ? g.ext_shifts : g.base_shifts;
          w.push_back((uint32_t)sh.size());
          for (auto& ss : sh) {
            w.push_back((uint32_t)ss.size());
            for (size_t q = 0; q < ss.size(); ++q, ++k) w.insert(w.end(), cv[k].v, cv[k].v + 6);
          }
        }
      }
    }
  }
  // setup.Setup on the device: uploads, RS-encodes and commits the setup
  // columns (this Job has one instance) into `out`, which then owns them.
  void commit_setup(const uint32_t* const* host_cols, SetupData& out);

 private:
  DevBuf dev(size_t words) {
    const size_t bytes = words * sizeof(uint32_t);
    if (!gb_) return ctx_.pool.get(bytes);
    if (next_buf_ == gb_->bufs.size()) {
      if (mode_ == REPLAY) throw std::runtime_error("graph bucket: allocation sequence changed");
      gb_->bufs.emplace_back(bytes);
      gb_->bytes += bytes;
    }
    DevBuf& b = gb_->bufs[next_buf_++];
    if (b.bytes() != bytes) throw std::runtime_error("graph bucket: allocation sequence changed");
    return DevBuf::view(b.as<void>(), bytes);
  }
  // Pinned host memory for device -> host copies of the device phase (the
  // bucket's, in graph mode: a recorded copy writes to a fixed address).
  uint32_t* pin(size_t bytes) {
    if (!gb_) {
      uint32_t* p = ctx_.pinned_ext(bytes, lane_, n_pins_);
      ++n_pins_;
      return p;
    }
    if (n_pins_ == gb_->pins.size()) {
      if (mode_ == REPLAY) throw std::runtime_error("graph bucket: pinned sequence changed");
      void* p = nullptr;
      LG_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
      gb_->pins.push_back({static_cast<uint32_t*>(p), bytes});
    }
    auto& e = gb_->pins[n_pins_++];
    if (e.second != bytes) throw std::runtime_error("graph bucket: pinned sequence changed");
    return e.first;
  }
  // hipMemcpyAsync on the Job's stream, not issued while launches are
  // suppressed (the recorded graph holds it).
  void copy_async(void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
    if (launch_suppressed() || !bytes) return;
    LG_HIP(hipMemcpyAsync(dst, src, bytes, kind, s_));
  }
  // REPLAY launches the graph at once: the host logic of the region then only
  // re-derives what the prove needs afterwards and rewrites the bucket's
  // staging area with the bytes it already holds (the bucket is this Job's
  // alone, and every staged word of the device phase is determined by the
  // plan, the batch size and the bucket's addresses).
  void region_begin() {
    stg().flush(s_);
    in_region_ = true;
    if (mode_ == CAPTURE) LG_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeRelaxed));
    if (mode_ == REPLAY) {
      LG_HIP(hipGraphLaunch(gb_->exec, s_));
      set_launch_suppression(true);
      stg().set_verify(true);
    }
  }
  void region_end() {
    if (!in_region_) return;
    if (mode_ == REPLAY) {
      stg().set_verify(false);
      // a staged table that differs from the recorded one means the graph ran
      // on stale addresses or values: fail (the caller proves again eagerly)
      if (stg().changed()) throw std::runtime_error("graph bucket: staged tables changed on replay");
    }
    if (mode_ == CAPTURE) {
      stg().flush_sync(ctx_.side_stream());
      hipGraph_t g = nullptr;
      LG_HIP(hipStreamEndCapture(s_, &g));
      const hipError_t e = hipGraphInstantiate(&gb_->exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      LG_HIP(e);
    }
    if (mode_ == REPLAY) {
      set_launch_suppression(false);
      stg().assume_flushed();  // the graph copies what the host logic staged
    }
    in_region_ = false;
    if (mode_ == CAPTURE) LG_HIP(hipGraphLaunch(gb_->exec, s_));
  }
  template <class T>
  const T* stage(const std::vector<T>& v) { return stg().put(v, s_); }
  const uint32_t* const* pres(const NttCfg& c) {
    std::vector<const uint32_t*> v(c.pre, c.pre + NTT_MAX_OUT);
    return stage(v);
  }
  void sync() { LG_HIP(hipStreamSynchronize(s_)); }
  // host timestamps of named points, printed at the end with LOOMGPU_MARKS
  void mark(const char* what) {
    static const bool on = getenv("LOOMGPU_MARKS") != nullptr;
    if (on) marks_.push_back({what, now_s()});
  }
  std::vector<std::pair<std::string, double>> marks_;
  double t0_run_ = 0;
  // Phases that end inside the device region do not wait for the device: an
  // eager prove records an event on its stream there and reports the device
  // time between the events once the first phase outside the region has
  // synchronised; that phase gets the rest of the wall time. A recorded or
  // replayed region reports 0 for its phases (the whole region lands in the
  // next phase).
  hipEvent_t ph_ev_[3] = {nullptr, nullptr, nullptr};
  std::vector<const char*> ph_pending_;
  void phase_start() {
    if (!cb_ || mode_ != EAGER) return;
    LG_HIP(hipEventCreate(&ph_ev_[0]));
    LG_HIP(hipEventRecord(ph_ev_[0], s_));
  }
  void phase(const char* name, double& t0) {
    if (cb_ && in_region_) {
      if (mode_ == EAGER && ph_ev_[0] && ph_pending_.size() < 2) {
        hipEvent_t& e = ph_ev_[1 + ph_pending_.size()];
        LG_HIP(hipEventCreate(&e));
        LG_HIP(hipEventRecord(e, s_));
        ph_pending_.push_back(name);
      } else {
        cb_(name, 0.0);
      }
      return;
    }
    if (cb_) {
      sync();
      double t = now_s();
      double dev = 0;
      for (size_t i = 0; i < ph_pending_.size(); ++i) {
        float ms = 0;
        LG_HIP(hipEventElapsedTime(&ms, ph_ev_[i], ph_ev_[i + 1]));
        cb_(ph_pending_[i], ms * 1e-3);
        dev += ms * 1e-3;
      }
      ph_pending_.clear();
      cb_(name, std::max(0.0, t - t0 - dev));
      t0 = t;
    }
  }

  // LOOMGPU_TRACE=1: host-side milestones of the run, to stderr.
  void trace(const char* what) {
    static const bool on = getenv("LOOMGPU_TRACE") != nullptr;
    if (!on) return;
    sync();
    double t = now_s();
    fprintf(stderr, "[loomgpu] B=%d %-24s %8.3f ms\n", B_, what, (t - t_trace_) * 1e3);
    t_trace_ = t;
  }
  double t_trace_ = now_s();

  void alloc_columns();
  void upload_inputs();
  void public_inputs(const std::vector<std::vector<const uint32_t*>>& vals_dev);
  void gencol();
  void run_steps(size_t r);
  // One step DAG of round r: DAG k of step s, on module mi.
  struct StepDag {
    size_t s, k;
    uint32_t mi;
  };
  Layer eval_step_dags(size_t r, const std::vector<StepDag>& items);
  bool same_dag(size_t r, const StepDag& a, const StepDag& b) const;
  void step_grand_products(size_t r, const std::vector<size_t>& steps);
  void step_multiplicity(size_t r, size_t s);
  void step_multiplicity_general(size_t r, size_t s);
  void step_logups(size_t r, const std::vector<size_t>& steps);
  void finish_round_steps(size_t r);
  void flush_err_flags();
  // Round and __zeta challenges on the device (k_fs_challenge); the host
  // replays them into its transcripts at its first synchronisation after
  // the claimed values (replay_device_fs).
  void round_challenge(size_t r);
  void init_segments(FsMsgBuilder& mb);
  void zeta_challenge();
  void replay_device_fs();
  const uint32_t* chal_operands(const std::vector<uint32_t>& chal_ids);
  void lde_columns(const std::vector<uint32_t>& col_ids);
  // Runs the transforms cfg describes on every job: one LDS-resident launch
  // up to 2^14, the two-pass four-step form above.
  void ntt(const NttCfg& cfg, const std::vector<NttJob>& jobs);
  void commit_tree(size_t t);
  std::vector<Digest> roots_of(TreeBuf& tb);
  void build_tree(TreeBuf& tb, const std::vector<std::vector<const uint32_t*>>& leaf_cols_per_group,
                  const std::vector<std::pair<int, int>>& group_shape, const std::vector<int>& group_rows);
  // A FRI fold challenge to derive from a tree's root in the launch that
  // finishes the tree (see launch_merkle_coop_challenge).
  struct RootChallenge {
    std::vector<uint32_t> prefix;
    const uint32_t* prev;
    uint32_t *dig, *alpha, *roots;
  };
  bool tree_levels(TreeBuf& tb, uint32_t* const* tp_d, const RootChallenge* chal = nullptr);
  // A challenge with a proof of work (fri.WithGrinding): the device salt
  // search over mb || [FSPW, bits, salt], then the digest (dig[b], and its
  // first six words at e6[6 b] when e6 is given); salts[b] gets the salt.
  void grind_challenge(FsMsgBuilder mb, uint32_t* dig, uint32_t* e6, uint32_t* salts);

  // AIR quotients of modules of one class (same size, chunk count and
  // relation kernel code), proved together as (module, instance) pairs.
  void quotient(const std::vector<size_t>& mods);
  void open();

  Plan& plan_;
  const PlanData& d_;
  Context& ctx_;
  const uint32_t grind_bits_ = plan_.grinding();
  const uint32_t* chain_dev_ = nullptr;  // [b][8] digests of the callers' running transcripts
  DevBuf fri_salts_;  // [fold round][b] salts of the fold challenges (with grinding)
  Tables& tb_;
  const std::vector<InstanceInput>& in_;
  std::vector<ProveOutput>& out_;
  PhaseCallback cb_;
  int B_;
  int lane_;
  hipStream_t s_;
  std::unique_ptr<Staging> stg_;
  GraphBucket* gb_;
  Mode mode_;
  Tuning tune_;  // the plan's launch choices for this prove's batch
  const bool skip_fri_;  // prover.SkipFRI: claimed values, no opening
  bool in_region_ = false;
  size_t next_buf_ = 0, n_pins_ = 0;
  // the fold rounds before the level enters, chained on the device
  struct {
    DevBuf dev;
    const uint32_t* prev = nullptr;
    int n = 0;
  } pre_fold_;
  size_t fold_pos_ = 0;  // schedule position of fri_fold_0

  std::vector<ColBuf> cols_;
  DevBuf in_arena_;                                 // uploaded trace columns + permutations
  std::vector<std::vector<const uint32_t*>> perm_ptr_;  // [perm][b] into in_arena_
  std::vector<TreeBuf> trees_;
  std::vector<DevBuf> keep_;
  std::vector<Transcript> ts_;
  std::vector<std::vector<E6>> totals_;  // [b][col] exposed logup totals (host, after the replay)
  std::vector<std::pair<uint32_t, const uint32_t*>> totals_dev_;  // (total column, B x 6 words) of the current round
  // Per-instance error flags (B words each) raised by kernels of the current
  // round, with the message an instance with a set flag fails with; copied
  // back after the round's commit (finish_round_steps).
  std::vector<std::pair<std::string, DevBuf>> pending_err_;
  std::vector<std::string> err_msg_;  // message of flag k in the pinned flag area
  std::vector<std::string> inst_err_;  // [b]: the first failure of instance b ("" if none)
  // Device transcript data.  chal_dev_: [b][col][6] challenge values and
  // logup averages T/N (the DAG kernels' per-instance constants); fs_dig_:
  // [pos][b][8] digests of the round challenges (pos = round) and __zeta
  // (pos = rounds); fs_h_: pinned copies of the digests, the tree roots, the
  // logup totals and the lookup flags, read by replay_device_fs.
  DevBuf chal_dev_, fs_dig_;
  uint32_t* fs_h_ = nullptr;
  size_t fs_roots_off_ = 0, fs_tot_off_ = 0, fs_err_off_ = 0;
  std::vector<uint32_t> fs_init_;                    // bindings of the first challenge (hash id, sizes, setup root)
  std::vector<std::vector<uint32_t>> pi_words_;      // [b] public-input transcript elements (Montgomery)
  const uint32_t* pi_dev_ = nullptr;                 // the same, [b][len] staged before the device phase
  std::map<uint32_t, const uint32_t*> total_ptr_;    // logup total column -> device B x 6 words
  std::vector<uint32_t> total_cols_;
  std::vector<std::vector<Digest>> roots_;  // [tree][b]
  std::vector<E6> zeta_;
};

void Job::alloc_columns() {
  cols_.resize(d_.cols.size());
  std::vector<char> is_total(d_.cols.size(), 0);
  for (auto& x : d_.exposed) is_total[x.col] = 1;
  std::vector<char> committed(d_.cols.size(), 0);
  for (auto& t : d_.trees)
    for (auto& g : t.groups) {
      for (auto c : g.base) committed[c] = 1;
      for (auto c : g.ext) committed[c] = 1;
    }
  for (size_t c = 0; c < d_.cols.size(); ++c) {
    const ColInfo& ci = d_.cols[c];
    ColBuf& cb = cols_[c];
    cb.limbs = ci.limbs();
    cb.n = d_.modules[ci.module].N;
    cb.committed = committed[c] != 0;
    if (ci.src == SRC_CHALLENGE || is_total[c]) continue;  // per-instance constants, kept in chal_dev_
    if (ci.src == SRC_SETUP) {  // the plan's resident setup column, shared by every instance
      const DevBuf& sb = plan_.setup()->cols.at(ci.a0);
      cb.buf = DevBuf::view(sb.as<void>(), sb.bytes());
      cb.blk = 0;
      cb.lde_ready = true;
      continue;
    }
    cb.blk = (size_t)cb.limbs * cb.n * (cb.committed ? 5 : 1);
    cb.buf = dev(cb.blk * B_);
  }
  trees_.resize(d_.trees.size());
  for (size_t t = 0; t < d_.trees.size(); ++t) {
    if (d_.trees[t].kind != TREE_SETUP) continue;
    const TreeBuf& st = plan_.setup()->tree;
    TreeBuf& tb = trees_[t];
    tb.shared = true;
    tb.nleaves = st.nleaves;
    tb.order = st.order;
    tb.inj_width = st.inj_width;
    tb.nodes = DevBuf::view(st.nodes.as<void>(), st.nodes.bytes());
    for (auto& x : st.inj) tb.inj.push_back(DevBuf::view(x.as<void>(), x.bytes()));
  }
}

void Job::commit_setup(const uint32_t* const* host_cols, SetupData& out) {
  const TreeInfo& ti = d_.trees.at(0);
  if (ti.kind != TREE_SETUP) throw std::runtime_error("set_setup: no setup tree");
  cols_.resize(d_.cols.size());
  out.cols.resize(d_.n_setup);
  std::vector<uint32_t> ids;
  for (size_t c = 0; c < d_.cols.size(); ++c) {
    const ColInfo& ci = d_.cols[c];
    if (ci.src != SRC_SETUP) continue;
    ColBuf& cb = cols_[c];
    cb.limbs = ci.field ? 6 : 1;  // an extension column: 6 limb arrays of N words (ExtSetup)
    cb.n = d_.modules[ci.module].N;
    cb.committed = true;
    cb.blk = (size_t)5 * cb.limbs * cb.n;
    out.cols[ci.a0].alloc(cb.blk * 4);
    cb.buf = DevBuf::view(out.cols[ci.a0].as<void>(), cb.blk * 4);
    LG_HIP(hipMemcpyAsync(cb.lag(0, 0), host_cols[ci.a0], (size_t)cb.limbs * cb.n * 4, hipMemcpyHostToDevice, s_));
  }
  for (auto& g : ti.groups) {
    ids.insert(ids.end(), g.base.begin(), g.base.end());
    ids.insert(ids.end(), g.ext.begin(), g.ext.end());
  }
  lde_columns(ids);
  std::vector<std::vector<const uint32_t*>> gcols;
  std::vector<std::pair<int, int>> shape;
  std::vector<int> rows;
  for (auto& g : ti.groups) {
    std::vector<const uint32_t*> cp;
    for (auto c : g.base) cp.push_back(cols_[c].lde(0, 0));
    for (auto c : g.ext)
      for (int l = 0; l < 6; ++l) cp.push_back(cols_[c].lde(0, l));
    gcols.push_back(cp);
    shape.push_back({(int)g.base.size(), (int)g.ext.size()});
    rows.push_back((int)g.N * 4);
  }
  trees_.resize(d_.trees.size());
  build_tree(trees_[0], gcols, shape, rows);
  sync();
  LG_HIP(hipMemcpyAsync(out.root, trees_[0].node_ptr(0) + 8, 32, hipMemcpyDeviceToHost, s_));
  sync();
  // the tree's buffers come from the context pool: the key takes them over
  out.tree.nleaves = trees_[0].nleaves;
  out.tree.order = trees_[0].order;
  out.tree.inj_width = trees_[0].inj_width;
  out.tree.nodes = std::move(trees_[0].nodes);
  out.tree.inj = std::move(trees_[0].inj);
  trees_[0].inj.clear();
}

// Trace columns and permutations go through one pinned buffer: the copies
// into it run on the host thread pool, then one DMA per buffer.
void Job::upload_inputs() {
  const size_t npub = d_.publics.size();
  for (int b = 0; b < B_; ++b) {
    if (in_[b].columns.size() != d_.inputs.size() + npub) throw std::runtime_error("wrong number of input columns");
    if (in_[b].perms.size() != d_.perms.size()) throw std::runtime_error("wrong number of permutations");
  }
  for (size_t k = 0; k < d_.inputs.size(); ++k)
    if (cols_[d_.inputs[k]].limbs != 1) throw std::runtime_error("extension-field input columns are not supported");
  const size_t nin = d_.inputs.size();
  if (in_[0].device) {
    // Inputs already in HBM: columns are copied into their buffers by one
    // launch per module size; permutations are read where they are.
    for (int b = 0; b < B_; ++b) {
      if (!in_[b].device) throw std::runtime_error("lg_prove: mixed host and device inputs");
      for (size_t k = 0; k < nin + npub + d_.perms.size(); ++k) {
        const void* ptr = k < nin + npub ? (const void*)in_[b].columns[k] : (const void*)in_[b].perms[k - nin - npub];
        if (k >= nin && k < nin + npub && d_.publics[k - nin].rows.empty()) continue;  // no entries: not read
        hipPointerAttribute_t at;
        if (hipPointerGetAttributes(&at, ptr) != hipSuccess || at.type != hipMemoryTypeDevice ||
            at.device != ctx_.device) {
          (void)hipGetLastError();
          throw std::runtime_error("lg_prove_device: input pointer is not device memory of the context's device");
        }
      }
    }
    std::map<uint32_t, std::pair<std::vector<const uint32_t*>, std::vector<uint32_t*>>> by_n;
    for (int b = 0; b < B_; ++b)
      for (size_t k = 0; k < nin; ++k) {
        const ColBuf& cb = cols_[d_.inputs[k]];
        auto& e = by_n[cb.n];
        e.first.push_back(in_[b].columns[k]);
        e.second.push_back(cb.lag(b, 0));
      }
    for (auto& kv : by_n)
      launch_copy_cols(stage(kv.second.first), const_cast<uint32_t* const*>(stage(kv.second.second)), (int)kv.first,
                       (int)kv.second.first.size(), s_);
    perm_ptr_.assign(d_.perms.size(), std::vector<const uint32_t*>(B_));
    for (size_t p = 0; p < d_.perms.size(); ++p)
      for (int b = 0; b < B_; ++b) perm_ptr_[p][b] = in_[b].perms[p];
    if (gb_ && !d_.perms.empty()) {
      // A recorded graph reads the addresses it was recorded with, and the
      // permutations are read inside it (k_gencol): copy them into the
      // bucket's own buffer, whose address every prove of the bucket shares.
      size_t words = 0;
      for (auto& pm : d_.perms) words += (size_t)pm.len * B_;
      in_arena_ = dev(words);
      std::map<uint32_t, std::pair<std::vector<const uint32_t*>, std::vector<uint32_t*>>> by_len;
      size_t off = 0;
      for (size_t p = 0; p < d_.perms.size(); ++p)
        for (int b = 0; b < B_; ++b) {
          uint32_t* dst = in_arena_.as<uint32_t>() + off;
          off += d_.perms[p].len;
          auto& e = by_len[d_.perms[p].len];
          e.first.push_back(in_[b].perms[p]);
          e.second.push_back(dst);
          perm_ptr_[p][b] = dst;
        }
      for (auto& kv : by_len)
        if (kv.first)
          launch_copy_cols(stage(kv.second.first), const_cast<uint32_t* const*>(stage(kv.second.second)),
                           (int)kv.first, (int)kv.second.first.size(), s_);
    }
    std::vector<std::vector<const uint32_t*>> pv(npub, std::vector<const uint32_t*>(B_));
    for (size_t i = 0; i < npub; ++i)
      for (int b = 0; b < B_; ++b) pv[i][b] = in_[b].columns[nin + i];
    public_inputs(pv);
    return;
  }
  // host layout: [b][input k] then [perm p][b], then [public i][b]
  size_t words = 0;
  std::vector<size_t> col_off((size_t)B_ * d_.inputs.size()), perm_off(d_.perms.size() * B_);
  for (int b = 0; b < B_; ++b)
    for (size_t k = 0; k < d_.inputs.size(); ++k) {
      col_off[(size_t)b * d_.inputs.size() + k] = words;
      words += cols_[d_.inputs[k]].n;
    }
  for (size_t p = 0; p < d_.perms.size(); ++p)
    for (int b = 0; b < B_; ++b) {
      perm_off[p * B_ + b] = words;
      words += d_.perms[p].len;
    }
  std::vector<size_t> pub_off(npub * B_);
  for (size_t i = 0; i < npub; ++i)
    for (int b = 0; b < B_; ++b) {
      pub_off[i * B_ + b] = words;
      words += d_.publics[i].words();
    }
  uint32_t* h = ctx_.pinned_in(words * 4, lane_);
  parallel_for((int)(B_ * (nin + d_.perms.size())), [&](int i) {
    const int b = i % B_;
    const size_t k = (size_t)i / B_;
    if (k < nin) std::memcpy(h + col_off[(size_t)b * nin + k], in_[b].columns[k], cols_[d_.inputs[k]].n * 4);
    else std::memcpy(h + perm_off[(k - nin) * B_ + b], in_[b].perms[k - nin], d_.perms[k - nin].len * 4);
  });
  for (size_t i = 0; i < npub; ++i)
    for (int b = 0; b < B_; ++b)
      std::memcpy(h + pub_off[i * B_ + b], in_[b].columns[nin + i], d_.publics[i].words() * 4);
  in_arena_ = dev(words);
  LG_HIP(hipMemcpyAsync(in_arena_.as<void>(), h, words * 4, hipMemcpyHostToDevice, s_));
  // columns of one module size share a scatter launch
  std::map<uint32_t, std::pair<std::vector<uint64_t>, std::vector<uint32_t*>>> by_n;
  for (int b = 0; b < B_; ++b)
    for (size_t k = 0; k < nin; ++k) {
      const ColBuf& cb = cols_[d_.inputs[k]];
      auto& e = by_n[cb.n];
      e.first.push_back(col_off[(size_t)b * nin + k]);
      e.second.push_back(cb.lag(b, 0));
    }
  for (auto& kv : by_n)
    launch_scatter(in_arena_.as<uint32_t>(), stage(kv.second.first), const_cast<uint32_t* const*>(stage(kv.second.second)),
                   (int)kv.first, (int)kv.second.first.size(), s_);
  perm_ptr_.assign(d_.perms.size(), std::vector<const uint32_t*>(B_));
  for (size_t p = 0; p < d_.perms.size(); ++p)
    for (int b = 0; b < B_; ++b) perm_ptr_[p][b] = in_arena_.as<uint32_t>() + perm_off[p * B_ + b];
  std::vector<std::vector<const uint32_t*>> pv(npub, std::vector<const uint32_t*>(B_));
  for (size_t i = 0; i < npub; ++i)
    for (int b = 0; b < B_; ++b) pv[i][b] = in_arena_.as<uint32_t>() + pub_off[i * B_ + b];
  public_inputs(pv);
}

// Public inputs (prover.go:236-273): each column is zero except at its entry
// rows, which take the instance's values; their transcript elements
// (public.Inputs.TranscriptElements, public/public_inputs.go:50-74: PUBI,
// #inputs, then per input its tagged name and module, #entries and
// (row, field 0, value) per entry) are bound to the first challenge.  The
// per-instance words are staged before the device phase, so a recorded
// graph reads this prove's values.
void Job::public_inputs(const std::vector<std::vector<const uint32_t*>>& vals_dev) {
  const size_t npub = d_.publics.size();
  pi_words_.assign(B_, std::vector<uint32_t>());
  if (!npub) return;
  const size_t nin = d_.inputs.size();
  for (size_t i = 0; i < npub; ++i) {
    const PublicCol& pc = d_.publics[i];
    const ColBuf& cb = cols_[pc.col];
    std::vector<uint32_t*> dst(B_);
    for (int b = 0; b < B_; ++b) dst[b] = cb.lag(b, 0);
    uint32_t* const* dst_d = const_cast<uint32_t* const*>(stage(dst));
    // every limb of the column is zero but the entries; value word j goes to
    // limb l of its row (limb arrays of n words each; a base entry in an
    // extension column keeps its other limbs zero: the lift)
    launch_fill(dst_d, (int)(cb.n * cb.limbs), 0u, B_, s_);
    std::vector<uint32_t> at;
    for (size_t e = 0; e < pc.rows.size(); ++e)
      for (uint32_t l = 0; l < (pc.fields[e] ? 6u : 1u); ++l) at.push_back(l * cb.n + pc.rows[e]);
    if (!at.empty()) launch_scatter_rows(stage(vals_dev[i]), stage(at), (int)at.size(), dst_d, B_, s_);
  }
  // the values on the host: the caller's arrays, or a copy of the device ones
  std::vector<std::vector<uint32_t>> hv(npub * B_);
  for (size_t i = 0; i < npub; ++i)
    for (int b = 0; b < B_; ++b) {
      const size_t k = d_.publics[i].words();
      hv[i * B_ + b].resize(k);
      if (!k) continue;
      if (in_[b].device) {
        LG_HIP(hipMemcpyAsync(hv[i * B_ + b].data(), in_[b].columns[nin + i], k * 4, hipMemcpyDeviceToHost, s_));
        sync();
      }
      else std::memcpy(hv[i * B_ + b].data(), in_[b].columns[nin + i], k * 4);
    }
  std::vector<uint32_t> all;
  for (int b = 0; b < B_; ++b) {
    std::vector<const uint32_t*> vals(npub);
    for (size_t i = 0; i < npub; ++i) vals[i] = hv[i * B_ + b].data();
    std::vector<uint32_t>& w = pi_words_[b];
    w = public_input_words(d_, vals);
    if (b && w.size() != pi_words_[0].size()) throw std::runtime_error("public inputs: instance shapes differ");
    all.insert(all.end(), w.begin(), w.end());
  }
  pi_dev_ = stage(all);
}

void Job::gencol() {
  std::vector<GenColJob> jobs;
  int blocks = 0;
  uint32_t* sig_err = nullptr;
  for (size_t c = 0; c < d_.cols.size(); ++c) {
    const ColInfo& ci = d_.cols[c];
    const ColBuf& cb = cols_[c];
    if (ci.src != SRC_LAGRANGE && ci.src != SRC_ID && ci.src != SRC_SIGMA && ci.src != SRC_RANGE) continue;
    const int logn = d_.modules[ci.module].log_n;
    std::vector<uint32_t*> outs(B_);
    for (int b = 0; b < B_; ++b) outs[b] = cb.lag(b, 0);
    GenColJob j{};
    j.n = (int)cb.n;
    j.logn = logn;
    j.outs = const_cast<uint32_t* const*>(stage(outs));
    if (ci.src == SRC_RANGE) {
      j.kind = GENCOL_RANGE;
      j.a0 = ci.a0;
    } else if (ci.src == SRC_LAGRANGE) {
      if (ci.a0 >= cb.n) throw std::runtime_error("Lagrange column position out of range");
      j.kind = GENCOL_LAGRANGE;
      j.a0 = ci.a0;
    } else if (ci.src == SRC_ID) {
      j.kind = GENCOL_ID;
      j.wpow = tb_.wpow(logn);
      j.gk = kb_pow(M(MULT_GEN), ci.a0);
    } else {
      const Perm& pm = d_.perms[ci.a1];
      std::vector<const uint32_t*> S(B_);
      for (int b = 0; b < B_; ++b) S[b] = perm_ptr_[ci.a1][b];
      if (!sig_err) {
        DevBuf e = dev((size_t)B_);
        sig_err = e.as<uint32_t>();
        std::vector<uint32_t*> ep(B_);
        for (int b = 0; b < B_; ++b) ep[b] = sig_err + b;
        launch_fill(const_cast<uint32_t* const*>(stage(ep)), 1, 0u, B_, s_);
        pending_err_.push_back({"copy constraint: a permutation entry is outside the permutation's cells", std::move(e)});
      }
      j.kind = GENCOL_SIGMA;
      j.a0 = ci.a0;
      j.wpow = tb_.wpow(logn);
      j.gpow = tb_.gpow_small(pm.len / cb.n);
      j.S = stage(S);
      j.len = pm.len;
      j.err = sig_err;
    }
    j.block0 = blocks;
    blocks += (j.n + 255) / 256;
    jobs.push_back(j);
  }
  if (!jobs.empty()) launch_gencol(stage(jobs), (int)jobs.size(), blocks, B_, s_);
}

}  // namespace

void run_ntt(Tables& tb, const NttCfg& cfg, const std::vector<NttJob>& jobs, hipStream_t s, const StageFn& stage,
             const ScratchFn& scratch) {
  const int logn = cfg.logn;
  if (jobs.empty()) return;
  if (logn <= NTT_MAX_LOG_LDS) {
    const NttJob* jd = static_cast<const NttJob*>(stage(jobs.data(), jobs.size() * sizeof(NttJob)));
    const uint32_t* const* pd = static_cast<const uint32_t* const*>(stage(cfg.pre, sizeof(cfg.pre)));
    launch_ntt(jd, (int)jobs.size(), cfg, pd, tb.tw_fwd(logn), tb.tw_inv(logn), tb.ntt_ladders(false),
               tb.ntt_ladders(true), s);
    return;
  }
  const size_t n = (size_t)1 << logn;
  const int la = (logn + 1) / 2, lb = logn - la;  // sub-transform sizes 2^la (pass B) and 2^lb (pass A)
  // Fused form of an inverse followed by forward outputs (the LDE, the
  // quotient cosets, the DEEP LDE), when every pre-scale is a geometric table
  // and no job has a post-scale: the inverse's pass A, one middle pass (the
  // inverse's row transform and every output's first pass, factors swapped:
  // launch_lde_mid) and every output's pass B: three passes per job instead
  // of 2 (1 + n_out).  Scratch per job: n words for the pass-A output and n
  // per output.
  bool fuse = cfg.inverse_first && cfg.n_out >= 1;
  uint32_t geo_s[NTT_MAX_OUT] = {0};
  for (int o = 0; fuse && o < cfg.n_out; ++o)
    fuse = cfg.pre[o] && tb.geometric(cfg.pre[o], &geo_s[o]) && cfg.bitrev_out[o] == cfg.bitrev_out[0];
  for (size_t k = 0; fuse && k < jobs.size(); ++k) fuse = jobs[k].post == nullptr;
  if (fuse) {
    const size_t per_job = n * (1 + (size_t)cfg.n_out);
    const size_t max_words = (size_t)256 << 20;
    const size_t G = std::min(jobs.size(), std::max<size_t>(1, max_words / per_job));
    uint32_t* sp = scratch(G * per_job);
    LdeMidArgs ma{};
    ma.a = la;
    ma.b = lb;
    ma.n_out = cfg.n_out;
    for (int o = 0; o < cfg.n_out; ++o) {
      ma.pre[o] = cfg.pre[o];
      ma.blk[o] = tb.geo_blocks(geo_s[o], la, lb);
    }
    ma.lad_inv = tb.ntt_ladders(true);
    ma.lad_fwd = tb.ntt_ladders(false);
    ma.tw_inv = tb.tw_inv(la);
    ma.tw_fwd = tb.tw_fwd(la);
    ma.wlo = tb.wpow(logn);
    ma.whi = tb.wpow(logn - 11);
    for (size_t j0 = 0; j0 < jobs.size(); j0 += G) {
      const size_t g = std::min(G, jobs.size() - j0);
      auto tmp = [&](size_t k) { return sp + k * per_job; };
      auto ubuf = [&](size_t k, int o) { return sp + k * per_job + n * (1 + (size_t)o); };
      std::vector<BigJob> ja(g), jb(g * cfg.n_out);
      std::vector<LdeMidJob> jm(g);
      for (size_t k = 0; k < g; ++k) {
        const NttJob& J = jobs[j0 + k];
        ja[k] = BigJob{J.in, nullptr, tmp(k), nullptr, nullptr};
        jm[k].tmp = tmp(k);
        for (int o = 0; o < cfg.n_out; ++o) {
          jm[k].u[o] = ubuf(k, o);
          jb[k * cfg.n_out + o] = BigJob{nullptr, J.out[o], ubuf(k, o), nullptr, nullptr};
        }
      }
      const BigJob* da = static_cast<const BigJob*>(stage(ja.data(), g * sizeof(BigJob)));
      launch_ntt_pass_a(da, (int)g, la, lb, tb.tw_inv(lb), tb.wpow_inv(logn), tb.wpow_inv(logn - 11), s);
      const LdeMidJob* dm = static_cast<const LdeMidJob*>(stage(jm.data(), g * sizeof(LdeMidJob)));
      launch_lde_mid(dm, (int)g, ma, s);
      const BigJob* db = static_cast<const BigJob*>(stage(jb.data(), jb.size() * sizeof(BigJob)));
      launch_ntt_pass_b(db, (int)jb.size(), lb, la, tb.tw_fwd(lb), cfg.bitrev_out[0], s);
    }
    return;
  }
  // Four-step form: every transform is two launches over a group of jobs;
  // per job an n-word transpose buffer, plus an n-word buffer for the
  // inverse result when forward outputs follow.  Groups bound the scratch.
  const bool need_x = cfg.inverse_first && cfg.n_out > 0;
  const size_t per_job = n * (need_x ? 2 : 1);
  const size_t max_words = (size_t)256 << 20;
  const size_t G = std::min(jobs.size(), std::max<size_t>(1, max_words / per_job));
  uint32_t* sp = scratch(G * per_job);
  for (size_t j0 = 0; j0 < jobs.size(); j0 += G) {
    const size_t g = std::min(G, jobs.size() - j0);
    std::vector<BigJob> bj(g);
    auto tmp = [&](size_t k) { return sp + k * per_job; };
    auto xbuf = [&](size_t k) { return sp + k * per_job + n; };
    if (cfg.inverse_first) {
      for (size_t k = 0; k < g; ++k) {
        const NttJob& J = jobs[j0 + k];
        bj[k] = BigJob{J.in, need_x ? xbuf(k) : J.out[0], tmp(k), nullptr, J.post};
      }
      const BigJob* d = static_cast<const BigJob*>(stage(bj.data(), g * sizeof(BigJob)));
      launch_ntt_big(d, d, (int)g, logn, tb.tw_inv(lb), tb.tw_inv(la), tb.wpow_inv(logn), tb.wpow_inv(logn - 11), 0, s);
    }
    for (int o = 0; o < cfg.n_out; ++o) {
      for (size_t k = 0; k < g; ++k) {
        const NttJob& J = jobs[j0 + k];
        bj[k] = BigJob{cfg.inverse_first ? xbuf(k) : J.in, J.out[o], tmp(k), cfg.pre[o], nullptr};
      }
      const BigJob* d = static_cast<const BigJob*>(stage(bj.data(), g * sizeof(BigJob)));
      launch_ntt_big(d, d, (int)g, logn, tb.tw_fwd(lb), tb.tw_fwd(la), tb.wpow(logn), tb.wpow(logn - 11),
                     cfg.bitrev_out[o], s);
    }
  }
}

namespace {

void Job::ntt(const NttCfg& cfg, const std::vector<NttJob>& jobs) {
  run_ntt(
      tb_, cfg, jobs, s_, [&](const void* p, size_t bytes) -> const void* { return stg().put(p, bytes, s_); },
      [&](size_t words) {
        keep_.push_back(dev(words));
        return keep_.back().as<uint32_t>();
      });
}

void Job::lde_columns(const std::vector<uint32_t>& col_ids) {
  std::map<int, std::vector<NttJob>> by_log;
  for (uint32_t c : col_ids) {
    ColBuf& cb = cols_[c];
    if (cb.lde_ready) continue;
    const int logn = d_.modules[d_.cols[c].module].log_n;
    for (int b = 0; b < B_; ++b)
      for (int l = 0; l < cb.limbs; ++l) {
        NttJob j{};
        j.in = cb.lag(b, l);
        j.post = nullptr;
        for (int o = 0; o < 4; ++o) j.out[o] = cb.lde(b, l) + (size_t)o * cb.n;
        by_log[logn].push_back(j);
      }
    cb.lde_ready = true;
  }
  for (auto& kv : by_log) {
    NttCfg cfg{};
    cfg.kind = NTT_KIND_LDE;
    cfg.logn = kv.first;
    cfg.inverse_first = 1;
    cfg.n_out = 4;
    for (int o = 0; o < 4; ++o) {
      cfg.pre[o] = tb_.lde_pre(kv.first, o, RATE_LOG, true);
      cfg.bitrev_out[o] = 1;
    }
    const double n = double(1u << kv.first);
    TimeIt t(ctx_, s_, "ntt_lde", kv.second.size() * 5.0 * n * 4, 0);
    ntt(cfg, kv.second);
  }
}

// Hash every group, then build the tree with the smaller groups as level
// injections (fri/commitment.go:259-386, merkle/tree.go:210-296).
void Job::build_tree(TreeBuf& tb, const std::vector<std::vector<const uint32_t*>>& gcols,
                     const std::vector<std::pair<int, int>>& gshape, const std::vector<int>& grows) {
  const size_t ng = gcols.size() / B_;
  tb.order.resize(ng);
  for (size_t k = 0; k < ng; ++k) tb.order[k] = k;
  std::stable_sort(tb.order.begin(), tb.order.end(), [&](size_t a, size_t b) { return grows[a] > grows[b]; });
  tb.nleaves = grows[tb.order[0]] / 2;
  tb.nodes = dev((size_t)B_ * 2 * tb.nleaves * 8);
  tb.inj.clear();
  tb.inj_width.clear();
  // Groups of one shape (base and extension columns) and width hash in one
  // launch: the (group, instance) pairs form the launch's batch.
  struct LeafClass {
    int nb, ne, width;
    std::vector<const uint32_t*> cp;  // [member][b][column]
    std::vector<uint32_t*> outs;      // [member][b]
  };
  std::vector<LeafClass> classes;
  for (size_t k = 0; k < ng; ++k) {
    const size_t g = tb.order[k];
    const int nb = gshape[g].first, ne = gshape[g].second;
    const int width = grows[g] / 2;
    if (k > 0) {
      tb.inj.push_back(dev((size_t)B_ * width * 8));
      tb.inj_width.push_back(width);
    }
    LeafClass* lc = nullptr;
    for (auto& c : classes)
      if (c.nb == nb && c.ne == ne && c.width == width) lc = &c;
    if (!lc) {
      classes.push_back(LeafClass{nb, ne, width, {}, {}});
      lc = &classes.back();
    }
    for (int b = 0; b < B_; ++b) {
      lc->cp.insert(lc->cp.end(), gcols[g * B_ + b].begin(), gcols[g * B_ + b].end());
      lc->outs.push_back(k == 0 ? tb.node_ptr(b) + (size_t)8 * tb.nleaves : tb.inj_ptr(tb.inj.size() - 1, b));
    }
  }
  for (auto& lc : classes) {
    const int nb = lc.nb, ne = lc.ne, ncols = nb + 6 * ne, width = lc.width;
    const int nbatch = (int)lc.outs.size();
    std::vector<int32_t> msg(3, 0);
    for (int i = 0; i < nb; ++i) msg.push_back(2 * i);
    for (int i = 0; i < nb; ++i) msg.push_back(2 * i + 1);
    for (int i = 0; i < 6 * ne; ++i) msg.push_back(2 * (nb + i));
    for (int i = 0; i < 6 * ne; ++i) msg.push_back(2 * (nb + i) + 1);
    const uint32_t* const* cp_d = stage(lc.cp);
    const int32_t* msg_d = stage(msg);
    uint32_t* const* outs_d = const_cast<uint32_t* const*>(stage(lc.outs));
    TimeIt tm(ctx_, s_, "leaf_hash", (double)nbatch * width * (2.0 * ncols + 8) * 4,
              (double)nbatch * width * ((msg.size() + 15) / 16));
    launch_leaf_hash(cp_d, ncols, msg_d, (int)msg.size(), M(2 * nb), M(2 * ne), outs_d, width, nbatch, s_,
                     tune_.hash_threads);
  }
  std::vector<uint32_t*> tp(B_);
  for (int b = 0; b < B_; ++b) tp[b] = tb.node_ptr(b);
  tb.nodes_d = const_cast<uint32_t* const*>(stage(tp));
  tree_levels(tb, tb.nodes_d);
}

// The inner levels of a tree whose leaves (and injection digests) are in
// place; tp_d: the staged node-array pointers of the batch.  With chal, the
// launch that reaches the root also derives the challenge; returns whether it
// did (a tree of one leaf pair has no such launch).
bool Job::tree_levels(TreeBuf& tb, uint32_t* const* tp_d, const RootChallenge* chal) {
  bool chal_done = false;
  size_t next = 0;
  // every injection at one level width in one launch, applied per node in
  // schedule order
  auto fold = [&](int width) {
    std::vector<const uint32_t*> ip;  // [injection][b]
    int count = 0;
    while (next < tb.inj.size() && tb.inj_width[next] == width) {
      for (int b = 0; b < B_; ++b) ip.push_back(tb.inj_ptr(next, b));
      ++count;
      ++next;
    }
    if (!count) return;
    const uint32_t* const* ip_d = stage(ip);
    TimeIt tm(ctx_, s_, "merkle_inject", (double)B_ * width * (2 + count) * 32, (double)B_ * width * count);
    launch_inject_seq(tp_d, ip_d, count, width, B_, s_);
  };
  auto has_inj = [&](int width) {
    for (size_t k = next; k < tb.inj.size(); ++k)
      if (tb.inj_width[k] == width) return true;
    return false;
  };
  fold(tb.nleaves);
  int w = tb.nleaves / 2;
  // Wide levels: one launch per level keeps every lane busy.  Once a level
  // has few nodes per batch, the remaining levels run in one fused launch
  // (workgroup-local subtrees), trading idle lanes for fewer launches.
  while (w >= 1 && (size_t)w * B_ >= (size_t)tune_.merkle_wide_min) {
    {
      TimeIt tm(ctx_, s_, "merkle", (double)B_ * w * 3 * 32, (double)B_ * w);
      launch_merkle_level(tp_d, w, B_, s_, tune_.hash_threads);
    }
    fold(w);
    w /= 2;
  }
  // Narrow levels: cooperative permutations (an octet of lanes per node),
  // up to 6 levels per launch on workgroup-local subtrees.
  while (w >= 1) {
    const int G = w < 32 ? w : 32;
    int lo = w;
    while (lo > 1 && !has_inj(lo) && (w / (lo / 2)) <= G) lo /= 2;
    {
      int span = 0;
      for (int x = w; x >= lo; x /= 2) span += x;
      TimeIt tm(ctx_, s_, "merkle_coop", (double)B_ * span * 3 * 32, (double)B_ * span);
      if (chal && lo == 1 && next == tb.inj.size()) {
        launch_merkle_coop_challenge(tp_d, w, chal->prefix.data(), (int)chal->prefix.size(), chal->prev, chal->dig,
                                     chal->alpha, chal->roots, B_, s_);
        chal_done = true;
      } else {
        launch_merkle_coop(tp_d, w, lo, B_, s_);
      }
    }
    fold(lo);
    w = lo / 2;
  }
  if (next != tb.inj.size()) throw std::runtime_error("merkle: injection width does not match a level");
  return chal_done;
}

std::vector<Digest> Job::roots_of(TreeBuf& tb) {
  const uint32_t* const* tp_d = tb.nodes_d;
  if (!tp_d) {
    std::vector<const uint32_t*> tp(B_);
    for (int b = 0; b < B_; ++b) tp[b] = tb.node_ptr(b);
    tp_d = stage(tp);
  }
  DevBuf r = dev((size_t)B_ * 8);
  launch_copy_roots(tp_d, r.as<uint32_t>(), B_, s_);
  std::vector<Digest> out(B_);
  LG_HIP(hipMemcpyAsync(out.data(), r.as<void>(), (size_t)B_ * 32, hipMemcpyDeviceToHost, s_));
  mark("roots enqueued");
  sync();
  mark("roots synced");
  keep_.push_back(std::move(r));
  return out;
}

// Diagnostic: with LOOMGPU_DEBUG_DIR set, writes instance 0's Lagrange
// values of every column of a committed tree (and its LDE) to that directory.
static void debug_dump(const char* tag, const std::vector<uint32_t>& ids, const std::vector<ColBuf>& cols,
                       hipStream_t s) {
  const char* dir = getenv("LOOMGPU_DEBUG_DIR");
  if (!dir) return;
  LG_HIP(hipStreamSynchronize(s));
  for (uint32_t c : ids) {
    const ColBuf& cb = cols[c];
    std::vector<uint32_t> h(cb.blk);
    LG_HIP(hipMemcpy(h.data(), cb.lag(0, 0), cb.blk * 4, hipMemcpyDeviceToHost));
    std::string fn = std::string(dir) + "/" + tag + "_col" + std::to_string(c) + ".bin";
    FILE* f = fopen(fn.c_str(), "wb");
    if (f) {
      fwrite(h.data(), 4, h.size(), f);
      fclose(f);
    }
  }
}

void Job::commit_tree(size_t t) {
  const TreeInfo& ti = d_.trees[t];
  std::vector<uint32_t> ids;
  for (auto& g : ti.groups) {
    ids.insert(ids.end(), g.base.begin(), g.base.end());
    ids.insert(ids.end(), g.ext.begin(), g.ext.end());
  }
  lde_columns(ids);
  std::vector<std::vector<const uint32_t*>> gcols;
  std::vector<std::pair<int, int>> shape;
  std::vector<int> rows;
  for (auto& g : ti.groups) {
    for (int b = 0; b < B_; ++b) {
      std::vector<const uint32_t*> cp;
      for (auto c : g.base) cp.push_back(cols_[c].lde(b, 0));
      for (auto c : g.ext)
        for (int l = 0; l < 6; ++l) cp.push_back(cols_[c].lde(b, l));
      gcols.push_back(std::move(cp));
    }
    shape.push_back({(int)g.base.size(), (int)g.ext.size()});
    rows.push_back((int)g.N * 4);
  }
  build_tree(trees_[t], gcols, shape, rows);
  {
    // the roots come back with the transcript data (replay_device_fs)
    DevBuf r = dev((size_t)B_ * 8);
    launch_copy_roots(trees_[t].nodes_d, r.as<uint32_t>(), B_, s_);
    copy_async(fs_h_ + fs_roots_off_ + t * (size_t)B_ * 8, r.as<void>(), (size_t)B_ * 32, hipMemcpyDeviceToHost);
    keep_.push_back(std::move(r));
  }
  debug_dump(("tree" + std::to_string(t)).c_str(), ids, cols_, s_);
}

// Evaluates input DAG k of step (r, s) on the rows of module mi (Lagrange
// values), into an E6 layer.
// Step DAGs of one launch: items that share the kernel code, the module size
// and the slot counts, with (item, instance) pairs as the launch batch; the
// output Layer holds item i's instance b at batch index i B + b.
Layer Job::eval_step_dags(size_t r, const std::vector<StepDag>& it) {
  const DagKernel& K0 = *plan_.step_kernel(r, it[0].s, it[0].k);
  const ModuleInfo& m = d_.modules[it[0].mi];
  const size_t NI = it.size(), nchal = K0.chal_ids.size();
  Layer out;
  out.len = m.N;
  out.buf = dev(NI * B_ * 6 * m.N);
  std::vector<const uint32_t*> cp;
  std::vector<uint32_t*> op;
  uint32_t* chal = nullptr;
  if (nchal) {
    keep_.push_back(dev(NI * B_ * nchal * 6));
    chal = keep_.back().as<uint32_t>();
  }
  for (size_t i = 0; i < NI; ++i) {
    const DagKernel& K = *plan_.step_kernel(r, it[i].s, it[i].k);
    for (int b = 0; b < B_; ++b) {
      for (auto c : K.col_ids) cp.push_back(cols_[c].lag(b, 0));
      op.push_back(out.ptr((int)(i * B_ + b)));
    }
    if (nchal)  // challenge operands of item i: [b][slot][6]
      launch_chal_compact(chal_dev_.as<uint32_t>(), (int)d_.cols.size(), stage(K.chal_ids), (int)nchal,
                          chal + i * B_ * nchal * 6, B_, s_);
  }
  DagLaunch a{};
  a.cols = stage(cp);
  a.chal = chal;
  a.rowscale = nullptr;
  a.out = const_cast<uint32_t* const*>(stage(op));
  a.col_stride = m.N;
  a.out_stride = m.N;
  a.logn = m.log_n;
  a.ncos = 1;
  a.batch = (int)(NI * B_);
  dag_kernel_launch(K0, a, s_);
  return out;
}

bool Job::same_dag(size_t r, const StepDag& a, const StepDag& b) const {
  const DagKernel& A = *plan_.step_kernel(r, a.s, a.k);
  const DagKernel& B = *plan_.step_kernel(r, b.s, b.k);
  return d_.modules[a.mi].N == d_.modules[b.mi].N && A.code == B.code && A.col_ids.size() == B.col_ids.size() &&
         A.chal_ids.size() == B.chal_ids.size();
}

// The steps of a round are independent (each writes its own columns); steps
// of one kind whose modules have one size and whose DAGs share their code
// run as one class: each phase is one launch for the whole class, the
// (step, instance) pairs its batch.
void Job::run_steps(size_t r) {
  const Round& rd = d_.rounds[r];
  std::vector<std::vector<size_t>> gp, lu;
  auto place = [&](std::vector<std::vector<size_t>>& cls, size_t s, size_t ndag) {
    for (auto& c : cls) {
      bool ok = true;
      for (size_t k = 0; ok && k < ndag; ++k)
        ok = same_dag(r, {c[0], k, rd.steps[c[0]].module}, {s, k, rd.steps[s].module});
      if (ok) {
        c.push_back(s);
        return;
      }
    }
    cls.push_back({s});
  };
  for (size_t s = 0; s < rd.steps.size(); ++s) {
    const Step& st = rd.steps[s];
    switch (st.kind) {
      case STEP_GRAND_PRODUCT: place(gp, s, 2); break;
      case STEP_LOOKUP_MULTIPLICITY: step_multiplicity(r, s); break;
      case STEP_CYCLIC_LOGUP: {
        bool placed = false;
        for (auto& c : lu)
          if (rd.steps[c[0]].params[0] == st.params[0]) {
            bool ok = true;
            for (size_t k = 0; ok && k < 2 * (size_t)st.params[0]; ++k)
              ok = same_dag(r, {c[0], k, rd.steps[c[0]].module}, {s, k, st.module});
            if (ok) {
              c.push_back(s);
              placed = true;
              break;
            }
          }
        if (!placed) lu.push_back({s});
        break;
      }
      default: throw std::runtime_error("unsupported prover step");
    }
  }
  for (auto& c : gp) step_grand_products(r, c);
  for (auto& c : lu) step_logups(r, c);
}

// board.GrandProductStep (board/step.go:301-327) for a class of steps.
void Job::step_grand_products(size_t r, const std::vector<size_t>& steps) {
  const size_t NI = steps.size();
  const uint32_t n = d_.modules[d_.rounds[r].steps[steps[0]].module].N;
  std::vector<StepDag> num, den;
  for (size_t s : steps) {
    num.push_back({s, 0, d_.rounds[r].steps[s].module});
    den.push_back({s, 1, d_.rounds[r].steps[s].module});
  }
  Layer nd[2] = {eval_step_dags(r, num), eval_step_dags(r, den)};
  const int NB = (int)(NI * B_);
  std::vector<const uint32_t*> np(NB), dp(NB);
  std::vector<uint32_t*> gpp(NB), rp(NB);
  Layer ratio;
  ratio.len = n;
  ratio.buf = dev((size_t)NB * 6 * n);
  for (size_t i = 0; i < NI; ++i)
    for (int b = 0; b < B_; ++b) {
      const int x = (int)(i * B_ + b);
      np[x] = nd[0].ptr(x);
      dp[x] = nd[1].ptr(x);
      gpp[x] = cols_[d_.rounds[r].steps[steps[i]].out_col].lag(b, 0);
      rp[x] = ratio.ptr(x);
    }
  uint32_t* scan_tmp = nullptr;
  if (scan_tmp_words((int)n, NB)) {
    keep_.push_back(dev(scan_tmp_words((int)n, NB)));
    scan_tmp = keep_.back().as<uint32_t>();
  }
  launch_grand_product(stage(np), stage(dp), const_cast<uint32_t* const*>(stage(rp)),
                       const_cast<uint32_t* const*>(stage(gpp)), n, NB, scan_tmp, s_);
  keep_.push_back(std::move(ratio.buf));
  keep_.push_back(std::move(nd[0].buf));
  keep_.push_back(std::move(nd[1].buf));
}

// board.LookupMultiplicityStep (board/step.go:257-295) for a range target:
// the target row holding v is row v, so the multiplicity column is the
// histogram of the source values; a value outside the table is an error
// (poly/lookup_multiplicity.go:90-96), reported after the round's commit.
// Sources of one size and DAG code are evaluated and counted in one launch
// each (the counts are atomic additions, so their order does not matter).
// The range shape of arguments.Range (lookup.go:34-51): width 1, no
// selectors, one target that is a generated range column [0, bound) of the
// step's module, where the first matching row of value v is row v.
static bool range_lookup(const PlanData& d, const Step& st) {
  if (st.params[1] != 1 || st.params[2] != 1 || st.params[3] || st.params[4]) return false;
  const uint32_t ns = st.params[0];
  const Dag& t = st.ins[ns];
  return t.nodes.size() == 1 && t.nodes[0].op == OP_COL && t.nodes[0].b == 0 && d.cols[t.nodes[0].a].src == SRC_RANGE &&
         d.cols[t.nodes[0].a].module == st.module && st.params[5 + ns] == st.module &&
         d.cols[t.nodes[0].a].a0 <= d.modules[st.module].N;
}

void Job::step_multiplicity(size_t r, size_t s) {
  const Step& st = d_.rounds[r].steps[s];
  if (!range_lookup(d_, st)) {
    step_multiplicity_general(r, s);
    return;
  }
  const uint32_t ns = st.params[0];
  const uint32_t bound = d_.cols[st.ins[ns].nodes[0].a].a0;
  const ColBuf& mc = cols_[st.out_col];
  std::vector<uint32_t*> cnt(B_), errp(B_);
  DevBuf err = dev((size_t)B_);
  for (int b = 0; b < B_; ++b) {
    cnt[b] = mc.lag(b, 0);
    errp[b] = err.as<uint32_t>() + b;
  }
  uint32_t* const* cnt_d = const_cast<uint32_t* const*>(stage(cnt));
  launch_fill(cnt_d, (int)mc.n, 0u, B_, s_);
  launch_fill(const_cast<uint32_t* const*>(stage(errp)), 1, 0u, B_, s_);
  std::vector<std::vector<StepDag>> cls;
  for (uint32_t j = 0; j < ns; ++j) {
    const StepDag it{s, j, st.params[5 + j]};
    bool placed = false;
    for (auto& c : cls)
      if (same_dag(r, c[0], it)) {
        c.push_back(it);
        placed = true;
        break;
      }
    if (!placed) cls.push_back({it});
  }
  for (auto& c : cls) {
    Layer src = eval_step_dags(r, c);
    const int NB = (int)(c.size() * B_);
    std::vector<const uint32_t*> sp(NB);
    std::vector<uint32_t*> cp(NB), ep(NB);
    for (int x = 0; x < NB; ++x) {  // (source, instance) pair x counts into instance x % B's table and flag
      sp[x] = src.ptr(x);  // limb 0: the base value
      cp[x] = cnt[x % B_];
      ep[x] = errp[x % B_];
    }
    launch_histogram(stage(sp), (int)d_.modules[c[0].mi].N, const_cast<uint32_t* const*>(stage(cp)), bound,
                     const_cast<uint32_t* const*>(stage(ep)), NB, s_);
    keep_.push_back(std::move(src.buf));
  }
  launch_counts_to_field(cnt_d, (int)mc.n, B_, s_);
  pending_err_.push_back({"lookup: a source value is not in the range table", std::move(err)});
}

// poly.BuildLookupMultiplicities (poly/lookup_multiplicity.go:37-100) for
// any shape: every selector and tuple entry evaluated on the device, then the
// hash-table kernels of launch_lookup_general per instance.
void Job::step_multiplicity_general(size_t r, size_t s) {
  const Step& st = d_.rounds[r].steps[s];
  const uint32_t ns = st.params[0], nt = st.params[1], w = st.params[2];
  const bool hs = st.params[3] != 0, ht = st.params[4] != 0;
  auto mod_of = [&](uint32_t x) { return st.params[5 + x]; };  // x < ns: source, else target x - ns
  // input DAG index of each operand
  const size_t i_sel_s = 0, i_sel_t = hs ? ns : 0, i_src = i_sel_t + (ht ? nt : 0), i_tgt = i_src + (size_t)ns * w;
  std::vector<Layer> val(st.ins.size());
  for (size_t k = 0; k < st.ins.size(); ++k) {
    uint32_t mi;
    if (k < i_sel_t) mi = mod_of((uint32_t)k);
    else if (k < i_src) mi = mod_of(ns + (uint32_t)(k - i_sel_t));
    else if (k < i_tgt) mi = mod_of((uint32_t)((k - i_src) / w));
    else mi = mod_of(ns + (uint32_t)((k - i_tgt) / w));
    val[k] = eval_step_dags(r, {StepDag{s, k, mi}});
  }
  uint32_t total_t = 0, total_s = 0;
  for (uint32_t k = 0; k < nt; ++k) total_t += d_.modules[mod_of(ns + k)].N;
  for (uint32_t j = 0; j < ns; ++j) total_s += d_.modules[mod_of(j)].N;
  auto cap_log_of = [](uint64_t rows) {
    int l = 10;
    while ((1ull << l) < 2 * rows) ++l;
    return (uint32_t)l;
  };
  const uint32_t cap_log = cap_log_of(total_t), miss_log = hs ? cap_log_of(total_s) : 0;
  const size_t cap = (size_t)1 << cap_log, mcap = hs ? (size_t)1 << miss_log : 0;
  DevBuf row = dev(B_ * cap), first = dev(B_ * cap), wsum = dev(B_ * cap * 2);
  DevBuf mrow = hs ? dev(B_ * mcap) : DevBuf(), mw = hs ? dev(B_ * mcap * 2) : DevBuf();
  DevBuf err = dev((size_t)B_);
  std::vector<uint32_t*> fill0, fill1;
  std::vector<LkTable> tabs((size_t)B_ * (nt + ns));
  std::vector<LkJob> jobs(B_);
  // the tuple columns of every table, w per table, in tabs' order
  std::vector<const uint32_t*> colp(tabs.size() * w);
  for (int b = 0; b < B_; ++b) {
    const size_t t0 = (size_t)b * (nt + ns);
    for (uint32_t k = 0; k < nt; ++k)
      for (uint32_t c = 0; c < w; ++c) colp[(t0 + k) * w + c] = val[i_tgt + (size_t)k * w + c].ptr(b);
    for (uint32_t j = 0; j < ns; ++j)
      for (uint32_t c = 0; c < w; ++c) colp[(t0 + nt + j) * w + c] = val[i_src + (size_t)j * w + c].ptr(b);
  }
  const uint32_t* const* colp_d = stage(colp);
  for (int b = 0; b < B_; ++b) {
    LkJob& J = jobs[b];
    J.slot_row = row.as<uint32_t>() + b * cap;
    J.slot_first = first.as<uint32_t>() + b * cap;
    J.slot_w = reinterpret_cast<unsigned long long*>(wsum.as<uint32_t>() + b * cap * 2);
    J.miss_row = hs ? mrow.as<uint32_t>() + b * mcap : nullptr;
    J.miss_w = hs ? reinterpret_cast<unsigned long long*>(mw.as<uint32_t>() + b * mcap * 2) : nullptr;
    J.err = err.as<uint32_t>() + b;
    J.cap_log = cap_log;
    J.miss_cap_log = miss_log;
    uint32_t off = 0;
    for (uint32_t k = 0; k < nt; ++k) {  // targets
      LkTable& t = tabs[(size_t)b * (nt + ns) + k];
      t.col = colp_d + ((size_t)b * (nt + ns) + k) * w;
      t.sel = ht ? val[i_sel_t + k].ptr(b) : nullptr;
      t.mult = cols_[st.outs[k]].lag(b, 0);
      t.n = d_.modules[mod_of(ns + k)].N;
      t.off = off;
      off += t.n;
    }
    off = 0;
    for (uint32_t j = 0; j < ns; ++j) {  // sources
      LkTable& t = tabs[(size_t)b * (nt + ns) + nt + j];
      t.col = colp_d + ((size_t)b * (nt + ns) + nt + j) * w;
      t.sel = hs ? val[i_sel_s + j].ptr(b) : nullptr;
      t.mult = nullptr;
      t.n = d_.modules[mod_of(j)].N;
      t.off = off;
      off += t.n;
    }
    fill0.push_back(J.slot_row);
    fill1.push_back(J.slot_first);
  }
  // cleared here by fill kernels (no memset inside a recorded graph)
  launch_fill(const_cast<uint32_t* const*>(stage(fill0)), (int)cap, 0u, B_, s_);
  launch_fill(const_cast<uint32_t* const*>(stage(fill1)), (int)cap, ~0u, B_, s_);
  {
    std::vector<uint32_t*> z;
    for (int b = 0; b < B_; ++b) z.push_back(wsum.as<uint32_t>() + b * cap * 2);
    launch_fill(const_cast<uint32_t* const*>(stage(z)), (int)(cap * 2), 0u, B_, s_);
    if (hs) {
      std::vector<uint32_t*> zr, zw;
      for (int b = 0; b < B_; ++b) {
        zr.push_back(mrow.as<uint32_t>() + b * mcap);
        zw.push_back(mw.as<uint32_t>() + b * mcap * 2);
      }
      launch_fill(const_cast<uint32_t* const*>(stage(zr)), (int)mcap, 0u, B_, s_);
      launch_fill(const_cast<uint32_t* const*>(stage(zw)), (int)(mcap * 2), 0u, B_, s_);
    }
    std::vector<uint32_t*> ze;
    for (int b = 0; b < B_; ++b) ze.push_back(err.as<uint32_t>() + b);
    launch_fill(const_cast<uint32_t* const*>(stage(ze)), 1, 0u, B_, s_);
  }
  launch_lookup_general(stage(tabs), stage(jobs), (int)nt, (int)ns, (int)w, total_t, total_s, (uint32_t)cap,
                        (uint32_t)mcap, B_, s_);
  for (auto& v : val) keep_.push_back(std::move(v.buf));
  keep_.push_back(std::move(row));
  keep_.push_back(std::move(first));
  keep_.push_back(std::move(wsum));
  if (hs) {
    keep_.push_back(std::move(mrow));
    keep_.push_back(std::move(mw));
  }
  pending_err_.push_back({"lookup: a source tuple is not in any target", std::move(err)});
}

// board.CyclicLogUpStep (board/logup.go:138-186) for a class of steps:
// L_i = sum_{r<=i} v_r - (i+1) T/N with v = sum_j M_j / E_j; the totals T
// come back with the round's roots.
void Job::step_logups(size_t r, const std::vector<size_t>& steps) {
  const size_t NI = steps.size();
  const Step& st0 = d_.rounds[r].steps[steps[0]];
  const ModuleInfo& m = d_.modules[st0.module];
  const uint32_t nt = st0.params[0];
  const int NB = (int)(NI * B_);
  Layer acc;
  acc.len = m.N;
  acc.buf = dev((size_t)NB * 6 * m.N);
  std::vector<uint32_t*> ap(NB), lp(NB);
  for (size_t i = 0; i < NI; ++i)
    for (int b = 0; b < B_; ++b) {
      const int x = (int)(i * B_ + b);
      ap[x] = acc.ptr(x);
      lp[x] = cols_[d_.rounds[r].steps[steps[i]].outs[0]].lag(b, 0);
    }
  uint32_t* const* ap_d = const_cast<uint32_t* const*>(stage(ap));
  for (uint32_t j = 0; j < nt; ++j) {
    std::vector<StepDag> ei, mi;
    for (size_t s : steps) {
      ei.push_back({s, 2 * (size_t)j, d_.rounds[r].steps[s].module});
      mi.push_back({s, 2 * (size_t)j + 1, d_.rounds[r].steps[s].module});
    }
    Layer e = eval_step_dags(r, ei);
    Layer mm = eval_step_dags(r, mi);
    std::vector<const uint32_t*> ep(NB), mp(NB);
    for (int x = 0; x < NB; ++x) {
      ep[x] = e.ptr(x);
      mp[x] = mm.ptr(x);
    }
    launch_logup_acc(stage(ep), stage(mp), ap_d, (int)m.N, j == 0 ? 1 : 0, NB, s_);
    keep_.push_back(std::move(e.buf));
    keep_.push_back(std::move(mm.buf));
  }
  DevBuf tot = dev((size_t)NB * 6);
  uint32_t* scan_tmp = nullptr;
  if (scan_tmp_words((int)m.N, NB)) {
    keep_.push_back(dev(scan_tmp_words((int)m.N, NB)));
    scan_tmp = keep_.back().as<uint32_t>();
  }
  launch_cyclic_logup_finish(ap_d, const_cast<uint32_t* const*>(stage(lp)), tot.as<uint32_t>(), (int)m.N,
                             kb_inv(M(m.N)), NB, scan_tmp, s_);
  keep_.push_back(std::move(acc.buf));
  for (size_t i = 0; i < NI; ++i)  // step i's totals: [b][6] at i B 6
    totals_dev_.push_back({d_.rounds[r].steps[steps[i]].outs[1], tot.as<uint32_t>() + i * B_ * 6});
  keep_.push_back(std::move(tot));
}

// After round r's commitment: lookup errors, then the logup totals of the
// round (exposed values, and the constant column T/N the relations read).
void Job::flush_err_flags() {
  for (auto& e : pending_err_) {
    copy_async(fs_h_ + fs_err_off_ + err_msg_.size() * (size_t)B_, e.second.as<void>(), (size_t)B_ * 4,
               hipMemcpyDeviceToHost);
    err_msg_.push_back(e.first);
    keep_.push_back(std::move(e.second));
  }
  pending_err_.clear();
}

void Job::finish_round_steps(size_t r) {
  flush_err_flags();
  for (auto& kv : totals_dev_) {
    const uint32_t col = kv.first;
    const uint32_t ninv = kb_inv(M(d_.modules[d_.cols[col].module].N));
    launch_scale_into(kv.second, ninv, chal_dev_.as<uint32_t>(), (int)d_.cols.size(), (int)col, B_, s_);
    copy_async(fs_h_ + fs_tot_off_ + (size_t)col * B_ * 6, kv.second, (size_t)B_ * 24, hipMemcpyDeviceToHost);
    total_ptr_[col] = kv.second;
    total_cols_.push_back(col);
  }
  totals_dev_.clear();
  const Round& rd = d_.rounds[r];
  if (rd.fs_hook != FS_NONE && rd.fs_hook != FS_BIND_EXPOSED_VALUES) throw std::runtime_error("unsupported FS hook");
}

// The DAG's challenge operands, gathered from chal_dev_ into [b][k][6].
const uint32_t* Job::chal_operands(const std::vector<uint32_t>& chal_ids) {
  if (chal_ids.empty()) return nullptr;
  keep_.push_back(dev((size_t)B_ * chal_ids.size() * 6));
  launch_chal_compact(chal_dev_.as<uint32_t>(), (int)d_.cols.size(), stage(chal_ids), (int)chal_ids.size(),
                      keep_.back().as<uint32_t>(), B_, s_);
  return keep_.back().as<uint32_t>();
}

// Message segments of a device challenge.
struct FsMsgBuilder {
  std::vector<FsSeg> seg;
  int len = 0;
  void add(const uint32_t* p, uint32_t len_words, uint32_t stride) {
    if (len_words == 0) return;
    seg.push_back(FsSeg{p, len_words, stride});
    len += (int)len_words;
  }
};

void Job::grind_challenge(FsMsgBuilder mb, uint32_t* dig, uint32_t* e6, uint32_t* salts) {
  std::vector<uint32_t*> sp(B_);
  for (int b = 0; b < B_; ++b) sp[b] = salts + b;
  launch_fill(const_cast<uint32_t* const*>(stage(sp)), 1, 0xFFFFFFFFu, B_, s_);
  // about one sweep of 256-salt blocks for the expected 2^bits trials
  const int wg = (int)std::min<uint64_t>(256, std::max<uint64_t>(1, (1ull << grind_bits_) / 256));
  keep_.push_back(dev((size_t)B_ * 3));
  uint32_t* tr = keep_.back().as<uint32_t>();
  launch_fs_grind(stage(mb.seg), (int)mb.seg.size(), mb.len, grind_bits_, salts, tr, wg, B_, s_);
  mb.add(tr, 3, 3);
  launch_fs_challenge(stage(mb.seg), (int)mb.seg.size(), mb.len, dig, e6, 6, B_, s_);
}

// The first challenge's bindings: the common words, then each instance's
// public-input words.
void Job::init_segments(FsMsgBuilder& mb) {
  if (chain_dev_) mb.add(chain_dev_, 8, 8);  // a running transcript's last digest precedes the bindings
  mb.add(stage(fs_init_), (uint32_t)fs_init_.size(), 0);
  if (!pi_words_.empty() && !pi_words_[0].empty())
    mb.add(pi_dev_, (uint32_t)pi_words_[0].size(), (uint32_t)pi_words_[0].size());
}

// Challenge of round r: StringToElements(FSID, name) || previous digest ||
// bindings: the first challenge's init words, the round tree's root
// (prover.go:360-412), then with FS_BIND_EXPOSED_VALUES [0, 6 limbs] per
// logup total in step order (prover.go:386-395, board/logup.go:213-225).
void Job::round_challenge(size_t r) {
  const Round& rd = d_.rounds[r];
  const ChallengeSchedule& sc = *plan_.schedule();
  FsMsgBuilder mb;
  mb.add(stage(sc.prefix[r]), (uint32_t)sc.prefix[r].size(), 0);
  if (r > 0) mb.add(fs_dig_.as<uint32_t>() + (r - 1) * (size_t)B_ * 8, 8, 8);
  else init_segments(mb);
  if (rd.tree != PLAN_NONE) {
    const TreeBuf& tb = trees_[rd.tree];
    mb.add(tb.node_ptr(0) + 8, 8, (uint32_t)(16 * tb.nleaves));
  }
  if (rd.fs_hook == FS_BIND_EXPOSED_VALUES) {
    static const std::vector<uint32_t> zero{0u};
    const uint32_t* z = stage(zero);
    for (auto& st : rd.steps) {
      if (st.kind != STEP_CYCLIC_LOGUP) continue;
      mb.add(z, 1, 0);
      mb.add(total_ptr_.at(st.outs[1]), 6, 6);
    }
  }
  uint32_t* dig = fs_dig_.as<uint32_t>() + r * (size_t)B_ * 8;
  launch_fs_challenge(stage(mb.seg), (int)mb.seg.size(), mb.len, dig,
                      chal_dev_.as<uint32_t>() + (size_t)rd.challenge_col * 6, (int)d_.cols.size() * 6, B_, s_);
  copy_async(fs_h_ + r * (size_t)B_ * 8, dig, (size_t)B_ * 32, hipMemcpyDeviceToHost);
}

// __zeta (prover/prover.go:545-554): previous digest (or the init words when there
// are no rounds) || the roots of the AIR trees in tree order.
void Job::zeta_challenge() {
  const size_t nr = d_.rounds.size();
  const ChallengeSchedule& sc = *plan_.schedule();
  FsMsgBuilder mb;
  mb.add(stage(sc.prefix[nr]), (uint32_t)sc.prefix[nr].size(), 0);
  if (nr > 0) mb.add(fs_dig_.as<uint32_t>() + (nr - 1) * (size_t)B_ * 8, 8, 8);
  else init_segments(mb);
  for (size_t t = 0; t < d_.trees.size(); ++t)
    if (d_.trees[t].kind == 1) mb.add(trees_[t].node_ptr(0) + 8, 8, (uint32_t)(16 * trees_[t].nleaves));
  uint32_t* dig = fs_dig_.as<uint32_t>() + nr * (size_t)B_ * 8;
  launch_fs_challenge(stage(mb.seg), (int)mb.seg.size(), mb.len, dig, nullptr, 0, B_, s_);
  copy_async(fs_h_ + nr * (size_t)B_ * 8, dig, (size_t)B_ * 32, hipMemcpyDeviceToHost);
}

// Host side of the device challenges, once the stream has passed their
// copies: lookup flags, logup totals and roots, then every instance's
// transcript gets each round's and __zeta's bindings and digest.
void Job::replay_device_fs() {
  // A raised flag fails its instance only: the batch runs on in lockstep (the
  // kernels read nothing out of bounds for it) and its proof is not written.
  for (size_t k = 0; k < err_msg_.size(); ++k)
    for (int b = 0; b < B_; ++b)
      if (fs_h_[fs_err_off_ + k * B_ + b] && inst_err_[b].empty()) inst_err_[b] = err_msg_[k];
  for (uint32_t col : total_cols_)
    for (int b = 0; b < B_; ++b)
      for (int l = 0; l < 6; ++l) totals_[b][col].v[l] = fs_h_[fs_tot_off_ + ((size_t)col * B_ + b) * 6 + l];
  for (size_t t = 0; t < d_.trees.size(); ++t) {
    roots_[t].resize(B_);
    for (int b = 0; b < B_; ++b)
      for (int l = 0; l < 8; ++l) roots_[t][b].v[l] = fs_h_[fs_roots_off_ + (t * B_ + b) * 8 + l];
  }
  const size_t nr = d_.rounds.size();
  zeta_.resize(B_);
  parallel_for(B_, [&](int b) {
    for (size_t r = 0; r < nr; ++r) {
      const Round& rd = d_.rounds[r];
      const std::string name = challenge_name(r);
      if (rd.tree != PLAN_NONE) ts_[b].bind(name, roots_[rd.tree][b].v, 8);
      if (rd.fs_hook == FS_BIND_EXPOSED_VALUES)
        for (auto& st : rd.steps) {
          if (st.kind != STEP_CYCLIC_LOGUP) continue;
          uint32_t w[7] = {0};
          for (int l = 0; l < 6; ++l) w[1 + l] = totals_[b][st.outs[1]].v[l];
          ts_[b].bind(name, w, 7);
        }
      ts_[b].check(name, fs_h_ + (r * B_ + b) * 8);
    }
    for (size_t t = 0; t < d_.trees.size(); ++t)
      if (d_.trees[t].kind == 1) ts_[b].bind("__zeta", roots_[t][b].v, 8);
    const uint32_t* zd = fs_h_ + (nr * B_ + b) * 8;
    ts_[b].check("__zeta", zd);
    for (int l = 0; l < 6; ++l) zeta_[b].v[l] = zd[l];
  });
}

// AIR quotient of one module (compute_quotient.go:223-424 and :484-499,
// prover.go:422-440): leaf columns on rho cosets g*w_{rho n}^i <w_n>, the
// generated DAG kernel divides by X^n - 1, per-coset inverse transforms and a
// rho-point combination give the coefficients of Q, each n-chunk is then
// evaluated on H (Lagrange form) and on the rate-4 LDE domain.
void Job::quotient(const std::vector<size_t>& mods) {
  const size_t NM = mods.size();
  const ModuleInfo& m = d_.modules[mods[0]];
  const DagKernel& K0 = *plan_.relation_kernel(mods[0]);
  const uint32_t n = m.N;
  const int logn = m.log_n;
  const int rho = (int)m.n_chunks;
  int rho_log = 0;
  while ((1 << rho_log) < rho) ++rho_log;
  if ((1 << rho_log) != rho) throw std::runtime_error("quotient: chunk count is not a power of two");
  const size_t rn = (size_t)rho * n;
  const int MB = (int)NM * B_;  // the launch batch: (member module, instance) pairs, mb = member * B + b
  const size_t nslot = K0.col_ids.size();
  // 1. coset evaluations of every leaf column of every member
  std::vector<std::vector<DevBuf>> cos(NM);
  std::vector<NttJob> jobs;
  for (size_t mm = 0; mm < NM; ++mm) {
    const DagKernel& K = *plan_.relation_kernel(mods[mm]);
    cos[mm].resize(nslot);
    for (size_t k = 0; k < nslot; ++k) {
      const ColBuf& cb = cols_[K.col_ids[k]];
      cos[mm][k] = dev((size_t)B_ * cb.limbs * rn);
      for (int b = 0; b < B_; ++b)
        for (int l = 0; l < cb.limbs; ++l) {
          NttJob j{};
          j.in = cb.lag(b, l);
          for (int i = 0; i < rho; ++i)
            j.out[i] = cos[mm][k].as<uint32_t>() + ((size_t)b * cb.limbs + l) * rn + (size_t)i * n;
          jobs.push_back(j);
        }
    }
  }
  NttCfg cfg{};
  cfg.kind = NTT_KIND_QUOT_COSET;
  cfg.logn = logn;
  cfg.inverse_first = 1;
  cfg.n_out = rho;
  for (int i = 0; i < rho; ++i) cfg.pre[i] = tb_.quot_pre(logn, i, rho_log);
  {
    TimeIt tm(ctx_, s_, "ntt_quot_coset", jobs.size() * (1.0 + rho) * n * 4, 0);
    ntt(cfg, jobs);
  }
  // 2. numerator / Z_H on the cosets
  DevBuf Q = dev((size_t)MB * 6 * rn);
  {
    std::vector<const uint32_t*> cp;
    std::vector<uint32_t*> op;
    for (size_t mm = 0; mm < NM; ++mm) {
      const DagKernel& K = *plan_.relation_kernel(mods[mm]);
      for (int b = 0; b < B_; ++b) {
        for (size_t k = 0; k < nslot; ++k)
          cp.push_back(cos[mm][k].as<uint32_t>() + (size_t)b * cols_[K.col_ids[k]].limbs * rn);
        op.push_back(Q.as<uint32_t>() + ((size_t)mm * B_ + b) * 6 * rn);
      }
    }
    // (g w_{rho n}^i)^n - 1 = g^n w_rho^i - 1
    std::vector<uint32_t> zh(rho);
    const uint32_t gn = kb_pow(M(MULT_GEN), n);
    for (int i = 0; i < rho; ++i) zh[i] = kb_inv(kb_sub(kb_mul(gn, kb_pow(omega(rho_log), i)), KB_R));
    // challenge operands of every member, [member][b][slot][6]
    const uint32_t* chal = nullptr;
    if (!K0.chal_ids.empty()) {
      keep_.push_back(dev((size_t)MB * K0.chal_ids.size() * 6));
      for (size_t mm = 0; mm < NM; ++mm) {
        const DagKernel& K = *plan_.relation_kernel(mods[mm]);
        launch_chal_compact(chal_dev_.as<uint32_t>(), (int)d_.cols.size(), stage(K.chal_ids), (int)K.chal_ids.size(),
                            keep_.back().as<uint32_t>() + mm * (size_t)B_ * K.chal_ids.size() * 6, B_, s_);
      }
      chal = keep_.back().as<uint32_t>();
    }
    DagLaunch a{};
    a.cols = stage(cp);
    a.chal = chal;
    a.rowscale = stage(zh);
    a.out = const_cast<uint32_t* const*>(stage(op));
    a.col_stride = rn;
    a.out_stride = rn;
    a.logn = logn;
    a.ncos = rho;
    a.batch = MB;
    double leaf_words = 0;
    for (auto c : K0.col_ids) leaf_words += cols_[c].limbs;
    TimeIt tm(ctx_, s_, "dag_quotient", (double)MB * rn * (leaf_words + 6) * 4, 0);
    dag_kernel_launch(K0, a, s_);
  }
  for (auto& cm : cos)
    for (auto& c : cm) keep_.push_back(std::move(c));
  // 3. per-coset inverse transforms with the w_{rho n}^{-a i} / (rho n) twiddle
  DevBuf Y = dev((size_t)MB * 6 * rn);
  jobs.clear();
  for (int mb = 0; mb < MB; ++mb)
    for (int l = 0; l < 6; ++l)
      for (int i = 0; i < rho; ++i) {
        NttJob j{};
        size_t off = ((size_t)mb * 6 + l) * rn + (size_t)i * n;
        j.in = Q.as<uint32_t>() + off;
        j.post = tb_.quot_post(logn, i, rho_log);
        j.out[0] = Y.as<uint32_t>() + off;
        jobs.push_back(j);
      }
  NttCfg c2{};
  c2.kind = NTT_KIND_QUOT_BACK;
  c2.logn = logn;
  c2.inverse_first = 1;
  c2.n_out = 0;
  {
    TimeIt tm(ctx_, s_, "ntt_quot_back", jobs.size() * 2.0 * n * 4, 0);
    ntt(c2, jobs);
  }
  // 4. rho-point combination -> chunk coefficients [mb][t][l][a]
  DevBuf coef = dev((size_t)MB * rho * 6 * n);
  {
    std::vector<uint32_t> wr(rho);
    for (int k = 0; k < rho; ++k) wr[k] = kb_inv(kb_pow(omega(rho_log), k));
    launch_quot_combine(Y.as<uint32_t>(), coef.as<uint32_t>(), tb_.ginv_pow(logn + rho_log), stage(wr), rho, n, MB, s_);
  }
  // 5. chunks: LDE (4 bit-reversed blocks) + Lagrange values on H
  jobs.clear();
  for (size_t mm = 0; mm < NM; ++mm) {
    const size_t mi = mods[mm];
    std::vector<uint32_t> chunk_col(rho, PLAN_NONE);
    for (size_t c = 0; c < d_.cols.size(); ++c)
      if (d_.cols[c].src == SRC_AIR && d_.cols[c].module == mi && d_.cols[c].a0 < (uint32_t)rho) chunk_col[d_.cols[c].a0] = c;
    for (int t = 0; t < rho; ++t) {
      if (chunk_col[t] == PLAN_NONE) throw std::runtime_error("quotient: missing chunk column");
      ColBuf& cb = cols_[chunk_col[t]];
      if (cb.limbs != 6 || !cb.committed) throw std::runtime_error("quotient: chunk column must be a committed E6 column");
      for (int b = 0; b < B_; ++b)
        for (int l = 0; l < 6; ++l) {
          NttJob j{};
          j.in = coef.as<uint32_t>() + (((mm * B_ + b) * rho + t) * 6 + l) * n;
          for (int o = 0; o < 4; ++o) j.out[o] = cb.lde(b, l) + (size_t)o * n;
          j.out[4] = cb.lag(b, l);
          jobs.push_back(j);
        }
      cb.lde_ready = true;
    }
  }
  NttCfg c3{};
  c3.kind = NTT_KIND_CHUNKS;
  c3.logn = logn;
  c3.inverse_first = 0;
  c3.n_out = 5;
  for (int o = 0; o < 4; ++o) {
    c3.pre[o] = tb_.lde_pre(logn, o, RATE_LOG, false);
    c3.bitrev_out[o] = 1;
  }
  c3.pre[4] = nullptr;
  c3.bitrev_out[4] = 0;
  {
    TimeIt tm(ctx_, s_, "ntt_chunks", jobs.size() * 6.0 * n * 4, 0);
    ntt(c3, jobs);
  }
  keep_.push_back(std::move(Q));
  keep_.push_back(std::move(Y));
  keep_.push_back(std::move(coef));
}

void Job::run() {
  double t0 = now_s();
  t0_run_ = t0;
  LG_HIP(hipSetDevice(ctx_.device));
  phase_start();
  alloc_columns();
  upload_inputs();
  trace("upload");
  trees_.resize(d_.trees.size());
  roots_.resize(d_.trees.size());
  totals_.assign(B_, std::vector<E6>(d_.cols.size(), e6_zero()));
  ts_.assign(B_, Transcript(plan_.schedule()));
  inst_err_.assign(B_, std::string());
  if (in_[0].chained) {  // prover.WithTranscript: every instance follows its caller's transcript
    std::vector<uint32_t> pv((size_t)B_ * 8);
    for (int b = 0; b < B_; ++b) {
      if (!in_[b].chained) throw std::runtime_error("prove: running transcripts for some instances of a batch only");
      ts_[b].chain_from(in_[b].fs_prev);
      std::copy(in_[b].fs_prev, in_[b].fs_prev + 8, pv.begin() + (size_t)8 * b);
    }
    chain_dev_ = stage(pv);
  } else {
    for (int b = 0; b < B_; ++b)
      if (in_[b].chained) throw std::runtime_error("prove: running transcripts for some instances of a batch only");
  }
  const size_t nr = d_.rounds.size();
  // the first challenge binds the hash backend id and the module sizes
  // (prover/prover.go:214-220)
  fs_init_ = string_to_elements(TAG_HASH, "poseidon2");
  fs_init_.push_back(M(TAG_SIZE));
  fs_init_.push_back(M((uint32_t)d_.modules.size()));
  for (auto& m : d_.modules) fs_init_.push_back(M(m.N));
  // the setup tree's root, then the public inputs (prover.go:224-234)
  if (d_.n_setup) fs_init_.insert(fs_init_.end(), plan_.setup()->root, plan_.setup()->root + 8);
  {
    const std::string init = nr ? challenge_name(0) : "__zeta";
    for (int b = 0; b < B_; ++b) {
      ts_[b].bind(init, fs_init_);
      if (!pi_words_[b].empty()) ts_[b].bind(init, pi_words_[b]);
    }
  }
  chal_dev_ = dev((size_t)B_ * d_.cols.size() * 6);
  fs_dig_ = dev((nr + 1) * (size_t)B_ * 8);
  fold_pos_ = plan_.schedule()->pos.at("fri_fold_0");
  mark("setup done");
  region_begin();  // the device phase: from the generated columns to the FRI commit phase
  mark("device phase launched");
  {
    // zeroed with a fill kernel (a recorded graph replays it like any launch)
    std::vector<uint32_t*> cp{chal_dev_.as<uint32_t>()};
    launch_fill(const_cast<uint32_t* const*>(stage(cp)), (int)((size_t)B_ * d_.cols.size() * 6), 0u, 1, s_);
  }
  {
    size_t n_mult = 0;
    for (auto& rd : d_.rounds)
      for (auto& st : rd.steps) n_mult += st.kind == STEP_LOOKUP_MULTIPLICITY;
    fs_roots_off_ = (nr + 1) * (size_t)B_ * 8;
    fs_tot_off_ = fs_roots_off_ + d_.trees.size() * (size_t)B_ * 8;
    fs_err_off_ = fs_tot_off_ + d_.cols.size() * (size_t)B_ * 6;
    fs_h_ = pin((fs_err_off_ + (n_mult + 1) * (size_t)B_) * 4);
  }
  gencol();
  trace("gencol");
  // ExecuteSteps (prover.go:360-412)
  for (size_t r = 0; r < nr; ++r) {
    const Round& rd = d_.rounds[r];
    run_steps(r);
    mark("steps enqueued");
    trace(("steps " + std::to_string(r)).c_str());
    if (rd.tree != PLAN_NONE) {
      commit_tree(rd.tree);
      trace(("commit " + std::to_string(r)).c_str());
    }
    finish_round_steps(r);
    mark("round finished");
    round_challenge(r);
    mark("challenge");
    trace(("round " + std::to_string(r)).c_str());
  }
  flush_err_flags();  // the generated columns' flag when there are no rounds
  phase("execute-steps", t0);
  // ComputeAIRQuotients (prover.go:461-561)
  {
    std::vector<std::vector<size_t>> classes;
    for (size_t mi = 0; mi < d_.modules.size(); ++mi) {
      if (!d_.modules[mi].n_chunks) continue;
      const DagKernel& K = *plan_.relation_kernel(mi);
      bool placed = false;
      for (auto& c : classes) {
        const size_t m0 = c[0];
        const DagKernel& K0 = *plan_.relation_kernel(m0);
        if (d_.modules[m0].N == d_.modules[mi].N && d_.modules[m0].n_chunks == d_.modules[mi].n_chunks &&
            K0.code == K.code && K0.col_ids.size() == K.col_ids.size() && K0.chal_ids.size() == K.chal_ids.size()) {
          c.push_back(mi);
          placed = true;
          break;
        }
      }
      if (!placed) classes.push_back({mi});
    }
    for (auto& c : classes) quotient(c);
  }
  trace("quotient");
  for (size_t t = 0; t < d_.trees.size(); ++t) {
    if (d_.trees[t].kind != 1) continue;
    commit_tree(t);
  }
  zeta_challenge();
  mark("zeta");
  trace("air commit + zeta");
  phase("compute-air-quotients", t0);
  open();
  // prover.go:692-716: SkipFRI reports the claimed-values-only opening
  phase(skip_fri_ ? "pcs-claimed-values-only" : "pcs-open", t0);
  if (!marks_.empty()) {
    std::string line = "[loomgpu] lane " + std::to_string(lane_) + " marks (ms):";
    char buf[96];
    snprintf(buf, sizeof buf, " t0 %.6f s;", t0_run_);
    line += buf;
    for (auto& m : marks_) {
      snprintf(buf, sizeof buf, " %s %.3f;", m.first.c_str(), (m.second - t0_run_) * 1e3);
      line += buf;
    }
    fprintf(stderr, "%s\n", line.c_str());
  }
}

// PCS.Open (fri/open.go:98-215) + multi-degree FRI (fri/fri.go:352-736) +
// query openings (open.go:290-416), then the canonical proof encoding.
void Job::open() {
  const size_t nt = d_.trees.size();
  const size_t nm = d_.modules.size();
  const int nq = (int)d_.num_queries;
  if (grind_bits_ && !skip_fri_) fri_salts_ = dev((size_t)FRI_ROUNDS * B_);
  // ---- claimed values; zeta is the device digest of __zeta
  const uint32_t* zdig = fs_dig_.as<uint32_t>() + d_.rounds.size() * (size_t)B_ * 8;
  // Lagrange weights L_i(zeta) depend on zeta and the size only: one vector
  // per distinct module size, shared by the modules of that size.
  std::map<uint32_t, DevBuf> W;  // by module size: [b][6][N]
  std::vector<uint32_t> wsizes, wninv;
  for (size_t mi = 0; mi < nm; ++mi) {
    bool used = false;
    for (auto& t : d_.trees)
      for (auto& g : t.groups) used |= g.module == mi;
    const uint32_t n = d_.modules[mi].N;
    if (!used || std::find(wsizes.begin(), wsizes.end(), n) != wsizes.end()) continue;
    wsizes.push_back(n);
    wninv.push_back(kb_inv(M(n)));
  }
  DevBuf zb = dev((size_t)B_ * 6), cmb = dev(wsizes.size() * (size_t)B_ * 6 + 6);
  launch_zeta_consts(zdig, stage(wsizes), stage(wninv), (int)wsizes.size(), zb.as<uint32_t>(), cmb.as<uint32_t>(), B_,
                     s_);
  for (size_t j = 0; j < wsizes.size(); ++j) {
    const uint32_t n = wsizes[j];
    DevBuf& w = W[n];
    w = dev((size_t)B_ * 6 * n);
    std::vector<uint32_t*> wp(B_);
    for (int b = 0; b < B_; ++b) wp[b] = w.as<uint32_t>() + (size_t)b * 6 * n;
    launch_lagrange_weights(zb.as<uint32_t>(), cmb.as<uint32_t>() + j * (size_t)B_ * 6, tb_.wpow(ilog2(n)),
                            const_cast<uint32_t* const*>(stage(wp)), n, B_, s_);
  }
  keep_.push_back(std::move(zb));
  keep_.push_back(std::move(cmb));
  // value index per (tree, group, rail, poly, shift), in canonical order
  struct CV {
    size_t t, g;
    int ext;
    size_t poly;
    uint32_t col, shift;
  };
  std::vector<CV> cvs;
  for (size_t t = 0; t < nt; ++t)
    for (size_t g = 0; g < d_.trees[t].groups.size(); ++g) {
      const Group& gr = d_.trees[t].groups[g];
      for (size_t i = 0; i < gr.base.size(); ++i)
        for (auto s : gr.base_shifts[i]) cvs.push_back({t, g, 0, i, gr.base[i], s});
      for (size_t i = 0; i < gr.ext.size(); ++i)
        for (auto s : gr.ext_shifts[i]) cvs.push_back({t, g, 1, i, gr.ext[i], s});
    }
  const size_t ncv = cvs.size();
  DevBuf cvd = dev((size_t)B_ * ncv * 6);
  // One claim set per (instance, module): its claims share the module's
  // Lagrange weights.  One launch per module size.
  struct SizeClaims {
    std::vector<ClaimSet> sets;
    std::vector<ClaimItem> items;
    double bytes = 0;  // algorithmic: W once per set, each distinct polynomial once
  };
  // one set per (instance, module size): the claims of every module of that
  // size share its weights
  std::map<uint32_t, SizeClaims> cl;
  for (int b = 0; b < B_; ++b)
    for (auto& wn : W) {
      const uint32_t n = wn.first;
      SizeClaims& sc = cl[n];
      ClaimSet set{wn.second.as<uint32_t>() + (size_t)b * 6 * n, (int32_t)sc.items.size(), 0};
      std::set<uint32_t> polys;
      for (size_t k = 0; k < ncv; ++k) {
        const CV& v = cvs[k];
        if (d_.modules[d_.cols[v.col].module].N != n) continue;
        ClaimItem it;
        it.poly = cols_[v.col].lag(b, 0);
        it.out = cvd.as<uint32_t>() + ((size_t)b * ncv + k) * 6;
        it.shift = (int32_t)v.shift;
        it.ext = v.ext;
        sc.items.push_back(it);
        if (polys.insert(v.col).second) sc.bytes += (double)n * 4 * (v.ext ? 6 : 1);
      }
      set.count = (int32_t)(sc.items.size() - set.first);
      if (!set.count) continue;
      sc.sets.push_back(set);
      sc.bytes += (double)n * 24 + set.count * 24.0;
    }
  for (auto& kv : cl) {
    const int S = claims_split((int)kv.first, tune_.claims_rows);
    uint32_t* part = nullptr;
    if (S > 1) {
      keep_.push_back(dev(kv.second.items.size() * (size_t)S * 6));
      part = keep_.back().as<uint32_t>();
    }
    TimeIt tm(ctx_, s_, "claims", kv.second.bytes, 0);
    launch_claims(stage(kv.second.sets), (int)kv.second.sets.size(), stage(kv.second.items),
                  (int)kv.second.items.size(), kv.first, part, s_, tune_.claims_rows);
  }
  if (skip_fri_) {
    // prover.SkipFRI (prover/prover.go:626-653, 691-700): the claimed values
    // and nothing of the opening after them
    std::vector<E6> cv((size_t)B_ * ncv);
    region_end();
    LG_HIP(hipMemcpyAsync(cv.data(), cvd.as<void>(), cv.size() * sizeof(E6), hipMemcpyDeviceToHost, s_));
    sync();
    replay_device_fs();
    keep_.push_back(std::move(cvd));
    for (auto& x : W) keep_.push_back(std::move(x.second));
    parallel_for(B_, [&](int b) {
      if (!inst_err_[b].empty()) {
        out_[b].error = inst_err_[b];
        out_[b].proof_len = 0;
        return;
      }
      auto emit = [&](auto& w) {
        emit_head(w, b, cv.data() + (size_t)b * ncv);
        // levels root (zero), no levels, no roots, an empty extension-field
        // final polynomial, no queries, no proofs of work, no samplings
        const uint32_t tail[] = {0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 0u, 1u, 0u, 0u, 0u, 0u};
        w.insert(w.end(), tail, tail + 15);
      };
      if (in_[b].proof_dst) {
        SpanWriter w{in_[b].proof_dst, in_[b].proof_cap};
        emit(w);
        out_[b].proof_len = w.n;
        out_[b].direct = true;
      } else {
        emit(out_[b].proof);
        out_[b].proof_len = out_[b].proof.size();
      }
      out_[b].transcript = ts_[b].log_words();
    });
    return;
  }
  // ---- alpha_DEEP: bind in class (module) order, then tree, group, base, ext
  std::vector<int32_t> bind_order;  // claimed-value indexes in binding order
  for (size_t mi = 0; mi < nm; ++mi)
    for (size_t k = 0; k < ncv; ++k)
      if (d_.trees[cvs[k].t].groups[cvs[k].g].module == mi) bind_order.push_back((int32_t)k);
  const std::vector<size_t> classes = deep_classes(d_);
  std::vector<size_t> lv_order(classes.size());
  for (size_t l = 0; l < lv_order.size(); ++l) lv_order[l] = l;
  std::stable_sort(lv_order.begin(), lv_order.end(), [&](size_t a, size_t b) {
    return d_.modules[classes[a]].N > d_.modules[classes[b]].N;
  });
  const size_t L = classes.size();
  // With one DEEP class (every module the same size), alpha_DEEP and the DEEP
  // constants are derived on the device (k_deep_alpha) and the host replays
  // the claimed values and the digest into its transcripts while the DEEP
  // codeword and the levels tree are computed; otherwise the host derives
  // alpha_DEEP from the claimed values it copies back.
  const bool deep_on_device = L == 1;
  std::vector<E6> cvals((size_t)B_ * ncv);
  std::vector<E6> alpha(B_);
  if (!deep_on_device) {
    LG_HIP(hipMemcpyAsync(cvals.data(), cvd.as<void>(), cvals.size() * sizeof(E6), hipMemcpyDeviceToHost, s_));
    sync();
    replay_device_fs();
    parallel_for(B_, [&](int b) {
      for (int32_t k : bind_order) ts_[b].bind("alpha_DEEP", cvals[(size_t)b * ncv + k].v, 6);
      const Digest& dg = ts_[b].compute("alpha_DEEP");
      for (int l = 0; l < 6; ++l) alpha[b].v[l] = dg.v[l];
    });
  }
  trace("claimed values + alpha");
  uint32_t* alpha_h = nullptr;  // pinned: [b][8] digests, then the claimed values
  // ---- DEEP quotient per class, RS-encoded (deep.go:71-266).  With several
  // classes, classes of one size and bundle structure run as one launch of
  // (class, instance) pairs (below the loop).
  std::vector<Layer> level(L);
  struct DeepPending {
    size_t ci;
    std::vector<DeepBundle> bundles;
    std::vector<int32_t> ptl;
    std::vector<uint32_t> pt_shift, bcols;
    std::vector<E6> scale, consts;  // [b][bundle], [b][point]
  };
  std::vector<DeepPending> pending;
  for (size_t ci = 0; ci < L; ++ci) {
    const size_t mi = classes[ci];
    const ModuleInfo& m = d_.modules[mi];
    std::vector<DeepBundle> bundles;
    std::vector<int32_t> ptl;
    std::vector<uint32_t> pt_shift;  // distinct shifts = distinct points
    std::vector<std::vector<std::pair<size_t, int>>> bundle_vals;  // (cv index, point)
    std::vector<uint32_t> bcols;
    for (size_t t = 0; t < nt; ++t)
      for (size_t g = 0; g < d_.trees[t].groups.size(); ++g) {
        const Group& gr = d_.trees[t].groups[g];
        if (gr.module != mi) continue;
        for (int rail = 0; rail < 2; ++rail) {
          const auto& polys = rail ? gr.ext : gr.base;
          for (size_t i = 0; i < polys.size(); ++i) {
            DeepBundle bd;
            bd.ext = rail;
            bd.col = (int32_t)bcols.size();
            bcols.push_back(polys[i]);
            bd.first_pt = (int32_t)ptl.size();
            std::vector<std::pair<size_t, int>> vals;
            for (size_t k = 0; k < ncv; ++k) {
              const CV& v = cvs[k];
              if (v.t != t || v.g != g || v.ext != rail || v.poly != i) continue;
              auto it = std::find(pt_shift.begin(), pt_shift.end(), v.shift);
              int p = (int)(it - pt_shift.begin());
              if (it == pt_shift.end()) pt_shift.push_back(v.shift);
              ptl.push_back(p);
              vals.push_back({k, p});
            }
            bd.n_pts = (int32_t)vals.size();
            bundles.push_back(bd);
            bundle_vals.push_back(vals);
          }
        }
      }
    const int np = (int)pt_shift.size();
    std::vector<E6> scale((size_t)B_ * bundles.size()), consts((size_t)B_ * np);
    std::vector<const uint32_t*> cp;
    const uint32_t* scale_d = nullptr;
    const uint32_t* consts_d = nullptr;
    for (int b = 0; b < B_; ++b)
      for (auto c : bcols) cp.push_back(cols_[c].lag(b, 0));
    // points zeta * omega^shift, on the device
    std::vector<uint32_t> wsh(np);
    for (int p = 0; p < np; ++p) wsh[p] = kb_pow(omega(m.log_n), pt_shift[p]);
    DevBuf pts = dev((size_t)B_ * np * 6 + 6);
    launch_zeta_points(zdig, stage(wsh), np, pts.as<uint32_t>(), B_, s_);
    if (deep_on_device) {
      std::vector<int3> terms;
      for (size_t i = 0; i < bundles.size(); ++i)
        for (auto& kv : bundle_vals[i]) terms.push_back(make_int3((int)i, (int)kv.first, kv.second));
      DevBuf dig = dev((size_t)B_ * 8);
      DevBuf sc = dev((size_t)B_ * bundles.size() * 6 + 6);
      DevBuf cs = dev((size_t)B_ * np * 6 + 6);
      DeepAlphaArgs da{};
      da.prev = zdig;  // alpha_DEEP follows __zeta
      da.cvals = cvd.as<uint32_t>();
      da.order = stage(bind_order);
      da.ncv = (int)ncv;
      da.terms = stage(terms);
      da.n_terms = (int)terms.size();
      da.n_bundles = (int)bundles.size();
      da.n_points = np;
      da.dig = dig.as<uint32_t>();
      da.scale = sc.as<uint32_t>();
      da.consts = cs.as<uint32_t>();
      const std::vector<uint32_t> pre = string_to_elements(TAG_FSID, "alpha_DEEP");
      launch_deep_alpha(da, pre.data(), (int)pre.size(), B_, s_);
      // The fold rounds before the level enters bind the zero digest
      // (fri.go:630-640): their digests are chained on the device too, so
      // nothing before the query phase waits for the host.
      const int first_j = FRI_LOG_D - m.log_n;
      pre_fold_.dev = dev((size_t)std::max(1, first_j) * B_ * 8);
      {
        static const std::vector<uint32_t> zero8(8, 0u);
        const uint32_t* z = stage(zero8);
        const uint32_t* prev = dig.as<uint32_t>();
        for (int j = 0; j < first_j; ++j) {
          const std::vector<uint32_t>& pf = plan_.schedule()->prefix[fold_pos_ + j];
          FsMsgBuilder mb;
          mb.add(stage(pf), (uint32_t)pf.size(), 0);
          mb.add(prev, 8, 8);
          mb.add(z, 8, 0);
          uint32_t* out = pre_fold_.dev.as<uint32_t>() + (size_t)j * B_ * 8;
          if (grind_bits_) grind_challenge(mb, out, nullptr, fri_salts_.as<uint32_t>() + (size_t)j * B_);
          else launch_fs_challenge(stage(mb.seg), (int)mb.seg.size(), mb.len, out, nullptr, 0, B_, s_);
          prev = out;
        }
        pre_fold_.prev = prev;
        pre_fold_.n = first_j;
      }
      alpha_h = pin(((size_t)B_ * 8 + (size_t)B_ * ncv * 6 + (size_t)std::max(1, first_j) * B_ * 8) * 4);
      copy_async(alpha_h, dig.as<void>(), (size_t)B_ * 8 * 4, hipMemcpyDeviceToHost);
      copy_async(alpha_h + (size_t)B_ * 8, cvd.as<void>(), (size_t)B_ * ncv * 6 * 4, hipMemcpyDeviceToHost);
      copy_async(alpha_h + (size_t)B_ * 8 + (size_t)B_ * ncv * 6, pre_fold_.dev.as<void>(),
                 (size_t)first_j * B_ * 8 * 4, hipMemcpyDeviceToHost);
      scale_d = sc.as<uint32_t>();
      consts_d = cs.as<uint32_t>();
      keep_.push_back(std::move(dig));
      keep_.push_back(std::move(sc));
      keep_.push_back(std::move(cs));
    } else {
      for (int b = 0; b < B_; ++b) {
        E6 a = e6_one();
        for (int p = 0; p < np; ++p) consts[(size_t)b * np + p] = e6_zero();
        for (size_t i = 0; i < bundles.size(); ++i) {
          scale[(size_t)b * bundles.size() + i] = a;
          for (auto& kv : bundle_vals[i]) {
            E6& c = consts[(size_t)b * np + kv.second];
            c = e6_add(c, e6_mul(a, cvals[(size_t)b * ncv + kv.first]));
          }
          a = e6_mul(a, alpha[b]);
        }
      }
      pending.push_back(DeepPending{ci, bundles, ptl, pt_shift, bcols, std::move(scale), std::move(consts)});
      keep_.push_back(std::move(pts));
      continue;
    }
    Layer dq;
    dq.len = m.N;
    dq.buf = dev((size_t)B_ * 6 * m.N);
    std::vector<uint32_t*> op(B_);
    for (int b = 0; b < B_; ++b) op[b] = dq.ptr(b);
    DeepDesc dd{};
    dd.n = m.N;
    dd.n_bundles = (int)bundles.size();
    dd.n_points = np;
    dd.bundles = stage(bundles);
    dd.pt_list = stage(ptl);
    dd.cols = stage(cp);
    dd.n_cols = (int)bcols.size();
    dd.scale = scale_d;
    dd.consts = consts_d;
    dd.points = pts.as<uint32_t>();
    dd.wpow = tb_.wpow(m.log_n);
    dd.out = const_cast<uint32_t* const*>(stage(op));
    {
      TimeIt tm(ctx_, s_, "deep", (double)B_ * m.N * (bcols.size() + 6) * 4, 0);
      launch_deep(dd, B_, s_);
    }
    level[ci].len = 4 * m.N;
    level[ci].buf = dev((size_t)B_ * 6 * 4 * m.N);
    std::vector<NttJob> jobs;
    for (int b = 0; b < B_; ++b)
      for (int l = 0; l < 6; ++l) {
        NttJob j{};
        j.in = dq.ptr(b) + (size_t)l * m.N;
        for (int o = 0; o < 4; ++o) j.out[o] = level[ci].ptr(b) + (size_t)l * 4 * m.N + (size_t)o * m.N;
        jobs.push_back(j);
      }
    NttCfg cfg{};
    cfg.kind = NTT_KIND_DEEP_LDE;
    cfg.logn = m.log_n;
    cfg.inverse_first = 1;
    cfg.n_out = 4;
    for (int o = 0; o < 4; ++o) {
      cfg.pre[o] = tb_.lde_pre(m.log_n, o, RATE_LOG, true);
      cfg.bitrev_out[o] = 1;
    }
    {
      TimeIt tm(ctx_, s_, "ntt_deep_lde", jobs.size() * 5.0 * m.N * 4, 0);
      ntt(cfg, jobs);
    }
    keep_.push_back(std::move(dq.buf));
    keep_.push_back(std::move(pts));
  }
  // the deferred classes, grouped by size and structure
  {
    auto same = [&](const DeepPending& a, const DeepPending& b) {
      if (d_.modules[classes[a.ci]].N != d_.modules[classes[b.ci]].N || a.pt_shift != b.pt_shift ||
          a.ptl != b.ptl || a.bundles.size() != b.bundles.size() || a.bcols.size() != b.bcols.size())
        return false;
      for (size_t i = 0; i < a.bundles.size(); ++i)
        if (a.bundles[i].ext != b.bundles[i].ext || a.bundles[i].col != b.bundles[i].col ||
            a.bundles[i].first_pt != b.bundles[i].first_pt || a.bundles[i].n_pts != b.bundles[i].n_pts)
          return false;
      return true;
    };
    std::vector<char> done(pending.size(), 0);
    for (size_t g0 = 0; g0 < pending.size(); ++g0) {
      if (done[g0]) continue;
      std::vector<size_t> grp;
      for (size_t k = g0; k < pending.size(); ++k)
        if (!done[k] && same(pending[g0], pending[k])) {
          grp.push_back(k);
          done[k] = 1;
        }
      const DeepPending& p0 = pending[grp[0]];
      const ModuleInfo& m = d_.modules[classes[p0.ci]];
      const int np = (int)p0.pt_shift.size();
      const int MB = (int)grp.size() * B_;  // mb = member * B + b
      std::vector<uint32_t> wsh(np);
      for (int p = 0; p < np; ++p) wsh[p] = kb_pow(omega(m.log_n), p0.pt_shift[p]);
      DevBuf pts = dev((size_t)B_ * np * 6 + 6);  // shared by the members: indexed by instance
      launch_zeta_points(zdig, stage(wsh), np, pts.as<uint32_t>(), B_, s_);
      std::vector<const uint32_t*> cp;
      std::vector<E6> scale, consts;
      for (size_t k : grp) {
        const DeepPending& pk = pending[k];
        for (int b = 0; b < B_; ++b)
          for (auto c : pk.bcols) cp.push_back(cols_[c].lag(b, 0));
        scale.insert(scale.end(), pk.scale.begin(), pk.scale.end());
        consts.insert(consts.end(), pk.consts.begin(), pk.consts.end());
      }
      Layer dq;
      dq.len = m.N;
      dq.buf = dev((size_t)MB * 6 * m.N);
      std::vector<uint32_t*> op(MB);
      for (int mb = 0; mb < MB; ++mb) op[mb] = dq.ptr(mb);
      DeepDesc dd{};
      dd.n = m.N;
      dd.n_bundles = (int)p0.bundles.size();
      dd.n_points = np;
      dd.bundles = stage(p0.bundles);
      dd.pt_list = stage(p0.ptl);
      dd.cols = stage(cp);
      dd.n_cols = (int)p0.bcols.size();
      dd.scale = reinterpret_cast<const uint32_t*>(stg().put(scale.data(), scale.size() * sizeof(E6), s_));
      dd.consts = reinterpret_cast<const uint32_t*>(stg().put(consts.data(), consts.size() * sizeof(E6), s_));
      dd.points = pts.as<uint32_t>();
      dd.points_mod = B_;
      dd.wpow = tb_.wpow(m.log_n);
      dd.out = const_cast<uint32_t* const*>(stage(op));
      {
        TimeIt tm(ctx_, s_, "deep", (double)MB * m.N * (p0.bcols.size() + 6) * 4, 0);
        launch_deep(dd, MB, s_);
      }
      std::vector<NttJob> jobs;
      for (size_t mm = 0; mm < grp.size(); ++mm) {
        Layer& lv = level[pending[grp[mm]].ci];
        lv.len = 4 * m.N;
        lv.buf = dev((size_t)B_ * 6 * 4 * m.N);
        for (int b = 0; b < B_; ++b)
          for (int l = 0; l < 6; ++l) {
            NttJob j{};
            j.in = dq.ptr((int)mm * B_ + b) + (size_t)l * m.N;
            for (int o = 0; o < 4; ++o) j.out[o] = lv.ptr(b) + (size_t)l * 4 * m.N + (size_t)o * m.N;
            jobs.push_back(j);
          }
      }
      NttCfg cfg{};
      cfg.kind = NTT_KIND_DEEP_LDE;
      cfg.logn = m.log_n;
      cfg.inverse_first = 1;
      cfg.n_out = 4;
      for (int o = 0; o < 4; ++o) {
        cfg.pre[o] = tb_.lde_pre(m.log_n, o, RATE_LOG, true);
        cfg.bitrev_out[o] = 1;
      }
      {
        TimeIt tm(ctx_, s_, "ntt_deep_lde", jobs.size() * 5.0 * m.N * 4, 0);
        ntt(cfg, jobs);
      }
      keep_.push_back(std::move(dq.buf));
      keep_.push_back(std::move(pts));
    }
  }
  trace("alpha + deep");
  // ---- FRI (fri.go:606-736)
  auto ext_tree = [&](TreeBuf& tb, const std::vector<const Layer*>& lv) {
    std::vector<std::vector<const uint32_t*>> gcols;
    std::vector<std::pair<int, int>> shape;
    std::vector<int> rows;
    for (const Layer* x : lv) {
      for (int b = 0; b < B_; ++b) {
        std::vector<const uint32_t*> cp;
        for (int l = 0; l < 6; ++l) cp.push_back(x->ptr(b) + (size_t)l * x->len);
        gcols.push_back(cp);
      }
      shape.push_back({0, 1});
      rows.push_back((int)x->len);
    }
    build_tree(tb, gcols, shape, rows);
  };
  std::vector<const Layer*> lvp;
  for (size_t l : lv_order) lvp.push_back(&level[l]);
  const int first = FRI_LOG_D - d_.modules[classes[lv_order[0]]].log_n;
  TreeBuf levels_tree;
  ext_tree(levels_tree, lvp);
  // With one DEEP class the host replays the claimed values, alpha_DEEP and
  // the fold rounds before the level after the FRI commit phase (below):
  // until then nothing the device does waits for the host.
  const uint32_t* salts_h = nullptr;  // [fold round][b] proof-of-work salts, once the device phase is done
  auto replay_alpha = [&] {
    replay_device_fs();
    const uint32_t* cv_h = alpha_h + (size_t)B_ * 8;
    const uint32_t* pre_h = cv_h + (size_t)B_ * ncv * 6;
    std::memcpy(cvals.data(), cv_h, cvals.size() * sizeof(E6));
    parallel_for(B_, [&](int b) {
      for (int32_t k : bind_order) ts_[b].bind("alpha_DEEP", cvals[(size_t)b * ncv + k].v, 6);
      ts_[b].check("alpha_DEEP", alpha_h + (size_t)8 * b);
      static const uint32_t zero[8] = {0};
      for (int j = 0; j < pre_fold_.n; ++j) {
        ts_[b].bind_at(fold_pos_ + j, zero, 8);
        if (grind_bits_) ts_[b].set_pow(fold_pos_ + j, grind_bits_, salts_h[(size_t)j * B_ + b]);
        ts_[b].check_at(fold_pos_ + j, pre_h + ((size_t)j * B_ + b) * 8);
      }
    });
  };
  keep_.push_back(std::move(cvd));
  // With one FRI level nothing on the host needs the levels root before the
  // fold rounds: round `first` then derives its challenge on the device like
  // the later rounds, and the root comes back with theirs.
  const bool first_on_host = L > 1;
  std::vector<Digest> levels_root(B_);
  if (first_on_host) levels_root = roots_of(levels_tree);
  std::vector<std::vector<E6>> gamma(L, std::vector<E6>(B_));
  if (first_on_host)
    parallel_for(B_, [&](int b) {
      for (size_t l = 1; l < L; ++l) {
        const std::string nmg = "fri_level_" + std::to_string(l) + "_gamma";
        if (l == 1) ts_[b].bind(nmg, levels_root[b].v, 8);
        const Digest& dg = ts_[b].compute(nmg);
        for (int k = 0; k < 6; ++k) gamma[l][b].v[k] = dg.v[k];
      }
      if (grind_bits_) return;  // the fold challenges carry proofs of work: found on the device
      for (int j = 0; j < first; ++j) {
        const std::string nf = "fri_fold_" + std::to_string(j);
        uint32_t zero[8] = {0};
        ts_[b].bind(nf, zero, 8);
        ts_[b].compute(nf);
      }
    });
  std::vector<Layer> layers(FRI_ROUNDS + 1);
  std::vector<TreeBuf> ftrees(FRI_ROUNDS);
  std::vector<std::vector<Digest>> fri_roots;
  std::vector<uint32_t> fri_dig;  // device fold-challenge digests, [round][b][8]
  // Every layer, tree and pointer table of the commit phase is allocated and
  // staged before its first launch (one staging copy for all rounds).  Round
  // j (> first) builds the tree of layer j and derives its fold challenge on
  // the device (k_fri_challenge: no host round trip); the fold of round j
  // also hashes the leaves of layer j+1's tree (k_fri_fold_leaf) unless a
  // level joins layer j+1 first.  The host replays roots and digests into the
  // transcripts after the last round.
  std::vector<char> joins(FRI_ROUNDS + 1, 0);
  for (size_t l = 1; l < L; ++l) joins[FRI_LOG_D - d_.modules[classes[lv_order[l]]].log_n] = 1;
  layers[first].len = lvp[0]->len;
  layers[first].buf = dev((size_t)B_ * 6 * layers[first].len);
  copy_async(layers[first].buf.as<void>(), lvp[0]->buf.as<void>(), (size_t)B_ * 6 * layers[first].len * 4,
             hipMemcpyDeviceToDevice);
  for (int j = first + 1; j <= FRI_ROUNDS; ++j) {
    layers[j].len = layers[j - 1].len / 2;
    layers[j].buf = dev((size_t)B_ * 6 * layers[j].len);
  }
  auto ptrs = [&](const Layer& x) {
    std::vector<uint32_t*> v(B_);
    for (int b = 0; b < B_; ++b) v[b] = const_cast<uint32_t*>(x.ptr(b));
    return const_cast<uint32_t* const*>(stage(v));
  };
  std::vector<uint32_t* const*> lay_d(FRI_ROUNDS + 1), tree_d(FRI_ROUNDS), leaf_d(FRI_ROUNDS);
  std::vector<const uint32_t* const*> limb_d(FRI_ROUNDS);  // 6 limb columns per instance, for leaf_hash
  for (int j = first; j <= FRI_ROUNDS; ++j) lay_d[j] = ptrs(layers[j]);
  for (int j = first + 1; j < FRI_ROUNDS; ++j) {
    TreeBuf& tb = ftrees[j];
    tb.order = {0};
    tb.nleaves = (int)(layers[j].len / 2);
    tb.nodes = dev((size_t)B_ * 2 * tb.nleaves * 8);
    std::vector<uint32_t*> tp(B_), lp(B_);
    std::vector<const uint32_t*> cp;
    for (int b = 0; b < B_; ++b) {
      tp[b] = tb.node_ptr(b);
      lp[b] = tb.node_ptr(b) + (size_t)8 * tb.nleaves;
      for (int l = 0; l < 6; ++l) cp.push_back(layers[j].ptr(b) + (size_t)l * layers[j].len);
    }
    tree_d[j] = const_cast<uint32_t* const*>(stage(tp));
    leaf_d[j] = const_cast<uint32_t* const*>(stage(lp));
    limb_d[j] = stage(cp);
  }
  std::vector<int32_t> ext_msg{0, 0, 0};
  for (int i = 0; i < 6; ++i) ext_msg.push_back(2 * i);
  for (int i = 0; i < 6; ++i) ext_msg.push_back(2 * i + 1);
  const int32_t* ext_msg_d = stage(ext_msg);
  std::vector<const uint32_t*> level_lp(L);
  std::vector<const uint32_t* const*> level_d(L);
  std::vector<const uint32_t*> gamma_d(L);
  for (size_t l = 1; l < L; ++l) {
    std::vector<const uint32_t*> lp(B_);
    for (int b = 0; b < B_; ++b) lp[b] = lvp[l]->ptr(b);
    level_d[l] = stage(lp);
    gamma_d[l] = reinterpret_cast<const uint32_t*>(stg().put(gamma[l].data(), gamma[l].size() * sizeof(E6), s_));
  }
  const size_t fb = (size_t)B_ * 8;
  DevBuf chal_dig = dev((size_t)FRI_ROUNDS * fb);
  DevBuf chal_alpha = dev((size_t)FRI_ROUNDS * B_ * 6);
  DevBuf root_dev = dev((size_t)FRI_ROUNDS * fb);
  // round `first` binds the levels root: with several levels the host has it
  // (the gammas bound it) and computes the challenge here
  const uint32_t* prev_dig = nullptr;
  const uint32_t* alpha_first = nullptr;
  const std::string first_name = "fri_fold_" + std::to_string(first);
  // a fold challenge with a proof of work: the message prefix, the previous
  // digest and the bound root (node 1 of each instance's node array)
  auto grind_fold = [&](int j, const uint32_t* prev, const TreeBuf& tb) {
    const std::vector<uint32_t>& pf = plan_.schedule()->prefix[fold_pos_ + j];
    FsMsgBuilder mb;
    mb.add(stage(pf), (uint32_t)pf.size(), 0);
    mb.add(prev, 8, 8);
    mb.add(tb.node_ptr(0) + 8, 8, (uint32_t)(16 * tb.nleaves));
    grind_challenge(mb, chal_dig.as<uint32_t>() + (size_t)j * fb, chal_alpha.as<uint32_t>() + (size_t)j * B_ * 6,
                    fri_salts_.as<uint32_t>() + (size_t)j * B_);
  };
  if (!first_on_host) {
    // the previous digest is the device's (alpha_DEEP or the last zero-bound
    // fold round, chained above)
    std::vector<uint32_t*> tp(B_);
    for (int b = 0; b < B_; ++b) tp[b] = levels_tree.node_ptr(b);
    uint32_t* dig_f = chal_dig.as<uint32_t>() + (size_t)first * fb;
    uint32_t* al_f = chal_alpha.as<uint32_t>() + (size_t)first * B_ * 6;
    if (grind_bits_) {
      launch_copy_roots(const_cast<uint32_t* const*>(stage(tp)), root_dev.as<uint32_t>() + (size_t)first * fb, B_, s_);
      grind_fold(first, pre_fold_.prev, levels_tree);
    } else {
      const std::vector<uint32_t> pre = string_to_elements(TAG_FSID, first_name);
      launch_fri_challenge(pre.data(), (int)pre.size(), pre_fold_.prev, const_cast<uint32_t* const*>(stage(tp)),
                           dig_f, al_f, root_dev.as<uint32_t>() + (size_t)first * fb, B_, s_);
    }
    prev_dig = dig_f;
    alpha_first = al_f;
  } else if (grind_bits_) {
    // the zero-bound rounds and round `first` chain from the last level
    // gamma's digest (the host's), each with its proof of work
    std::vector<uint32_t> pv((size_t)B_ * 8);
    for (int b = 0; b < B_; ++b) {
      const Digest& d = ts_[b].previous_value("fri_fold_0");
      std::copy(d.v, d.v + 8, pv.begin() + (size_t)8 * b);
    }
    const uint32_t* prev = stage(pv);
    static const std::vector<uint32_t> zero8(8, 0u);
    const uint32_t* z = stage(zero8);
    for (int j = 0; j < first; ++j) {
      const std::vector<uint32_t>& pf = plan_.schedule()->prefix[fold_pos_ + j];
      FsMsgBuilder mb;
      mb.add(stage(pf), (uint32_t)pf.size(), 0);
      mb.add(prev, 8, 8);
      mb.add(z, 8, 0);
      grind_challenge(mb, chal_dig.as<uint32_t>() + (size_t)j * fb, nullptr, fri_salts_.as<uint32_t>() + (size_t)j * B_);
      prev = chal_dig.as<uint32_t>() + (size_t)j * fb;
    }
    grind_fold(first, prev, levels_tree);
    prev_dig = chal_dig.as<uint32_t>() + (size_t)first * fb;
    alpha_first = chal_alpha.as<uint32_t>() + (size_t)first * B_ * 6;
  } else {
    std::vector<uint32_t> dg((size_t)B_ * 8), al((size_t)B_ * 6);
    parallel_for(B_, [&](int b) {
      const std::string nf = "fri_fold_" + std::to_string(first);
      ts_[b].bind(nf, levels_root[b].v, 8);
      const Digest& d = ts_[b].compute(nf);
      for (int k = 0; k < 8; ++k) dg[(size_t)8 * b + k] = d.v[k];
      for (int k = 0; k < 6; ++k) al[(size_t)6 * b + k] = d.v[k];
    });
    prev_dig = stage(dg);
    alpha_first = stage(al);
  }
  const uint32_t hdr_ext = M(2);  // 2 ne for one E6 column
  std::vector<char> leaves_done(FRI_ROUNDS, 0);
  for (int j = first; j < FRI_ROUNDS; ++j) {
    for (size_t l = 1; l < L; ++l) {
      const int jl = FRI_LOG_D - d_.modules[classes[lv_order[l]]].log_n;
      if (jl == j) launch_fri_combine(lay_d[j], level_d[l], gamma_d[l], (int)layers[j].len, B_, s_);
    }
    const uint32_t* alpha_j = alpha_first;
    if (j > first) {
      TreeBuf& tb = ftrees[j];
      if (!leaves_done[j]) {
        TimeIt tm(ctx_, s_, "leaf_hash", (double)B_ * tb.nleaves * (2.0 * 6 + 8) * 4, (double)B_ * tb.nleaves);
        launch_leaf_hash(limb_d[j], 6, ext_msg_d, (int)ext_msg.size(), 0u, hdr_ext, leaf_d[j], tb.nleaves, B_, s_,
                         tune_.hash_threads);
      }
      uint32_t* dig_j = chal_dig.as<uint32_t>() + (size_t)j * fb;
      uint32_t* al_j = chal_alpha.as<uint32_t>() + (size_t)j * B_ * 6;
      const RootChallenge rc{string_to_elements(TAG_FSID, "fri_fold_" + std::to_string(j)), prev_dig, dig_j, al_j,
                             root_dev.as<uint32_t>() + (size_t)j * fb};
      if (grind_bits_) {
        tree_levels(tb, tree_d[j]);
        launch_copy_roots(tree_d[j], rc.roots, B_, s_);
        grind_fold(j, prev_dig, tb);
      } else if (!tree_levels(tb, tree_d[j], &rc)) {
        launch_fri_challenge(rc.prefix.data(), (int)rc.prefix.size(), prev_dig, tree_d[j], dig_j, al_j, rc.roots, B_,
                             s_);
      }
      prev_dig = dig_j;
      alpha_j = al_j;
    }
    int log_len = 0;
    while ((1u << log_len) < layers[j].len) ++log_len;
    const uint32_t* xi = tb_.fri_xinv2(log_len);
    const int half = (int)layers[j + 1].len;
    if (j + 1 < FRI_ROUNDS && !joins[j + 1]) {
      TimeIt tm(ctx_, s_, "fri_fold_leaf", (double)B_ * layers[j].len * 1.5 * 24, (double)B_ * half / 2);
      launch_fri_fold_leaf(lay_d[j], lay_d[j + 1], alpha_j, xi, half, hdr_ext, leaf_d[j + 1], B_, s_,
                           tune_.fri_coop_pairs);
      leaves_done[j + 1] = 1;
    } else {
      TimeIt tm(ctx_, s_, "fri_fold", (double)B_ * layers[j].len * 1.5 * 24, 0);
      launch_fri_fold(lay_d[j], lay_d[j + 1], alpha_j, xi, half, B_, s_);
    }
  }
  Layer running = std::move(layers[FRI_ROUNDS]);
  // final polynomial (running after the last round): 4 E6 values
  const uint32_t flen = running.len;
  const size_t fw_words = (size_t)B_ * 6 * flen;
  // digests, roots, final polynomial, proof-of-work salts
  const size_t salt_words = grind_bits_ ? (size_t)FRI_ROUNDS * B_ : 0;
  uint32_t* fri_h = pin((2 * (size_t)FRI_ROUNDS * fb + fw_words + salt_words) * 4);
  copy_async(fri_h, chal_dig.as<void>(), (size_t)FRI_ROUNDS * fb * 4, hipMemcpyDeviceToHost);
  copy_async(fri_h + FRI_ROUNDS * fb, root_dev.as<void>(), (size_t)FRI_ROUNDS * fb * 4, hipMemcpyDeviceToHost);
  copy_async(fri_h + 2 * FRI_ROUNDS * fb, running.buf.as<void>(), fw_words * 4, hipMemcpyDeviceToHost);
  if (salt_words)
    copy_async(fri_h + 2 * FRI_ROUNDS * fb + fw_words, fri_salts_.as<void>(), salt_words * 4, hipMemcpyDeviceToHost);
  region_end();  // the device phase of the prove ends here
  mark("device phase enqueued");
  sync();
  mark("device phase done");
  if (grind_bits_) salts_h = fri_h + 2 * FRI_ROUNDS * fb + fw_words;
  if (deep_on_device) replay_alpha();
  mark("transcripts replayed");
  std::vector<uint32_t> final_w(fri_h + 2 * FRI_ROUNDS * fb, fri_h + 2 * FRI_ROUNDS * fb + fw_words);
  {
    std::vector<uint32_t> dg(fri_h, fri_h + FRI_ROUNDS * fb), rt(fri_h + FRI_ROUNDS * fb, fri_h + 2 * FRI_ROUNDS * fb);
    for (int j = first + 1; j < FRI_ROUNDS; ++j) {
      std::vector<Digest> root(B_);
      for (int b = 0; b < B_; ++b)
        for (int k = 0; k < 8; ++k) root[b].v[k] = rt[(size_t)j * fb + (size_t)8 * b + k];
      fri_roots.push_back(root);
    }
    if (!first_on_host)
      for (int b = 0; b < B_; ++b)
        for (int k = 0; k < 8; ++k) levels_root[b].v[k] = rt[(size_t)first * fb + (size_t)8 * b + k];
    fri_dig = std::move(dg);
  }
  keep_.push_back(std::move(chal_dig));
  keep_.push_back(std::move(chal_alpha));
  keep_.push_back(std::move(root_dev));
  trace("fri rounds");
  // ---- query openings.  The gather ops (one record per (instance, query))
  // are built and staged first; the query chains then run on the host, and
  // as soon as every position of a chunk of instances is known, that chunk's
  // positions go to the device, its records are gathered and copied back,
  // and an event marks them ready for the proof assembly.
  // The top PATH_TOP levels of every tree deeper than that are shared by the
  // queries' authentication paths: they are copied back once per tree
  // (nodes 0 .. 2^(PATH_TOP+1) of the heap layout) while the query chains
  // run, and the records carry only the lower siblings of each path.
  const int chunk = std::min(B_, 8);
  constexpr int PATH_TOP = 7;
  constexpr size_t TAB_WORDS = (size_t)8 << (PATH_TOP + 1);
  int ntab = 0;
  std::vector<const uint32_t*> tab_src;  // [table][b] node arrays
  auto path_split = [&](int depth, int& low, int& tab, const TreeBuf& tb) {
    if (depth > PATH_TOP) {
      low = depth - PATH_TOP;
      tab = ntab++;
      for (int b = 0; b < B_; ++b) tab_src.push_back(tb.node_ptr(b));
    } else {
      low = depth;
      tab = -1;
    }
  };
  size_t stride = 0;
  auto seg = [&](size_t words) {  // 16-byte aligned segment of the record
    size_t o = (stride + 3) & ~size_t(3);
    stride = o + words;
    return o;
  };
  struct TreeRec {
    size_t rows_off, path_off, injl_off;
    int ncols_top, depth, shift, low, tab;
    std::vector<int> inj_ncols, inj_shift;
    std::vector<size_t> inj_rows_off;
  };
  std::vector<TreeRec> trec(nt);
  auto cols_of_group = [&](const Group& g, int b) {
    std::vector<const uint32_t*> cp;
    for (auto c : g.base) cp.push_back(cols_[c].lde(b, 0));
    for (auto c : g.ext)
      for (int l = 0; l < 6; ++l) cp.push_back(cols_[c].lde(b, l));
    return cp;
  };
  std::vector<GatherOp> ops;
  int blocks = 0;
  auto add_op = [&](int kind, int shift, int ncols, int d, int nleaves, size_t off,
                    const std::vector<const uint32_t*>& ptrs) {
    GatherOp op{};
    op.kind = kind;
    op.shift = shift;
    op.ncols = ncols;
    op.d = d;
    op.nleaves = nleaves;
    op.out_off = (int32_t)off;
    op.ptrs = stage(ptrs);
    const long per = kind == 0 ? 2 * ncols : (kind == 1 ? ncols : 1);
    const long total = (long)chunk * nq * per;  // one launch per chunk of instances
    op.block0 = blocks;
    op.nblocks = (int)((total + 255) / 256);
    blocks += op.nblocks;
    if (op.nblocks) ops.push_back(op);
  };
  for (size_t t = 0; t < nt; ++t) {
    const TreeInfo& ti = d_.trees[t];
    const TreeBuf& tb = trees_[t];
    TreeRec& r = trec[t];
    const Group& top = ti.groups[tb.order[0]];
    r.ncols_top = (int)(top.base.size() + 6 * top.ext.size());
    r.depth = 0;
    while ((1 << r.depth) < tb.nleaves) ++r.depth;
    r.shift = FRI_LOG_N - (r.depth + 1);
    path_split(r.depth, r.low, r.tab, tb);
    r.rows_off = seg(2 * r.ncols_top);
    r.path_off = seg(8 * r.low);
    r.injl_off = seg(8 * tb.inj.size());
    std::vector<const uint32_t*> cp, tp;
    for (int b = 0; b < B_; ++b) {
      auto c = cols_of_group(top, b);
      cp.insert(cp.end(), c.begin(), c.end());
      tp.push_back(tb.node_ptr(b));
    }
    add_op(0, r.shift, r.ncols_top, 0, 0, r.rows_off, cp);
    add_op(1, r.shift, r.low, 0, tb.nleaves, r.path_off, tp);
    for (size_t k = 0; k < tb.inj.size(); ++k) {
      const Group& g = ti.groups[tb.order[k + 1]];
      const int nc = (int)(g.base.size() + 6 * g.ext.size());
      int lr = 0;
      while ((1u << lr) < g.N * 4) ++lr;
      r.inj_ncols.push_back(nc);
      r.inj_shift.push_back(FRI_LOG_N - lr);
      r.inj_rows_off.push_back(seg(2 * nc));
      std::vector<const uint32_t*> ip, gp;
      for (int b = 0; b < B_; ++b) {
        ip.push_back(tb.inj_ptr(k, b));
        auto c = cols_of_group(g, b);
        gp.insert(gp.end(), c.begin(), c.end());
      }
      int dsh = 0;
      while ((tb.nleaves >> dsh) > tb.inj_width[k]) ++dsh;
      add_op(2, r.shift, 1, dsh, 0, r.injl_off + 8 * k, ip);
      add_op(0, r.inj_shift.back(), nc, 0, 0, r.inj_rows_off.back(), gp);
    }
  }
  struct FriRec {
    size_t rows_off, path_off;
    int depth, low, tab, nleaves;
  };
  std::vector<FriRec> frec(FRI_ROUNDS);
  for (int j = first; j < FRI_ROUNDS; ++j) {
    FriRec& r = frec[j];
    const Layer& lay = (j == first) ? *lvp[0] : layers[j];
    const TreeBuf& tb = (j == first) ? levels_tree : ftrees[j];
    r.depth = FRI_LOG_N - j - 1;
    r.nleaves = tb.nleaves;
    path_split(r.depth, r.low, r.tab, tb);
    r.rows_off = seg(12);
    r.path_off = seg(8 * r.low);
    std::vector<const uint32_t*> cp, tp;
    for (int b = 0; b < B_; ++b) {
      for (int l = 0; l < 6; ++l) cp.push_back(lay.ptr(b) + (size_t)l * lay.len);
      tp.push_back(tb.node_ptr(b));
    }
    add_op(0, j, 6, 0, 0, r.rows_off, cp);
    add_op(1, j, r.low, 0, tb.nleaves, r.path_off, tp);
  }
  std::vector<size_t> lq_off(L);
  for (size_t l = 1; l < L; ++l) {
    lq_off[l] = seg(12);
    const int jl = FRI_LOG_D - d_.modules[classes[lv_order[l]]].log_n;
    std::vector<const uint32_t*> cp;
    for (int b = 0; b < B_; ++b)
      for (int k = 0; k < 6; ++k) cp.push_back(lvp[l]->ptr(b) + (size_t)k * lvp[l]->len);
    add_op(0, jl, 6, 0, 0, lq_off[l], cp);
  }
  stride = (stride + 3) & ~size_t(3);
  const size_t gwords = (size_t)B_ * nq * stride;
  const size_t twords = (size_t)ntab * B_ * TAB_WORDS;
  DevBuf gat = dev(gwords + twords);
  DevBuf pos_dev = dev((size_t)B_ * nq);
  const GatherOp* ops_d = stage(ops);
  std::vector<uint32_t*> tab_dst(tab_src.size());
  for (size_t i = 0; i < tab_src.size(); ++i) tab_dst[i] = gat.as<uint32_t>() + gwords + i * TAB_WORDS;
  const uint32_t* const* tab_src_d = tab_src.empty() ? nullptr : stage(tab_src);
  uint32_t* const* tab_dst_d = tab_src.empty() ? nullptr : const_cast<uint32_t* const*>(stage(tab_dst));
  stg().flush(s_);  // the chunk launches below stage nothing
  uint32_t* gh = ctx_.pinned((gwords + twords) * 4, lane_);
  const uint32_t* tabs_h = gh + gwords;  // [table][b][TAB_WORDS]
  if (twords) {
    launch_copy_cols(tab_src_d, tab_dst_d, (int)TAB_WORDS, (int)tab_src.size(), s_);
    LG_HIP(hipMemcpyAsync(gh + gwords, gat.as<uint32_t>() + gwords, twords * 4, hipMemcpyDeviceToHost, s_));
  }
  uint32_t* pos = ctx_.pinned_in((size_t)B_ * nq * 4, lane_);  // the inputs' upload finished long ago
  struct Events {
    std::vector<hipEvent_t> e;
    ~Events() {
      for (auto x : e) (void)hipEventDestroy(x);
    }
  } ev;
  const int nchunks = (B_ + chunk - 1) / chunk;
  std::vector<std::atomic<int>> left(nchunks);
  for (int c = 0; c < nchunks; ++c) {
    left[c] = std::min(chunk, B_ - c * chunk);
    hipEvent_t e;
    LG_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    ev.e.push_back(e);
  }
  std::mutex launch_mu;
  auto launch_chunk = [&](int c) {
    const int b0 = c * chunk, nb = std::min(chunk, B_ - b0);
    const size_t q0 = (size_t)b0 * nq, w0 = q0 * stride;
    std::lock_guard<std::mutex> lk(launch_mu);
    LG_HIP(hipSetDevice(ctx_.device));
    LG_HIP(hipMemcpyAsync(pos_dev.as<uint32_t>() + q0, pos + q0, (size_t)nb * nq * 4, hipMemcpyHostToDevice, s_));
    {
      TimeIt tm(ctx_, s_, "gather", (double)nb * nq * stride * 8, 0);
      launch_gather(ops_d, (int)ops.size(), blocks, pos_dev.as<uint32_t>(), nq, b0, nb, gat.as<uint32_t>(),
                    (int)stride, s_);
    }
    LG_HIP(hipMemcpyAsync(gh + w0, gat.as<uint32_t>() + w0, (size_t)nb * nq * stride * 4, hipMemcpyDeviceToHost, s_));
    LG_HIP(hipEventRecord(ev.e[c], s_));
  };
  // A chunk's records may be awaited only once its event has been recorded.
  std::mutex ready_mu;
  std::condition_variable ready_cv;
  std::vector<char> ready(nchunks, 0);
  bool failed = false;
  auto mark = [&](int c, bool ok) {
    {
      std::lock_guard<std::mutex> lk(ready_mu);
      if (ok) ready[c] = 1;
      else failed = true;
    }
    ready_cv.notify_all();
  };
  // host timestamps of the phase (printed with LOOMGPU_TRACE)
  const double q_t0 = now_s();
  std::vector<double> q_chain(B_), q_back(B_), q_done(B_), q_cstart(B_), q_astart(B_);
  // The host transcript binds the fold roots and computes every fold
  // challenge itself, checking the device's digests (each query task does
  // its instance before the query chain).
  const size_t fold0 = ts_.empty() ? 0 : ts_[0].index(first_name);  // fri_fold_j at fold0 + j - first
  // with grinding, fold round j's message ends with its proof of work (the
  // device's salt); the transcript checks the digest's zero bits
  auto pow_at = [&](int b, int j, size_t p) {
    if (grind_bits_) ts_[b].set_pow(p, grind_bits_, salts_h[(size_t)j * B_ + b]);
  };
  std::vector<int> fold_name_order(FRI_ROUNDS);
  for (int j = 0; j < FRI_ROUNDS; ++j) fold_name_order[j] = j;
  std::sort(fold_name_order.begin(), fold_name_order.end(),
            [](int a, int b) { return "fri_fold_" + std::to_string(a) < "fri_fold_" + std::to_string(b); });
  auto replay_fri = [&](int b) {
    if (!first_on_host || grind_bits_) {
      if (first_on_host) {  // the zero-bound rounds were found on the device too
        static const uint32_t zero[8] = {0};
        for (int j = 0; j < first; ++j) {
          ts_[b].bind_at(fold_pos_ + j, zero, 8);
          pow_at(b, j, fold_pos_ + j);
          ts_[b].check_at(fold_pos_ + j, &fri_dig[(size_t)j * fb + (size_t)8 * b]);
        }
      }
      ts_[b].bind_at(fold0, levels_root[b].v, 8);
      pow_at(b, first, fold0);
      ts_[b].check_at(fold0, &fri_dig[(size_t)first * fb + (size_t)8 * b]);
    }
    for (int j = first + 1; j < FRI_ROUNDS; ++j) {
      const size_t p = fold0 + (size_t)(j - first);
      ts_[b].bind_at(p, fri_roots[j - first - 1][b].v, 8);
      pow_at(b, j, p);
      ts_[b].check_at(p, &fri_dig[(size_t)j * fb + (size_t)8 * b]);
    }
  };
  auto query_task = [&](int b) {
    q_cstart[b] = now_s();
    try {
      replay_fri(b);
      std::vector<uint32_t> fp{M(TAG_EXTP), M(flen)};
      for (uint32_t k = 0; k < flen; ++k)
        for (int l = 0; l < 6; ++l) fp.push_back(final_w[(size_t)b * 6 * flen + (size_t)l * flen + k]);
      ts_[b].bind("fri_query_0", fp);
      const size_t p0 = ts_[b].index("fri_query_0");
      for (int k = 0; k < nq; ++k) {
        const Digest& dg = ts_[b].compute_at(p0 + k);
        uint64_t c0 = kb_from_mont(dg.v[0]), c1 = kb_from_mont(dg.v[1]);
        pos[(size_t)b * nq + k] = (uint32_t)(((c0 << 31) ^ c1) % (1ull << FRI_LOG_N));
        if (k < nq - 1) ts_[b].bind_at(p0 + k + 1, dg.v, 8);
      }
      q_chain[b] = now_s();
      if (--left[b / chunk] == 0) {
        launch_chunk(b / chunk);
        mark(b / chunk, true);
      }
    } catch (...) {
      mark(0, false);
      throw;
    }
  };
  // ---- proof assembly (canonical encoding, see include/loomgpu.h)
  auto assemble_task = [&](int b) {
   q_astart[b] = now_s();
   {
     std::unique_lock<std::mutex> lk(ready_mu);
     ready_cv.wait(lk, [&] { return failed || ready[b / chunk]; });
     if (!ready[b / chunk]) return;  // a query task failed; its error is rethrown
   }
   LG_HIP(hipEventSynchronize(ev.e[b / chunk]));
   q_back[b] = now_s();
   if (!inst_err_[b].empty()) {  // a failed instance gets no proof
     out_[b].error = inst_err_[b];
     out_[b].proof_len = 0;
     q_done[b] = now_s();
     return;
   }
   // The words go straight to the caller's buffer when it gave one
   // (SpanWriter counts past its capacity so the caller learns the length),
   // else to out_[b].proof.
   // an authentication path: the gathered lower siblings, then the top
   // levels from the tree's table (heap index (nleaves + leaf) >> e, ^ 1)
   auto put_path = [&](auto& w, const uint32_t* low_rec, int depth, int low, int tab, uint32_t leaf,
                       uint32_t nleaves) {
    w.insert(w.end(), low_rec, low_rec + 8 * low);
    if (tab < 0) return;
    const uint32_t* tt = tabs_h + ((size_t)tab * B_ + b) * TAB_WORDS;
    for (int e = low; e < depth; ++e) {
      const uint32_t node = ((nleaves + leaf) >> e) ^ 1u;
      w.insert(w.end(), tt + 8 * (size_t)node, tt + 8 * (size_t)node + 8);
    }
   };
   auto emit = [&](auto& w) {
    emit_head(w, b, cvals.data() + (size_t)b * ncv);
    const uint32_t* rec0 = gh + (size_t)b * nq * stride;
    // FRI proof
    w.insert(w.end(), levels_root[b].v, levels_root[b].v + 8);
    w.push_back((uint32_t)(L - 1));
    for (size_t l = 1; l < L; ++l) {
      const int jl = FRI_LOG_D - d_.modules[classes[lv_order[l]]].log_n;
      w.push_back((uint32_t)nq);
      for (int q = 0; q < nq; ++q) {
        const uint32_t* rec = rec0 + (size_t)q * stride;
        const uint32_t row = pos[(size_t)b * nq + q] >> jl;
        w.push_back(1);
        w.push_back(row);
        for (int x = 0; x < 6; ++x) w.push_back(rec[lq_off[l] + x]);
        for (int x = 0; x < 6; ++x) w.push_back(rec[lq_off[l] + 6 + x]);
        w.push_back(0);
        w.push_back(0);
        w.push_back(0);
      }
    }
    w.push_back((uint32_t)fri_roots.size());
    for (auto& r : fri_roots) w.insert(w.end(), r[b].v, r[b].v + 8);
    w.push_back(1);
    w.push_back(flen);
    for (uint32_t e = 0; e < flen; ++e)
      for (int l = 0; l < 6; ++l) w.push_back(final_w[(size_t)b * 6 * flen + (size_t)l * flen + e]);
    w.push_back((uint32_t)nq);
    for (int q = 0; q < nq; ++q) {
      const uint32_t* rec = rec0 + (size_t)q * stride;
      const uint32_t s = pos[(size_t)b * nq + q];
      w.push_back((uint32_t)(FRI_ROUNDS - first));
      for (int j = first; j < FRI_ROUNDS; ++j) {
        const uint32_t row = s >> j;
        w.push_back(1);
        w.push_back(row);
        for (int x = 0; x < 6; ++x) w.push_back(rec[frec[j].rows_off + x]);
        for (int x = 0; x < 6; ++x) w.push_back(rec[frec[j].rows_off + 6 + x]);
        w.push_back((row & ~1u) >> 1);
        w.push_back((uint32_t)frec[j].depth);
        put_path(w, rec + frec[j].path_off, frec[j].depth, frec[j].low, frec[j].tab, row >> 1,
                 (uint32_t)frec[j].nleaves);
        w.push_back(0);  // layer r0's injection leaves are dropped (fri.go:719); others have none
      }
    }
    // proofs of work (fri.Proof.PoW, recordFoldProofsOfWork, fri.go:1075-1090):
    // one per fold challenge, in ascending byte order of the names
    if (grind_bits_) {
      w.push_back((uint32_t)FRI_ROUNDS);
      for (int j : fold_name_order) {
        const std::string nm = "fri_fold_" + std::to_string(j);
        w.push_back((uint32_t)nm.size());
        for (size_t i = 0; i < nm.size(); i += 4) {
          uint32_t v = 0;
          for (size_t k = 0; k < 4 && i + k < nm.size(); ++k) v |= (uint32_t)(uint8_t)nm[i + k] << (8 * k);
          w.push_back(v);
        }
        w.push_back(grind_bits_);
        w.push_back(kb_to_mont(salts_h[(size_t)j * B_ + b]));
      }
    } else {
      w.push_back(0);
    }
    w.push_back((uint32_t)nq);
    for (int q = 0; q < nq; ++q) {
      const uint32_t* rec = rec0 + (size_t)q * stride;
      const uint32_t s = pos[(size_t)b * nq + q];
      w.push_back((uint32_t)nt);
      for (size_t t = 0; t < nt; ++t) {
        const TreeInfo& ti = d_.trees[t];
        const TreeBuf& tb = trees_[t];
        const TreeRec& r = trec[t];
        auto rawrow = [&](const Group& g, const uint32_t* vals, int ncols, int hi) {
          const uint32_t* v = vals + (hi ? ncols : 0);
          w.push_back((uint32_t)g.base.size());
          w.insert(w.end(), v, v + g.base.size());
          w.push_back((uint32_t)g.ext.size());
          w.insert(w.end(), v + g.base.size(), v + ncols);
        };
        const Group& top = ti.groups[tb.order[0]];
        rawrow(top, rec + r.rows_off, r.ncols_top, 0);
        rawrow(top, rec + r.rows_off, r.ncols_top, 1);
        w.push_back(((s >> r.shift) & ~1u) >> 1);
        w.push_back((uint32_t)r.depth);
        put_path(w, rec + r.path_off, r.depth, r.low, r.tab, (s >> r.shift) >> 1, (uint32_t)tb.nleaves);
        w.push_back((uint32_t)tb.inj.size());
        w.insert(w.end(), rec + r.injl_off, rec + r.injl_off + 8 * tb.inj.size());
        w.push_back((uint32_t)tb.inj.size());
        for (size_t k2 = 0; k2 < tb.inj.size(); ++k2) {
          const Group& g = ti.groups[tb.order[k2 + 1]];
          rawrow(g, rec + r.inj_rows_off[k2], r.inj_ncols[k2], 0);
          rawrow(g, rec + r.inj_rows_off[k2], r.inj_ncols[k2], 1);
        }
      }
    }
   };
    if (in_[b].proof_dst) {
      SpanWriter w{in_[b].proof_dst, in_[b].proof_cap};
      emit(w);
      out_[b].proof_len = w.n;
      out_[b].direct = true;
    } else {
      std::vector<uint32_t>& w = out_[b].proof;
      w.clear();
      w.reserve(64 + (size_t)nq * (stride + 64 * (nt + FRI_ROUNDS)) + 16 * ncv);
      emit(w);
      out_[b].proof_len = w.size();
    }
    out_[b].transcript = ts_[b].log_words();
    q_done[b] = now_s();
  };
  // Query chains first (tasks are handed out in index order), then the
  // assemblies, which start as soon as threads free up and their chunk is
  // back, overlapping the last query chains.
  this->mark("query phase start");
  parallel_for(2 * B_, [&](int t) {
    if (t < B_) query_task(t);
    else assemble_task(t - B_);
  });
  this->mark("proofs assembled");
  if (getenv("LOOMGPU_TRACE")) {
    auto last = [&](const std::vector<double>& v) { return (*std::max_element(v.begin(), v.end()) - q_t0) * 1e3; };
    double chain = 0, wait = 0, assemble = 0;
    for (int b = 0; b < B_; ++b) {
      chain += q_chain[b] - q_cstart[b];
      wait += q_back[b] - q_astart[b];
      assemble += q_done[b] - q_back[b];
    }
    fprintf(stderr,
            "[loomgpu] B=%d queries: chains %.3f ms, records back %.3f ms, assembled %.3f ms, %zu words/proof, "
            "%zu record words; per instance: chain %.3f ms, wait %.3f ms, assembly %.3f ms\n",
            B_, last(q_chain), last(q_back), last(q_done), out_[0].proof_len, gwords, chain / B_ * 1e3,
            wait / B_ * 1e3, assemble / B_ * 1e3);
  }
  trace("queries, gathers, proof assembly");
  for (auto& x : W) keep_.push_back(std::move(x.second));
  for (auto& x : level) keep_.push_back(std::move(x.buf));
  for (auto& x : layers) keep_.push_back(std::move(x.buf));
  keep_.push_back(std::move(running.buf));
  keep_.push_back(std::move(gat));
  keep_.push_back(std::move(levels_tree.nodes));
  for (auto& x : levels_tree.inj) keep_.push_back(std::move(x));
  for (auto& t : ftrees) {
    keep_.push_back(std::move(t.nodes));
    for (auto& x : t.inj) keep_.push_back(std::move(x));
  }
}

}  // namespace

Plan::~Plan() = default;

void Plan::set_setup(const uint32_t* const* cols, size_t n_cols, uint32_t* root) {
  if (n_cols != d_.n_setup) throw std::runtime_error("set_setup: the plan has " + std::to_string(d_.n_setup) +
                                                     " setup columns");
  if (!d_.n_setup) throw std::runtime_error("set_setup: the plan has no setup columns");
  drop_graphs();  // recorded graphs hold the previous key's root and buffers
  auto sd = std::make_unique<SetupData>();
  LG_HIP(hipSetDevice(ctx_->device));
  std::vector<InstanceInput> in(1);
  std::vector<ProveOutput> out(1);
  {
    Job job(*this, in, out, nullptr, 0);
    job.commit_setup(cols, *sd);
  }
  LG_HIP(hipStreamSynchronize(ctx_->stream));
  std::memcpy(root, sd->root, 32);
  setup_ = std::move(sd);
}

void Plan::prove(const std::vector<InstanceInput>& inputs, std::vector<ProveOutput>& outputs, const PhaseCallback& cb) {
  if (inputs.empty()) return;
  if (d_.n_setup && !setup_) throw std::runtime_error("prove: the plan's setup columns are not committed (lg_plan_set_setup)");
  outputs.assign(inputs.size(), ProveOutput());
  const int B = (int)inputs.size();
  cur_batch_ = B;
  constexpr int kMinLaneBatch = 8;
  const int want_lanes = ctx_->lanes_forced ? ctx_->lanes : tuning(B).lanes;
  int lanes = std::min(want_lanes, std::max(1, B / kMinLaneBatch));
  if (ctx_->timer.on) lanes = 1;  // per-kernel timing wants one stream
  struct BucketHold {  // returns a lane's bucket however the lane ends
    Plan* p;
    GraphBucket* b;
    ~BucketHold() { p->give_bucket(b); }
  };
  // A recording can fail for reasons outside the prove (another thread's
  // legacy-stream work invalidates a capture): the bucket is then marked
  // broken and the same prove runs again eagerly.
  auto run_job = [&](const std::vector<InstanceInput>& in, std::vector<ProveOutput>& out, const PhaseCallback& pcb,
                     int lane, GraphBucket* gb, int mode) {
    if (mode == Job::CAPTURE || mode == Job::REPLAY) {
      // a failed recording or replay breaks the bucket; the prove runs again eagerly
      try {
        Job job(*this, in, out, pcb, lane, gb, (Job::Mode)mode);
        job.run();
        ++(mode == Job::CAPTURE ? graph_records : graph_replays);
        return;
      } catch (const std::exception&) {
        gb->broken = true;
        out.assign(in.size(), ProveOutput());
        ++graph_fallbacks;
      }
    }
    Job job(*this, in, out, pcb, lane);
    job.run();
  };
  if (lanes <= 1) {
    int mode = 0;
    BucketHold hold{this, take_bucket(B, 0, inputs[0].device, &mode)};
    run_job(inputs, outputs, cb, 0, hold.b, mode);
    LG_HIP(hipStreamSynchronize(ctx_->stream));
    ctx_->timer.resolve();
    return;
  }
  // Contiguous sub-batches, one per lane; instances are independent, so the
  // proofs are those of one batch.
  std::vector<std::vector<InstanceInput>> part(lanes);
  std::vector<std::vector<ProveOutput>> out(lanes);
  std::vector<int> first(lanes + 1, 0);
  for (int l = 0; l < lanes; ++l) first[l + 1] = first[l] + B / lanes + (l < B % lanes ? 1 : 0);
  for (int l = 0; l < lanes; ++l) {
    part[l].assign(inputs.begin() + first[l], inputs.begin() + first[l + 1]);
    out[l].assign(part[l].size(), ProveOutput());  // a Job writes its outputs in place
  }
  std::vector<std::exception_ptr> err(lanes);
  // host timestamps of the lanes' ends (LOOMGPU_MARKS)
  static const bool marks_on = getenv("LOOMGPU_MARKS") != nullptr;
  const double t_call = marks_on ? now_s() : 0.0;
  std::vector<double> t_job(lanes), t_sync(lanes);
  auto run_lane = [&](int l) {
    try {
      current_pool() = &ThreadPool::for_lane(l, lanes);
      int mode = 0;
      BucketHold hold{this, take_bucket((int)part[l].size(), l, inputs[0].device, &mode)};
      run_job(part[l], out[l], l == 0 ? cb : PhaseCallback(), l, hold.b, mode);  // phases: lane 0's
      if (marks_on) t_job[l] = now_s();
      current_pool() = nullptr;
      LG_HIP(hipStreamSynchronize(ctx_->lane_stream(l)));
      if (marks_on) t_sync[l] = now_s();
    } catch (...) {
      current_pool() = nullptr;
      err[l] = std::current_exception();
    }
  };
  std::vector<std::thread> th;
  for (int l = 1; l < lanes; ++l) th.emplace_back(run_lane, l);
  run_lane(0);
  for (auto& t : th) t.join();
  for (auto& e : err)
    if (e) std::rethrow_exception(e);
  for (int l = 0; l < lanes; ++l)
    for (int b = 0; b < (int)out[l].size(); ++b) outputs[first[l] + b] = std::move(out[l][b]);
  if (marks_on) {
    std::string line = "[loomgpu] prove marks (ms): t0 " + std::to_string(t_call) + " s;";
    char buf[96];
    for (int l = 0; l < lanes; ++l) {
      snprintf(buf, sizeof buf, " lane %d job returned %.3f, synced %.3f;", l, (t_job[l] - t_call) * 1e3,
               (t_sync[l] - t_call) * 1e3);
      line += buf;
    }
    snprintf(buf, sizeof buf, " outputs moved %.3f;", (now_s() - t_call) * 1e3);
    line += buf;
    fprintf(stderr, "%s\n", line.c_str());
  }
}

}  // namespace lg
