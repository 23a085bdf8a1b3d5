// extern "C" boundary of libloomgpu.so (declarations and contracts in
// include/loomgpu.h).  Exceptions never cross it: they become a negative
// return code and a thread-local message.
#include "../../include/loomgpu.h"

#include <algorithm>
#include <atomic>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>

#include "kernels.h"
#include "p2.h"
#include "pool.h"
#include "prover.h"
#include "server.h"
#include "verifier.h"

struct lg_ctx {
  lg::Context c;
  std::mutex mu;
};

struct lg_server {
  std::unique_ptr<lg::Server> s;
};

struct lg_plan {
  lg_ctx* ctx;
  std::unique_ptr<lg::Plan> p;
  double phases[3] = {0, 0, 0};
  std::vector<std::string> verify_msgs;  // of the last lg_verify
};

namespace {
thread_local std::string g_err;

template <class F>
int guarded(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const std::exception& e) {
    g_err = e.what();
    return -1;
  } catch (...) {
    g_err = "unknown error";
    return -1;
  }
}

lg::DevBuf stage_sync(const void* src, size_t bytes) {
  lg::DevBuf b(bytes);
  LG_HIP(hipMemcpy(b.as<void>(), src, bytes, hipMemcpyHostToDevice));
  return b;
}
}  // namespace

extern "C" {

const char* lg_last_error(void) { return g_err.c_str(); }

int lg_ctx_create(int device, lg_ctx** out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("lg_ctx_create: null out");
    int n = 0;
    LG_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw std::runtime_error("lg_ctx_create: no such HIP device");
    LG_HIP(hipSetDevice(device));
    lg::noalloc_init(device);  // this device's scratch address for lg_grand_product_noalloc
    auto* c = new lg_ctx();
    c->c.device = device;
    if (hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking) != hipSuccess) {
      delete c;
      throw std::runtime_error("lg_ctx_create: stream creation failed");
    }
    *out = c;
    return 0;
  });
}

void lg_ctx_destroy(lg_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->c.device);
  (void)hipStreamSynchronize(ctx->c.stream);
  (void)hipStreamDestroy(ctx->c.stream);
  delete ctx;
}

int lg_plan_create(lg_ctx* ctx, const uint32_t* words, size_t n_words, lg_plan** out) {
  return guarded([&] {
    if (!ctx || !words || !out) throw std::runtime_error("lg_plan_create: null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    auto* p = new lg_plan();
    p->ctx = ctx;
    try {
      p->p = std::make_unique<lg::Plan>(&ctx->c, words, n_words);
    } catch (...) {
      delete p;
      throw;
    }
    *out = p;
    return 0;
  });
}

void lg_plan_destroy(lg_plan* plan) { delete plan; }

int lg_plan_check(const uint32_t* words, size_t n_words, size_t* code_bytes) {
  return guarded([&] {
    if (!words) throw std::runtime_error("lg_plan_check: null words");
    lg::PlanData d = lg::parse_plan(words, n_words);
    size_t total = 0;
    auto one = [&](const lg::Dag& dag) {
      std::vector<uint32_t> c, h;
      total += lg::dag_kernel_compile_only(lg::dag_kernel_source(dag, d.cols, c, h));
    };
    for (auto& m : d.modules)
      if (m.n_chunks) one(m.relation);
    for (auto& r : d.rounds)
      for (auto& st : r.steps)
        for (auto& dag : st.ins) one(dag);
    if (code_bytes) *code_bytes = total;
    return 0;
  });
}

int lg_plan_set_setup(lg_plan* plan, const uint32_t* const* setup_columns, size_t n_columns, uint32_t* root) {
  return guarded([&] {
    if (!plan || !setup_columns || !root) throw std::runtime_error("lg_plan_set_setup: null argument");
    for (size_t k = 0; k < n_columns; ++k)
      if (!setup_columns[k]) throw std::runtime_error("lg_plan_set_setup: null column pointer");
    std::lock_guard<std::mutex> lock(plan->ctx->mu);
    plan->p->set_setup(setup_columns, n_columns, root);
    return 0;
  });
}

int lg_plan_num_inputs(const lg_plan* plan, size_t* n_columns, size_t* n_perms) {
  return guarded([&] {
    if (!plan) throw std::runtime_error("lg_plan_num_inputs: null plan");
    if (n_columns) *n_columns = plan->p->num_inputs();
    if (n_perms) *n_perms = plan->p->num_perms();
    return 0;
  });
}

static int prove_impl(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
                      bool device, uint32_t* const* proofs, const size_t* proof_caps, size_t* proof_lens,
                      uint32_t* const* transcripts, const size_t* transcript_caps, size_t* transcript_lens,
                      const uint32_t* fs_prev = nullptr) {
  return guarded([&] {
    if (!plan || batch <= 0 || !proofs || !proof_caps || !proof_lens) throw std::runtime_error("lg_prove: bad arguments");
    const size_t nc = plan->p->num_inputs(), np = plan->p->num_perms();
    if ((nc && !columns) || (np && !perms)) throw std::runtime_error("lg_prove: missing inputs");
    std::vector<lg::InstanceInput> in(batch);
    for (int b = 0; b < batch; ++b) {
      for (size_t k = 0; k < nc; ++k) {
        if (!columns[b * nc + k]) throw std::runtime_error("lg_prove: null column pointer");
        in[b].columns.push_back(columns[b * nc + k]);
      }
      for (size_t k = 0; k < np; ++k) {
        if (!perms[b * np + k]) throw std::runtime_error("lg_prove: null permutation pointer");
        in[b].perms.push_back(perms[b * np + k]);
      }
      in[b].device = device;
      in[b].proof_dst = proofs[b];
      in[b].proof_cap = proof_caps[b];
      if (fs_prev) {
        in[b].chained = true;
        std::memcpy(in[b].fs_prev, fs_prev + (size_t)8 * b, 32);
      }
    }
    std::vector<lg::ProveOutput> out;
    {
      std::lock_guard<std::mutex> lock(plan->ctx->mu);
      int ph = 0;
      double* phases = plan->phases;
      plan->p->prove(in, out, [&](const char*, double dt) {
        if (ph < 3) phases[ph++] = dt;
      });
    }
    std::atomic<int> rc{0};
    std::string failed;
    for (int b = 0; b < batch; ++b)
      if (!out[b].error.empty()) failed += (failed.empty() ? "" : "; ") + std::string("instance ") + std::to_string(b) +
                                           ": " + out[b].error;
    // copies only for proofs assembled into the library's own buffers and
    // for transcripts; otherwise the loop is too short for the thread pool
    bool copies = false;
    for (int b = 0; b < batch; ++b)
      if (out[b].error.empty() && (!out[b].direct || (transcripts && transcripts[b]))) copies = true;
    auto finish = [&](int b) {
      proof_lens[b] = out[b].proof_len;
      if (!out[b].error.empty()) {
        proof_lens[b] = 0;
        if (transcripts && transcripts[b] && transcript_lens) transcript_lens[b] = 0;
        return;
      } else if (out[b].proof_len > proof_caps[b]) {
        rc = -2;
      } else if (!out[b].direct) {
        std::memcpy(proofs[b], out[b].proof.data(), out[b].proof.size() * 4);
      }
      if (transcripts && transcripts[b]) {
        if (transcript_lens) transcript_lens[b] = out[b].transcript.size();
        if (!transcript_caps || out[b].transcript.size() > transcript_caps[b]) rc = -2;
        else std::memcpy(transcripts[b], out[b].transcript.data(), out[b].transcript.size() * 4);
      }
    };
    if (copies) lg::parallel_for(batch, finish);
    else
      for (int b = 0; b < batch; ++b) finish(b);
    if (!failed.empty()) {
      // the other instances' proofs are complete (unless a buffer was short)
      g_err = "lg_prove: " + failed;
      return rc.load() ? rc.load() : -3;
    }
    if (rc.load()) g_err = "lg_prove: output buffer too small";
    return rc.load();
  });
}

int lg_prove(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
             uint32_t* const* proofs, const size_t* proof_caps, size_t* proof_lens, uint32_t* const* transcripts,
             const size_t* transcript_caps, size_t* transcript_lens) {
  return prove_impl(plan, batch, columns, perms, false, proofs, proof_caps, proof_lens, transcripts, transcript_caps,
                    transcript_lens);
}

int lg_prove_device(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
                    uint32_t* const* proofs, const size_t* proof_caps, size_t* proof_lens,
                    uint32_t* const* transcripts, const size_t* transcript_caps, size_t* transcript_lens) {
  return prove_impl(plan, batch, columns, perms, true, proofs, proof_caps, proof_lens, transcripts, transcript_caps,
                    transcript_lens);
}

int lg_prove_chained(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
                     int device_inputs, const uint32_t* fs_prev, uint32_t* const* proofs, const size_t* proof_caps,
                     size_t* proof_lens, uint32_t* const* transcripts, const size_t* transcript_caps,
                     size_t* transcript_lens) {
  if (!fs_prev) {
    g_err = "lg_prove_chained: null fs_prev";
    return -1;
  }
  return prove_impl(plan, batch, columns, perms, device_inputs != 0, proofs, proof_caps, proof_lens, transcripts,
                    transcript_caps, transcript_lens, fs_prev);
}

static std::vector<lg::InstanceInput> instance_inputs(lg_plan* plan, int batch, const uint32_t* const* columns,
                                                     const uint32_t* const* perms, bool device) {
  const size_t nc = plan->p->num_inputs(), np = plan->p->num_perms();
  if ((nc && !columns) || (np && !perms)) throw std::runtime_error("missing inputs");
  std::vector<lg::InstanceInput> in(batch);
  for (int b = 0; b < batch; ++b) {
    for (size_t k = 0; k < nc; ++k) {
      if (!columns[b * nc + k]) throw std::runtime_error("null column pointer");
      in[b].columns.push_back(columns[b * nc + k]);
    }
    for (size_t k = 0; k < np; ++k) {
      if (!perms[b * np + k]) throw std::runtime_error("null permutation pointer");
      in[b].perms.push_back(perms[b * np + k]);
    }
    in[b].device = device;
  }
  return in;
}

int lg_plan_autotune(lg_plan* plan, int batch, const uint32_t* const* columns, const uint32_t* const* perms,
                     int device_inputs, int reps, int32_t* tuning_out, double* ms_out) {
  return guarded([&] {
    if (!plan || batch <= 0) throw std::runtime_error("lg_plan_autotune: bad arguments");
    std::vector<lg::InstanceInput> in = instance_inputs(plan, batch, columns, perms, device_inputs != 0);
    std::lock_guard<std::mutex> lock(plan->ctx->mu);
    double ms[2] = {0, 0};
    int trials = 0;
    const lg::Tuning t = plan->p->autotune(in, reps, ms, &trials);
    if (tuning_out) {
      const int32_t v[6] = {t.lanes, t.merkle_wide_min, t.hash_threads, t.fri_coop_pairs, t.claims_rows, trials};
      for (int i = 0; i < 6; ++i) tuning_out[i] = v[i];
    }
    if (ms_out) {
      ms_out[0] = ms[0];
      ms_out[1] = ms[1];
    }
    return 0;
  });
}

int lg_plan_set_options(lg_plan* plan, uint32_t flags) {
  return guarded([&] {
    if (!plan) throw std::runtime_error("lg_plan_set_options: null plan");
    if (flags & ~lg::PROVE_SKIP_FRI) throw std::runtime_error("lg_plan_set_options: unknown option");
    std::lock_guard<std::mutex> lock(plan->ctx->mu);
    plan->p->set_options(flags);
    return 0;
  });
}

int lg_plan_set_grinding(lg_plan* plan, int bits) {
  return guarded([&] {
    if (!plan || bits < 0 || bits > 31) throw std::runtime_error("lg_plan_set_grinding: bad arguments");
    std::lock_guard<std::mutex> lock(plan->ctx->mu);
    plan->p->set_grinding((uint32_t)bits);
    return 0;
  });
}

int lg_plan_set_tuning(lg_plan* plan, int batch, const int32_t* tuning) {
  return guarded([&] {
    if (!plan || batch <= 0 || !tuning) throw std::runtime_error("lg_plan_set_tuning: bad arguments");
    lg::Tuning t;
    t.lanes = tuning[0];
    t.merkle_wide_min = tuning[1];
    t.hash_threads = tuning[2];
    t.fri_coop_pairs = tuning[3];
    t.claims_rows = tuning[4];
    if (t.lanes < 1 || t.lanes > 8 || t.merkle_wide_min < 1 || t.hash_threads < 64 || t.hash_threads > 256 ||
        t.hash_threads % 64 || t.fri_coop_pairs < 1 || t.claims_rows < 256)
      throw std::runtime_error("lg_plan_set_tuning: value out of range");
    std::lock_guard<std::mutex> lock(plan->ctx->mu);
    plan->p->set_tuning(batch, t);
    return 0;
  });
}

int lg_ctx_set_graphs(lg_ctx* ctx, int on) {
  return guarded([&] {
    if (!ctx) throw std::runtime_error("lg_ctx_set_graphs: null ctx");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->c.graphs = on != 0;
    return 0;
  });
}

static int verify_impl(lg_plan* plan, int n, const uint32_t* const* proofs, const size_t* lens,
                       const uint32_t* setup_roots, size_t n_setup_roots, const uint32_t* const* public_values,
                       int skip_fri, const uint32_t* fs_prev, int32_t* status) {
  return guarded([&] {
    if (!plan || n < 0 || (n && (!proofs || !lens || !status)) || (n_setup_roots && !setup_roots))
      throw std::runtime_error("lg_verify: bad arguments");
    const size_t npub = plan->p->data().publics.size();
    if (npub && n && !public_values) throw std::runtime_error("lg_verify: the plan has public inputs");
    std::vector<lg::VerifyInput> in(n);
    for (int i = 0; i < n; ++i) {
      in[i].words = proofs[i];
      in[i].len = lens[i];
      for (size_t j = 0; j < npub; ++j) in[i].publics.push_back(public_values[(size_t)i * npub + j]);
      if (fs_prev) {
        in[i].chained = true;
        std::memcpy(in[i].fs_prev, fs_prev + (size_t)8 * i, 32);
      }
    }
    std::vector<lg::VerifyResult> out;
    std::lock_guard<std::mutex> lock(plan->ctx->mu);
    lg::verify_batch(*plan->p, in, setup_roots, n_setup_roots, skip_fri != 0, out);
    plan->verify_msgs.assign(n, std::string());
    for (int i = 0; i < n; ++i) {
      status[i] = out[i].status;
      plan->verify_msgs[i] = out[i].message;
    }
    return 0;
  });
}

int lg_verify(lg_plan* plan, int n, const uint32_t* const* proofs, const size_t* lens, const uint32_t* setup_roots,
              size_t n_setup_roots, const uint32_t* const* public_values, int skip_fri, int32_t* status) {
  return verify_impl(plan, n, proofs, lens, setup_roots, n_setup_roots, public_values, skip_fri, nullptr, status);
}

int lg_verify_chained(lg_plan* plan, int n, const uint32_t* const* proofs, const size_t* lens,
                      const uint32_t* setup_roots, size_t n_setup_roots, const uint32_t* const* public_values,
                      int skip_fri, const uint32_t* fs_prev, int32_t* status) {
  if (!fs_prev) {
    g_err = "lg_verify_chained: null fs_prev";
    return -1;
  }
  return verify_impl(plan, n, proofs, lens, setup_roots, n_setup_roots, public_values, skip_fri, fs_prev, status);
}

int lg_plan_verify_message(const lg_plan* plan, int i, char* buf, size_t cap, size_t* len) {
  return guarded([&] {
    if (!plan || i < 0 || (size_t)i >= plan->verify_msgs.size() || !len)
      throw std::runtime_error("lg_plan_verify_message: bad arguments");
    const std::string& m = plan->verify_msgs[i];
    *len = m.size() + 1;
    if (buf && cap) {
      const size_t k = std::min(cap - 1, m.size());
      std::memcpy(buf, m.data(), k);
      buf[k] = 0;
    }
    return 0;
  });
}

int lg_plan_graph_stats(const lg_plan* plan, uint64_t* out4) {
  return guarded([&] {
    if (!plan || !out4) throw std::runtime_error("lg_plan_graph_stats: null argument");
    out4[0] = plan->p->graph_bytes();
    out4[1] = plan->p->graph_records.load();
    out4[2] = plan->p->graph_replays.load();
    out4[3] = plan->p->graph_fallbacks.load();
    return 0;
  });
}

int lg_ctx_set_timing(lg_ctx* ctx, int on) {
  return guarded([&] {
    if (!ctx) throw std::runtime_error("lg_ctx_set_timing: null ctx");
    std::lock_guard<std::mutex> lock(ctx->mu);
    ctx->c.timer.on = on != 0;
    return 0;
  });
}

int lg_ctx_timing_report(lg_ctx* ctx, char* buf, size_t cap, size_t* len) {
  return guarded([&] {
    if (!ctx) throw std::runtime_error("lg_ctx_timing_report: null ctx");
    std::lock_guard<std::mutex> lock(ctx->mu);
    std::string out;
    char line[256];
    for (auto& kv : ctx->c.timer.stats) {
      snprintf(line, sizeof line, "%s %llu %.6f %.0f %.0f\n", kv.first.c_str(), (unsigned long long)kv.second.count,
               kv.second.ms, kv.second.bytes, kv.second.perms);
      out += line;
    }
    if (len) *len = out.size() + 1;
    if (buf && cap) {
      size_t k = std::min(cap - 1, out.size());
      std::memcpy(buf, out.data(), k);
      buf[k] = 0;
    }
    ctx->c.timer.stats.clear();
    return 0;
  });
}

int lg_plan_last_phases(const lg_plan* plan, double* out3) {
  return guarded([&] {
    if (!plan || !out3) throw std::runtime_error("lg_plan_last_phases: null argument");
    for (int i = 0; i < 3; ++i) out3[i] = plan->phases[i];
    return 0;
  });
}

int lg_hash_leaf_pairs(lg_ctx* ctx, const uint32_t* const* base_cols, int nb, const uint32_t* const* ext_cols, int ne,
                       size_t rows, uint32_t* digests) {
  return guarded([&] {
    if (!ctx || !digests || rows < 2 || (rows & (rows - 1)) || nb < 0 || ne < 0 || (nb + ne) == 0)
      throw std::runtime_error("lg_hash_leaf_pairs: bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    LG_HIP(hipSetDevice(ctx->c.device));
    std::vector<const uint32_t*> cp;
    for (int i = 0; i < nb; ++i) cp.push_back(base_cols[i]);
    for (int e = 0; e < ne; ++e)
      for (int l = 0; l < 6; ++l) cp.push_back(ext_cols[e] + (size_t)l * rows);
    std::vector<int32_t> msg(3, 0);
    for (int i = 0; i < nb; ++i) msg.push_back(2 * i);
    for (int i = 0; i < nb; ++i) msg.push_back(2 * i + 1);
    for (int i = 0; i < 6 * ne; ++i) msg.push_back(2 * (nb + i));
    for (int i = 0; i < 6 * ne; ++i) msg.push_back(2 * (nb + i) + 1);
    lg::DevBuf cpd = stage_sync(cp.data(), cp.size() * sizeof(void*));
    lg::DevBuf md = stage_sync(msg.data(), msg.size() * 4);
    uint32_t* o = digests;
    lg::DevBuf od = stage_sync(&o, sizeof(void*));
    lg::launch_leaf_hash(cpd.as<const uint32_t*>(), (int)cp.size(), md.as<int32_t>(), (int)msg.size(),
                         lg::kb_to_mont(2 * nb), lg::kb_to_mont(2 * ne), od.as<uint32_t*>(), (int)(rows / 2), 1,
                         ctx->c.stream);
    LG_HIP(hipStreamSynchronize(ctx->c.stream));
    return 0;
  });
}

int lg_merkle_build(lg_ctx* ctx, uint32_t* nodes, size_t n_leaves, const uint32_t* const* inj, const size_t* inj_width,
                    int n_inj) {
  return guarded([&] {
    if (!ctx || !nodes || n_leaves < 1 || (n_leaves & (n_leaves - 1)) || n_inj < 0)
      throw std::runtime_error("lg_merkle_build: bad arguments");
    size_t prev = n_leaves;
    for (int k = 0; k < n_inj; ++k) {
      if (!inj[k] || inj_width[k] == 0 || (inj_width[k] & (inj_width[k] - 1)) || inj_width[k] > prev)
        throw std::runtime_error("lg_merkle_build: injection widths must be non-increasing powers of two <= n_leaves");
      prev = inj_width[k];
    }
    std::lock_guard<std::mutex> lock(ctx->mu);
    LG_HIP(hipSetDevice(ctx->c.device));
    hipStream_t s = ctx->c.stream;
    uint32_t* tp = nodes;
    lg::DevBuf tpd = stage_sync(&tp, sizeof(void*));
    uint32_t* const* tp_d = tpd.as<uint32_t*>();
    std::vector<lg::DevBuf> keep;
    int next = 0;
    auto fold = [&](size_t width) {
      while (next < n_inj && inj_width[next] == width) {
        const uint32_t* ip = inj[next];
        keep.push_back(stage_sync(&ip, sizeof(void*)));
        lg::launch_inject(tp_d, keep.back().as<const uint32_t*>(), (int)width, 1, s);
        ++next;
      }
    };
    fold(n_leaves);
    for (size_t w = n_leaves / 2; w >= 1; w /= 2) {
      lg::launch_merkle_level(tp_d, (int)w, 1, s);
      fold(w);
    }
    LG_HIP(hipStreamSynchronize(s));
    return 0;
  });
}

int lg_grand_product(lg_ctx* ctx, const uint32_t* num, const uint32_t* den, uint32_t* gp, size_t n) {
  return guarded([&] {
    if (!ctx || !num || !den || !gp || n < 1 || (n & (n - 1))) throw std::runtime_error("lg_grand_product: bad arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    LG_HIP(hipSetDevice(ctx->c.device));
    const uint32_t* np_ = num;
    const uint32_t* dp_ = den;
    uint32_t* gp_ = gp;
    lg::DevBuf ratio(6 * n * sizeof(uint32_t));
    uint32_t* rp_ = ratio.as<uint32_t>();
    lg::DevBuf a = stage_sync(&np_, sizeof(void*)), b = stage_sync(&dp_, sizeof(void*)), c = stage_sync(&gp_, sizeof(void*));
    lg::DevBuf r = stage_sync(&rp_, sizeof(void*));
    const size_t tmp_words = lg::scan_tmp_words((int)n, 1);
    lg::DevBuf tmp(tmp_words ? tmp_words * sizeof(uint32_t) : 0);
    lg::launch_grand_product(a.as<const uint32_t*>(), b.as<const uint32_t*>(), r.as<uint32_t*>(), c.as<uint32_t*>(),
                             (int)n, 1, tmp_words ? tmp.as<uint32_t>() : nullptr, ctx->c.stream);
    LG_HIP(hipStreamSynchronize(ctx->c.stream));
    return 0;
  });
}

int lg_probe_issue_rate(lg_ctx* ctx, double* out5) {
  return guarded([&] {
    if (!ctx || !out5) throw std::runtime_error("lg_probe_issue_rate: null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    lg::issue_probe(ctx->c.device, out5, out5 + 4);
    return 0;
  });
}

int lg_grand_product_noalloc(lg_ctx* ctx, const uint32_t* num, const uint32_t* den, uint32_t* gp, size_t n) {
  return guarded([&] {
    if (!ctx || !num || !den || !gp) throw std::runtime_error("lg_grand_product_noalloc: null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    LG_HIP(hipSetDevice(ctx->c.device));
    lg::grand_product_noalloc(ctx->c.device, ctx->c.stream, num, den, gp, n);
    LG_HIP(hipStreamSynchronize(ctx->c.stream));
    return 0;
  });
}

int lg_rs_encode(lg_ctx* ctx, const uint32_t* in, size_t n, int n_cols, uint32_t* out) {
  return guarded([&] {
    if (!ctx || !in || !out || n < 2 || (n & (n - 1)) || n > ((size_t)1 << 22) || n_cols <= 0)
      throw std::runtime_error("lg_rs_encode: bad arguments");
    int logn = 0;
    while (((size_t)1 << logn) < n) ++logn;
    std::lock_guard<std::mutex> lock(ctx->mu);
    LG_HIP(hipSetDevice(ctx->c.device));
    std::vector<lg::NttJob> jobs(n_cols);
    for (int c = 0; c < n_cols; ++c) {
      jobs[c] = lg::NttJob{};
      jobs[c].in = in + (size_t)c * n;
      for (int o = 0; o < 4; ++o) jobs[c].out[o] = out + (size_t)c * 4 * n + (size_t)o * n;
    }
    lg::NttCfg cfg{};
    cfg.logn = logn;
    cfg.inverse_first = 1;
    cfg.n_out = 4;
    for (int o = 0; o < 4; ++o) {
      cfg.pre[o] = ctx->c.tables.lde_pre(logn, o, 2, true);
      cfg.bitrev_out[o] = 1;
    }
    std::vector<lg::DevBuf> held;
    lg::run_ntt(
        ctx->c.tables, cfg, jobs, ctx->c.stream,
        [&](const void* p, size_t bytes) -> const void* {
          held.push_back(stage_sync(p, bytes));
          return held.back().as<void>();
        },
        [&](size_t words) {
          held.emplace_back(words * 4);
          return held.back().as<uint32_t>();
        });
    LG_HIP(hipStreamSynchronize(ctx->c.stream));
    return 0;
  });
}

int lg_server_create(int device, int slots, uint64_t hbm_budget_bytes, int max_batch, lg_server** out) {
  return guarded([&] {
    if (!out) throw std::runtime_error("lg_server_create: null out");
    auto* s = new lg_server();
    try {
      s->s = std::make_unique<lg::Server>(device, slots, (size_t)hbm_budget_bytes, max_batch);
    } catch (...) {
      delete s;
      throw;
    }
    *out = s;
    return 0;
  });
}

void lg_server_destroy(lg_server* s) { delete s; }

int lg_server_submit(lg_server* s, const uint32_t* plan_words, size_t n_words, int batch,
                     const uint32_t* const* columns, const uint32_t* const* perms, uint64_t* ticket) {
  return guarded([&] {
    if (!s || !ticket) throw std::runtime_error("lg_server_submit: null argument");
    *ticket = s->s->submit(plan_words, n_words, batch, columns, perms);
    return 0;
  });
}

int lg_server_wait(lg_server* s, uint64_t ticket, int timeout_ms, uint32_t* const* proofs, const size_t* proof_caps,
                   size_t* proof_lens) {
  return guarded([&] {
    if (!s) throw std::runtime_error("lg_server_wait: null server");
    std::vector<lg::ProveOutput> out;
    std::string err;
    if (!s->s->wait(ticket, timeout_ms, &out, &err, true)) return 1;
    if (!err.empty()) {
      s->s->wait(ticket, 0, nullptr, nullptr, false);
      throw std::runtime_error("request failed: " + err);
    }
    int rc = 0;
    for (size_t b = 0; b < out.size(); ++b) {
      if (proof_lens) proof_lens[b] = out[b].proof.size();
      if (!proofs || !proof_caps || out[b].proof.size() > proof_caps[b]) rc = -2;
    }
    if (rc) {
      g_err = "lg_server_wait: output buffer too small";
      return rc;
    }
    for (size_t b = 0; b < out.size(); ++b) std::memcpy(proofs[b], out[b].proof.data(), out[b].proof.size() * 4);
    s->s->wait(ticket, 0, nullptr, nullptr, false);
    return 0;
  });
}

int lg_server_stats(lg_server* s, uint64_t* out, size_t n) {
  return guarded([&] {
    if (!s || !out) throw std::runtime_error("lg_server_stats: null argument");
    lg::Server::Stats st = s->s->stats();
    const uint64_t v[10] = {st.submitted, st.completed, st.failed, st.queued, st.running,
                            st.peak_running, st.peak_bytes, st.proves, st.merged, (uint64_t)s->s->budget()};
    for (size_t i = 0; i < n && i < 10; ++i) out[i] = v[i];
    return 0;
  });
}

int lg_plan_bytes_per_instance(const uint32_t* words, size_t n_words, uint64_t* bytes) {
  return guarded([&] {
    if (!words || !bytes) throw std::runtime_error("lg_plan_bytes_per_instance: null argument");
    *bytes = lg::plan_bytes_per_instance(lg::parse_plan(words, n_words));
    return 0;
  });
}

}  // extern "C"
