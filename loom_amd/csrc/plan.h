// Parsed prover plan: the program metadata a prove run needs (module sizes,
// column sources, rounds and steps, commitment trees, opening shifts, folded
// constraint DAGs).  The word format is documented in loom_amd/plan.py and in
// include/loomgpu.h.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace lg {

constexpr uint32_t PLAN_MAGIC = 0x31504C4Cu;
constexpr uint32_t PLAN_NONE = 0xFFFFFFFFu;

constexpr uint32_t PLAN_VERSION = 5;

enum ColSrc : uint32_t {
  SRC_INPUT = 0, SRC_LAGRANGE, SRC_ID, SRC_SIGMA, SRC_STEP, SRC_CHALLENGE, SRC_AIR, SRC_RANGE,
  SRC_SETUP,   // a precommitted setup column (a0: its index in the setup order)
  SRC_PUBLIC   // a public-input column (a0: its index in the public list)
};
enum TreeKind : uint32_t { TREE_TRACE = 0, TREE_AIR = 1, TREE_SETUP = 2 };
enum DagOp : uint32_t { OP_COL = 0, OP_CONST, OP_CHAL, OP_ADD, OP_SUB, OP_MUL, OP_POW };
enum StepKind : uint32_t { STEP_GRAND_PRODUCT = 1, STEP_LOOKUP_MULTIPLICITY = 2, STEP_CYCLIC_LOGUP = 3 };
enum FsHook : uint32_t { FS_NONE = 0, FS_BIND_EXPOSED_VALUES = 1 };

struct DagNode {
  uint32_t op, field, a, b;
};

struct Dag {
  std::vector<DagNode> nodes;
  uint32_t root = 0;
};

struct ModuleInfo {
  uint32_t N, degree, n_chunks;
  int log_n;
  Dag relation;
};

struct ColInfo {
  uint32_t field, module, src, a0, a1;
  int limbs() const { return field ? 6 : 1; }
};

struct Perm {
  uint32_t module;
  uint32_t len;  // the permutation values are per-instance inputs
};

struct Step {
  uint32_t kind, module, out_col;  // out_col = outs[0]
  std::vector<uint32_t> outs;
  std::vector<uint32_t> params;    // kind-specific, see loom_amd/plan.py
  std::vector<Dag> ins;
};

struct Exposed {
  uint32_t col;
  std::string name;
};

struct Round {
  uint32_t challenge_col, tree, fs_hook;
  std::vector<Step> steps;
};

struct Group {
  uint32_t module, N;
  std::vector<uint32_t> base, ext;
  std::vector<std::vector<uint32_t>> base_shifts, ext_shifts;
};

struct TreeInfo {
  uint32_t kind;  // TreeKind
  std::vector<Group> groups;
};

// A public-input column (public.Input): its entry rows are fixed by the plan,
// each instance passes the values.
struct PublicCol {
  uint32_t col, module;
  std::string name, module_name;
  std::vector<uint32_t> rows;    // ascending
  std::vector<uint32_t> fields;  // per entry: 0 base (one value word), 1 extension (six)
  // value words an instance passes for this input (entries in row order)
  size_t words() const {
    size_t w = 0;
    for (auto f : fields) w += f ? 6 : 1;
    return w;
  }
};

struct PlanData {
  uint32_t num_queries = 0;
  std::vector<ModuleInfo> modules;
  std::vector<ColInfo> cols;
  std::vector<uint32_t> inputs;
  std::vector<Perm> perms;
  std::vector<Round> rounds;
  std::vector<Exposed> exposed;  // sorted by name
  std::vector<PublicCol> publics;  // sorted by name
  std::vector<TreeInfo> trees;
  uint32_t n_setup = 0;            // setup columns (SRC_SETUP, a0 = 0 .. n_setup-1)
  // logup buses: the exposed totals (indices into `exposed`) that must sum to
  // zero (the verifier's checkLogupBus); the prover does not read them
  std::vector<std::vector<uint32_t>> buses;
};

class WordReader {
 public:
  WordReader(const uint32_t* w, size_t n) : w_(w), n_(n) {}
  uint32_t next() {
    if (pos_ >= n_) throw std::runtime_error("plan: truncated word stream");
    return w_[pos_++];
  }
  size_t pos() const { return pos_; }
  size_t size() const { return n_; }

 private:
  const uint32_t* w_;
  size_t n_;
  size_t pos_ = 0;
};

inline Dag read_dag(WordReader& r, size_t n_cols) {
  Dag d;
  uint32_t n = r.next();
  d.root = r.next();
  if (n == 0 || d.root >= n) throw std::runtime_error("plan: bad dag root");
  d.nodes.resize(n);
  for (uint32_t i = 0; i < n; ++i) {
    DagNode& x = d.nodes[i];
    x.op = r.next(); x.field = r.next(); x.a = r.next(); x.b = r.next();
    switch (x.op) {
      case OP_COL:
      case OP_CHAL:
        if (x.a >= n_cols) throw std::runtime_error("plan: dag leaf column out of range");
        break;
      case OP_CONST:
        break;
      case OP_ADD: case OP_SUB: case OP_MUL:
        if (x.a >= i || x.b >= i) throw std::runtime_error("plan: dag not topological");
        break;
      case OP_POW:
        if (x.a >= i) throw std::runtime_error("plan: dag not topological");
        break;
      default:
        throw std::runtime_error("plan: unknown dag op");
    }
  }
  return d;
}

inline int ilog2(uint64_t n) {
  int k = 0;
  while ((1ull << k) < n) ++k;
  return k;
}

// Every cross reference of a parsed plan, checked before anything indexes
// with it: column sources, permutations, rounds, trees and groups, opening
// shifts and DAG leaves.  A plan that passes addresses only what it declares.
inline void validate_plan(const PlanData& p) {
  auto bad = [](const char* what) { throw std::runtime_error(std::string("plan: ") + what); };
  const size_t nm = p.modules.size(), nc = p.cols.size(), np = p.perms.size();
  const size_t nr = p.rounds.size(), nt = p.trees.size();
  for (auto& m : p.modules) {
    if (m.n_chunks > 8 || (m.n_chunks & (m.n_chunks - 1))) bad("chunk count must be 0 or a power of two <= 8");
  }
  for (auto& pm : p.perms)
    if (pm.len == 0 || pm.len / p.modules[pm.module].N > 64) bad("permutation must hold 1..64 chunks");
  for (auto& c : p.cols) {
    if (c.field > 1) bad("column field must be 0 (base) or 1 (extension)");
    const uint32_t N = p.modules[c.module].N;
    switch (c.src) {
      case SRC_INPUT:
        if (c.field != 0) bad("input columns must be base-field");
        break;
      case SRC_LAGRANGE:
        if (c.field != 0 || c.a0 >= N) bad("Lagrange column position out of range");
        break;
      case SRC_ID:
        if (c.field != 0 || c.a0 >= 64) bad("identity support index out of range");
        break;
      case SRC_SIGMA:
        if (c.field != 0 || c.a1 >= np || p.perms[c.a1].module != c.module || c.a0 >= p.perms[c.a1].len / N)
          bad("sigma column refers to a missing permutation chunk");
        break;
      case SRC_STEP:
        if (c.a0 >= nr || c.a1 >= p.rounds[c.a0].steps.size()) bad("step column refers to a missing step");
        break;
      case SRC_CHALLENGE:
        if (c.field != 1) bad("challenge columns must be extension-field");
        break;
      case SRC_AIR:
        if (c.field != 1 || c.a0 >= p.modules[c.module].n_chunks) bad("AIR chunk column out of range");
        break;
      case SRC_RANGE:
        if (c.field != 0 || c.a0 > N) bad("range table larger than its module");
        break;
      case SRC_SETUP:
        if (c.field > 1 || c.a0 >= p.n_setup) bad("setup column index out of range");
        break;
      case SRC_PUBLIC:
        if (c.field > 1 || c.a0 >= p.publics.size() || p.publics[c.a0].col != (uint32_t)(&c - &p.cols[0]))
          bad("public-input column does not match its record");
        for (auto f : p.publics[c.a0].fields)
          if (f > 1 || (f == 1 && c.field != 1)) bad("public-input entry field (extension entries need an extension column)");
        break;
      default:
        bad("unknown column source");
    }
  }
  for (auto x : p.inputs)
    if (p.cols[x].src != SRC_INPUT) bad("input list names a generated column");
  auto check_dag = [&](const Dag& d) {
    for (auto& x : d.nodes) {
      if (x.field > 1) bad("dag node field must be 0 or 1");
      if (x.op == OP_COL && x.b >= p.modules[p.cols[x.a].module].N) bad("dag leaf shift out of range");
    }
  };
  for (auto& rd : p.rounds) {
    if (rd.challenge_col >= nc || p.cols[rd.challenge_col].src != SRC_CHALLENGE) bad("round challenge column");
    if (rd.tree != PLAN_NONE && (rd.tree >= nt || p.trees[rd.tree].kind != 0)) bad("round tree out of range");
    if (rd.fs_hook > FS_BIND_EXPOSED_VALUES) bad("unknown Fiat-Shamir hook");
    for (auto& st : rd.steps) {
      for (auto o : st.outs)
        if (p.cols[o].module != st.module && st.kind != STEP_LOOKUP_MULTIPLICITY) bad("step output of another module");
      for (auto& d : st.ins) check_dag(d);
    }
  }
  for (auto& m : p.modules) check_dag(m.relation);
  for (size_t i = 0; i < p.publics.size(); ++i) {
    const PublicCol& pc = p.publics[i];
    if (pc.col >= nc || p.cols[pc.col].src != SRC_PUBLIC || p.cols[pc.col].a0 != i) bad("public-input record column");
    if (pc.module >= nm || p.cols[pc.col].module != pc.module) bad("public-input module");
    if (i && !(p.publics[i - 1].name < pc.name)) bad("public inputs not sorted by name");
    for (size_t k = 0; k < pc.rows.size(); ++k)
      if (pc.rows[k] >= p.modules[pc.module].N || (k && pc.rows[k] <= pc.rows[k - 1]))
        bad("public-input rows must ascend within the module");
  }
  std::vector<uint32_t> setup_seen(p.n_setup, 0);
  for (size_t ti = 0; ti < p.trees.size(); ++ti) {
    const TreeInfo& t = p.trees[ti];
    if (t.kind > TREE_SETUP) bad("tree kind must be 0 (trace), 1 (AIR) or 2 (setup)");
    if (t.kind == TREE_SETUP && ti != 0) bad("the setup tree must come first");
    if (t.groups.empty()) bad("tree without groups");
    for (auto& g : t.groups)
      for (int rail = 0; rail < 2; ++rail)
        for (auto c : rail ? g.ext : g.base)
          if (c < nc && ((p.cols[c].src == SRC_SETUP) != (t.kind == TREE_SETUP)))
            bad("setup columns belong to the setup tree only");
    if (t.kind == TREE_SETUP)
      for (auto& g : t.groups)
        for (int rail = 0; rail < 2; ++rail)
          for (auto c : rail ? g.ext : g.base)
            if (c < nc && p.cols[c].a0 < p.n_setup) ++setup_seen[p.cols[c].a0];
  }
  for (auto v : setup_seen)
    if (v != 1) bad("every setup column must be committed once in the setup tree");
  for (auto& t : p.trees) {
    for (auto& g : t.groups) {
      if (g.module >= nm || g.N != p.modules[g.module].N) bad("group module or size mismatch");
      if (g.base.empty() && g.ext.empty()) bad("empty group");
      if (g.base_shifts.size() != g.base.size() || g.ext_shifts.size() != g.ext.size()) bad("shift lists");
      for (int rail = 0; rail < 2; ++rail) {
        const auto& ids = rail ? g.ext : g.base;
        const auto& sh = rail ? g.ext_shifts : g.base_shifts;
        for (size_t i = 0; i < ids.size(); ++i) {
          if (ids[i] >= nc) bad("group column out of range");
          const ColInfo& c = p.cols[ids[i]];
          if (c.module != g.module || c.field != (uint32_t)rail) bad("group column of another module or rail");
          if (c.src == SRC_CHALLENGE) bad("a challenge column cannot be committed");
          if (sh[i].size() > 8) bad("more than 8 opening shifts per polynomial");
          for (auto s : sh[i])
            if (s >= g.N) bad("opening shift out of range");
        }
      }
    }
  }
}

inline PlanData parse_plan(const uint32_t* words, size_t n_words) {
  WordReader r(words, n_words);
  if (r.next() != PLAN_MAGIC) throw std::runtime_error("plan: bad magic");
  if (r.next() != PLAN_VERSION) throw std::runtime_error("plan: unsupported version");
  PlanData p;
  p.num_queries = r.next();
  uint32_t nm = r.next();
  p.modules.resize(nm);
  for (auto& m : p.modules) {
    m.N = r.next(); m.degree = r.next(); m.n_chunks = r.next();
    if (m.N < 2 || (m.N & (m.N - 1)) || m.N > (1u << 22)) throw std::runtime_error("plan: module size must be a power of two in [2, 2^22]");
    m.log_n = ilog2(m.N);
  }
  uint32_t nc = r.next();
  p.cols.resize(nc);
  for (auto& c : p.cols) {
    c.field = r.next(); c.module = r.next(); c.src = r.next(); c.a0 = r.next(); c.a1 = r.next();
    if (c.module >= nm) throw std::runtime_error("plan: column module out of range");
  }
  uint32_t ni = r.next();
  p.inputs.resize(ni);
  for (auto& x : p.inputs) {
    x = r.next();
    if (x >= nc) throw std::runtime_error("plan: input column out of range");
  }
  uint32_t np = r.next();
  p.perms.resize(np);
  for (auto& pm : p.perms) {
    pm.module = r.next();
    pm.len = r.next();
    if (pm.module >= nm || pm.len % p.modules[pm.module].N) throw std::runtime_error("plan: bad permutation shape");
  }
  uint32_t nr = r.next();
  p.rounds.resize(nr);
  for (auto& rd : p.rounds) {
    rd.challenge_col = r.next(); rd.tree = r.next(); rd.fs_hook = r.next();
    uint32_t ns = r.next();
    rd.steps.resize(ns);
    for (auto& st : rd.steps) {
      st.kind = r.next(); st.module = r.next();
      if (st.module >= nm) throw std::runtime_error("plan: step module out of range");
      uint32_t no = r.next();
      if (no == 0 || no > 64) throw std::runtime_error("plan: bad step output count");
      for (uint32_t k = 0; k < no; ++k) {
        st.outs.push_back(r.next());
        if (st.outs.back() >= nc) throw std::runtime_error("plan: step output out of range");
      }
      st.out_col = st.outs[0];
      uint32_t npar = r.next();
      if (npar > 4096) throw std::runtime_error("plan: bad step parameter count");
      for (uint32_t k = 0; k < npar; ++k) st.params.push_back(r.next());
      uint32_t nin = r.next();
      for (uint32_t k = 0; k < nin; ++k) st.ins.push_back(read_dag(r, nc));
    }
  }
  uint32_t nx = r.next();
  p.exposed.resize(nx);
  for (auto& x : p.exposed) {
    x.col = r.next();
    if (x.col >= nc) throw std::runtime_error("plan: exposed column out of range");
    uint32_t len = r.next();
    if (len > 4096) throw std::runtime_error("plan: exposed name too long");
    std::string s;
    for (uint32_t k = 0; k < (len + 3) / 4; ++k) {
      uint32_t wd = r.next();
      for (int b = 0; b < 4; ++b) s.push_back((char)((wd >> (8 * b)) & 0xff));
    }
    s.resize(len);
    x.name = s;
  }
  auto rd_string = [&](size_t max_len) {
    uint32_t len = r.next();
    if (len > max_len) throw std::runtime_error("plan: string too long");
    std::string s;
    for (uint32_t k = 0; k < (len + 3) / 4; ++k) {
      uint32_t wd = r.next();
      for (int b = 0; b < 4; ++b) s.push_back((char)((wd >> (8 * b)) & 0xff));
    }
    s.resize(len);
    return s;
  };
  uint32_t npub = r.next();
  if (npub > 4096) throw std::runtime_error("plan: too many public inputs");
  p.publics.resize(npub);
  for (auto& pc : p.publics) {
    pc.col = r.next();
    pc.module = r.next();
    pc.name = rd_string(4096);
    pc.module_name = rd_string(4096);
    uint32_t k = r.next();
    if (k > (1u << 22)) throw std::runtime_error("plan: too many public-input entries");
    for (uint32_t j = 0; j < k; ++j) pc.rows.push_back(r.next());
    for (uint32_t j = 0; j < k; ++j) pc.fields.push_back(r.next());
  }
  for (auto& c : p.cols) p.n_setup += c.src == SRC_SETUP;
  uint32_t nt = r.next();
  p.trees.resize(nt);
  for (auto& t : p.trees) {
    t.kind = r.next();
    uint32_t ng = r.next();
    t.groups.resize(ng);
    for (auto& g : t.groups) {
      g.module = r.next(); g.N = r.next();
      uint32_t nb = r.next();
      for (uint32_t k = 0; k < nb; ++k) g.base.push_back(r.next());
      uint32_t ne = r.next();
      for (uint32_t k = 0; k < ne; ++k) g.ext.push_back(r.next());
      auto rd_shifts = [&](size_t cnt, std::vector<std::vector<uint32_t>>& out) {
        out.resize(cnt);
        for (auto& v : out) {
          uint32_t k = r.next();
          for (uint32_t j = 0; j < k; ++j) v.push_back(r.next());
        }
      };
      rd_shifts(nb, g.base_shifts);
      rd_shifts(ne, g.ext_shifts);
    }
  }
  for (auto& m : p.modules) m.relation = read_dag(r, nc);
  uint32_t nbus = r.next();
  if (nbus > 4096) throw std::runtime_error("plan: too many buses");
  p.buses.resize(nbus);
  for (auto& bus : p.buses) {
    uint32_t k = r.next();
    if (k == 0 || k > 4096) throw std::runtime_error("plan: bad bus size");
    for (uint32_t j = 0; j < k; ++j) {
      bus.push_back(r.next());
      if (bus.back() >= nx) throw std::runtime_error("plan: bus member out of range");
    }
  }
  if (r.pos() != r.size()) throw std::runtime_error("plan: trailing words");
  validate_plan(p);
  return p;
}

}  // namespace lg
