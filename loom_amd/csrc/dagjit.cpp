// Lowering of constraint DAGs to HIP source and their compilation with hiprtc.
#include "dagjit.h"

#include <hip/hiprtc.h>

#include <sstream>
#include <stdexcept>

#include "gpu.h"
#include "kb.h"
#include "kernels.h"

namespace lg {

static const char* kKbSource =
#include "kb_src.inc"
    ;

namespace {

struct NodeCode {
  bool ext;
  std::string name;
};

std::string hex32(uint32_t v) {
  char b[16];
  snprintf(b, sizeof b, "0x%08xu", v);
  return b;
}

}  // namespace

std::string dag_kernel_source(const Dag& dag, const std::vector<ColInfo>& cols, std::vector<uint32_t>& col_ids,
                              std::vector<uint32_t>& chal_ids) {
  col_ids.clear();
  chal_ids.clear();
  std::map<uint32_t, int> col_slot, chal_slot;
  std::ostringstream body;
  std::vector<NodeCode> nc(dag.nodes.size());
  // Only nodes reachable from the root are emitted.
  std::vector<char> live(dag.nodes.size(), 0);
  live[dag.root] = 1;
  for (size_t k = dag.nodes.size(); k-- > 0;) {
    if (!live[k]) continue;
    const DagNode& x = dag.nodes[k];
    if (x.op == OP_ADD || x.op == OP_SUB || x.op == OP_MUL) live[x.a] = live[x.b] = 1;
    if (x.op == OP_POW) live[x.a] = 1;
  }
  for (size_t k = 0; k < dag.nodes.size(); ++k) {
    if (!live[k]) continue;
    const DagNode& x = dag.nodes[k];
    std::string v = "v" + std::to_string(k);
    nc[k].name = v;
    switch (x.op) {
      case OP_COL: {
        auto it = col_slot.find(x.a);
        int slot;
        if (it == col_slot.end()) {
          slot = (int)col_ids.size();
          col_slot[x.a] = slot;
          col_ids.push_back(x.a);
        } else {
          slot = it->second;
        }
        const bool ext = cols[x.a].field != 0;
        nc[k].ext = ext;
        std::string idx = "(ci + ((j + " + std::to_string(x.b) + "u) & nm1))";
        std::string ptr = "cp[" + std::to_string(slot) + "]";
        if (ext)
          body << "  E6 " << v << " = ld6(" << ptr << ", cs, " << idx << ");\n";
        else
          body << "  uint32_t " << v << " = " << ptr << "[" << idx << "];\n";
        break;
      }
      case OP_CONST:
        nc[k].ext = false;
        body << "  uint32_t " << v << " = " << hex32(kb_to_mont(x.a)) << ";\n";
        break;
      case OP_CHAL: {
        auto it = chal_slot.find(x.a);
        int slot;
        if (it == chal_slot.end()) {
          slot = (int)chal_ids.size();
          chal_slot[x.a] = slot;
          chal_ids.push_back(x.a);
        } else {
          slot = it->second;
        }
        nc[k].ext = true;
        body << "  E6 " << v << " = ldc6(ch + 6 * " << slot << ");\n";
        break;
      }
      case OP_ADD:
      case OP_SUB:
      case OP_MUL: {
        const NodeCode& a = nc[x.a];
        const NodeCode& b = nc[x.b];
        const bool ext = a.ext || b.ext;
        nc[k].ext = ext;
        if (!ext) {
          const char* f = x.op == OP_ADD ? "kb_add" : x.op == OP_SUB ? "kb_sub" : "kb_mul";
          body << "  uint32_t " << v << " = " << f << "(" << a.name << ", " << b.name << ");\n";
        } else if (x.op == OP_MUL) {
          if (a.ext && b.ext) body << "  E6 " << v << " = e6_mul(" << a.name << ", " << b.name << ");\n";
          else if (a.ext) body << "  E6 " << v << " = e6_scale(" << a.name << ", " << b.name << ");\n";
          else body << "  E6 " << v << " = e6_scale(" << b.name << ", " << a.name << ");\n";
        } else {
          std::string an = a.ext ? a.name : "e6_from_base(" + a.name + ")";
          std::string bn = b.ext ? b.name : "e6_from_base(" + b.name + ")";
          body << "  E6 " << v << " = " << (x.op == OP_ADD ? "e6_add" : "e6_sub") << "(" << an << ", " << bn
               << ");\n";
        }
        break;
      }
      case OP_POW: {
        const NodeCode& a = nc[x.a];
        nc[k].ext = a.ext;
        const char* ty = a.ext ? "E6" : "uint32_t";
        const char* mul = a.ext ? "e6_mul" : "kb_mul";
        uint32_t e = x.b;
        if (e == 0) {
          body << "  " << ty << " " << v << " = " << (a.ext ? "e6_one()" : hex32(KB_R)) << ";\n";
          break;
        }
        // left-to-right square and multiply
        int top = 31;
        while (!((e >> top) & 1)) --top;
        body << "  " << ty << " " << v << " = " << a.name << ";\n";
        for (int bit = top - 1; bit >= 0; --bit) {
          body << "  " << v << " = " << mul << "(" << v << ", " << v << ");\n";
          if ((e >> bit) & 1) body << "  " << v << " = " << mul << "(" << v << ", " << a.name << ");\n";
        }
        break;
      }
      default:
        throw std::runtime_error("dag codegen: unknown op");
    }
  }
  const NodeCode& r = nc[dag.root];
  std::ostringstream src;
  src << "#define LG_HIPRTC 1\n"
         "typedef unsigned int uint32_t;\n"
         "typedef unsigned long long uint64_t;\n"
         "typedef unsigned long size_t;\n";
  src << kKbSource << "\n";
  src << "using namespace lg;\n"
         "__device__ __forceinline__ E6 ld6(const uint32_t* p, uint64_t st, uint64_t i) {\n"
         "  E6 r;\n"
         "  for (int l = 0; l < 6; ++l) r.v[l] = p[l * st + i];\n"
         "  return r;\n"
         "}\n"
         "__device__ __forceinline__ E6 ldc6(const uint32_t* p) {\n"
         "  E6 r;\n"
         "  for (int l = 0; l < 6; ++l) r.v[l] = p[l];\n"
         "  return r;\n"
         "}\n"
         "extern \"C\" __global__ __launch_bounds__(256) void lg_dag(const uint32_t* const* cols, int ncols,\n"
         "    const uint32_t* chal, int nchal, const uint32_t* rowscale, uint32_t* const* out, uint64_t cs,\n"
         "    uint64_t os, int logn, int ncos) {\n"
         "  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;\n"
         "  const uint32_t n = 1u << logn;\n"
         "  if (t >= (uint32_t)ncos * n) return;\n"
         "  const int b = blockIdx.y;\n"
         "  const uint32_t* const* cp = cols + (uint64_t)b * ncols;\n"
         "  const uint32_t* ch = chal + (uint64_t)b * nchal * 6;\n"
         "  const uint32_t nm1 = n - 1u;\n"
         "  const uint32_t i = t >> logn;\n"
         "  const uint32_t j = t & nm1;\n"
         "  const uint32_t ci = i << logn;\n"
         "  (void)cp; (void)ch; (void)j; (void)ci;\n";
  src << body.str();
  src << "  E6 res = " << (r.ext ? r.name : "e6_from_base(" + r.name + ")") << ";\n"
      << "  if (rowscale) res = e6_scale(res, rowscale[i]);\n"
         "  uint32_t* o = out[b];\n"
         "  for (int l = 0; l < 6; ++l) o[l * os + t] = res.v[l];\n"
         "}\n";
  return src.str();
}

static std::vector<char> hiprtc_compile(const std::string& source) {
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, source.c_str(), "lg_dag.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw std::runtime_error("hiprtcCreateProgram failed");
  const char* opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17",
                        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult rc = hiprtcCompileProgram(prog, (int)(sizeof opts / sizeof opts[0]), opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw std::runtime_error("hiprtc compile failed: " + log);
  }
  size_t cs = 0;
  hiprtcGetCodeSize(prog, &cs);
  std::vector<char> code(cs);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  return code;
}

size_t dag_kernel_compile_only(const std::string& source) { return hiprtc_compile(source).size(); }

std::shared_ptr<DagKernel> dag_kernel_get(const Dag& dag, const std::vector<ColInfo>& cols) {
  static std::mutex mu;
  static std::map<std::pair<int, std::string>, std::shared_ptr<DagModule>> cache;
  auto k = std::make_shared<DagKernel>();
  k->source = dag_kernel_source(dag, cols, k->col_ids, k->chal_ids);
  int dev = 0;
  LG_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  auto key = std::make_pair(dev, k->source);
  auto it = cache.find(key);
  if (it != cache.end()) {
    k->code = it->second;
    return k;
  }
  std::vector<char> code = hiprtc_compile(k->source);
  auto m = std::make_shared<DagModule>();
  LG_HIP(hipModuleLoadData(&m->module, code.data()));
  LG_HIP(hipModuleGetFunction(&m->fn, m->module, "lg_dag"));
  cache[key] = m;
  k->code = m;
  return k;
}

void dag_kernel_launch(const DagKernel& k, const DagLaunch& a, hipStream_t s) {
  int ncols = (int)k.col_ids.size();
  int nchal = (int)k.chal_ids.size();
  uint64_t cs = a.col_stride, os = a.out_stride;
  int logn = a.logn, ncos = a.ncos;
  const void* cols = a.cols;
  const void* chal = a.chal;
  const void* rowscale = a.rowscale;
  const void* out = a.out;
  void* args[] = {&cols, &ncols, &chal, &nchal, &rowscale, &out, &cs, &os, &logn, &ncos};
  const uint64_t total = (uint64_t)ncos << logn;
  const unsigned T = 256;
  const unsigned gx = (unsigned)((total + T - 1) / T);
  if (!prelaunch()) return;
  LG_HIP(hipModuleLaunchKernel(k.code->fn, gx, a.batch, 1, T, 1, 1, 0, s, args, nullptr));
}

}  // namespace lg
