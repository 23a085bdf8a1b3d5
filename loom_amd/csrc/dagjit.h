// Run-time compilation of folded constraint DAGs into straight-line HIP
// kernels (hiprtc), one kernel per DAG.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plan.h"

namespace lg {

// A DAG lowered to a kernel.  The kernel evaluates the DAG at every point
// t in [0, ncos * n): coset i = t >> logn, row j = t & (n-1).  A column leaf
// with row shift s reads its column at i*n + ((j + s) mod n), i.e. the row
// shift moves along the natural-order evaluations of one coset.  The result
// (E6, Montgomery) times rowscale[i] (when rowscale is non-null) goes to
// out[b] + l*out_stride + t for limb l.
struct DagModule {
  hipModule_t module = nullptr;
  hipFunction_t fn = nullptr;
};

// The compiled code depends only on the DAG's shape; which columns feed its
// slots is per DAG, so two DAGs of one shape share a module but not slots.
struct DagKernel {
  std::shared_ptr<DagModule> code;
  std::vector<uint32_t> col_ids;    // leaf columns, in kernel column-slot order
  std::vector<uint32_t> chal_ids;   // challenge columns, in kernel slot order
  std::string source;
};

// Generates the kernel source for `dag`; col_ids / chal_ids receive the slot
// order.  cols gives every column's storage rail (1 or 6 limbs).
std::string dag_kernel_source(const Dag& dag, const std::vector<ColInfo>& cols, std::vector<uint32_t>& col_ids,
                              std::vector<uint32_t>& chal_ids);

// Compiles (hiprtc, gfx950) and caches by source text.  Needs a current device.
std::shared_ptr<DagKernel> dag_kernel_get(const Dag& dag, const std::vector<ColInfo>& cols);

// hiprtc only (no device needed): returns the code object size, throws with
// the compiler log on failure.  Used by build checks.
size_t dag_kernel_compile_only(const std::string& source);

struct DagLaunch {
  const uint32_t* const* cols;   // [b * n_slots + slot] device pointers
  const uint32_t* chal;          // [b][n_chal][6]
  const uint32_t* rowscale;      // [ncos] or nullptr
  uint32_t* const* out;          // [b]
  uint64_t col_stride;           // limb stride of ext columns
  uint64_t out_stride;           // limb stride of the output
  int logn;
  int ncos;
  int batch;
};

void dag_kernel_launch(const DagKernel& k, const DagLaunch& a, hipStream_t s);

}  // namespace lg
