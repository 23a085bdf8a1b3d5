// The allocation-free grand product: lg_grand_product with its scratch (the
// ratio column, the scan's tile aggregates and the kernels' pointer tables)
// in one file-scope device array instead of device allocations.
//
// Each device holds its own copy of `scratch` at its own address; the host
// never uses the array's name as an address.  noalloc_init (called by
// lg_ctx_create with the device current) looks the address up once per
// device with hipGetSymbolAddress and keeps it; every use below goes through
// that address.  The array keeps the last call's contents: the kernels write
// every word they read (pointer table, ratios, tile aggregates) before
// reading it, so nothing needs clearing.  Calls that overlap on one device
// (two contexts of the same device, or two threads without a shared context)
// share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/loomgpu.h"
#include "gpu.h"
#include "kernels.h"

namespace lg {

namespace {

constexpr size_t PTR_WORDS = 64;  // pointer table: 4 device pointers, padded to 256 bytes
constexpr size_t RATIO_WORDS = (size_t)6 * LG_NOALLOC_MAX_ROWS;
constexpr size_t SCRATCH_WORDS = PTR_WORDS + RATIO_WORDS + (size_t)6 * (LG_NOALLOC_MAX_ROWS / 2048) + 64;

static __device__ __align__(256) uint32_t scratch[SCRATCH_WORDS];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
uint32_t* g_ptr[MAX_DEVICES] = {};

// Writes the four pointers of a grand-product launch (num, den, ratio, gp)
// into the pointer table of the scratch array.
__global__ void k_put_ptrs(const uint32_t** table, const uint32_t* num, const uint32_t* den, const uint32_t* ratio,
                           const uint32_t* gp) {
  table[0] = num;
  table[1] = den;
  table[2] = ratio;
  table[3] = gp;
}

}  // namespace

// Device address of this device's scratch array (the device must be current).
uint32_t* noalloc_init(int device) {
  if (device < 0 || device >= MAX_DEVICES) throw std::runtime_error("noalloc: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      throw std::runtime_error(std::string("noalloc: hipGetSymbolAddress failed: ") + hipGetErrorString(e));
    g_ptr[device] = static_cast<uint32_t*>(p);
  }
  return g_ptr[device];
}

// Same kernels as lg_grand_product (k_ratio, then the single-workgroup or
// tiled prefix product), every scratch word inside `scratch`.
void grand_product_noalloc(int device, hipStream_t s, const uint32_t* num, const uint32_t* den, uint32_t* gp,
                           size_t n) {
  if (n < 1 || (n & (n - 1))) throw std::runtime_error("lg_grand_product_noalloc: n must be a power of two");
  if (n > LG_NOALLOC_MAX_ROWS)
    throw std::runtime_error("lg_grand_product_noalloc: n = " + std::to_string(n) + " exceeds the scratch limit of " +
                             std::to_string((size_t)LG_NOALLOC_MAX_ROWS) + " rows");
  uint32_t* base;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    base = (device >= 0 && device < MAX_DEVICES) ? g_ptr[device] : nullptr;
  }
  if (!base) throw std::runtime_error("lg_grand_product_noalloc: scratch not set up on this device");
  const uint32_t** table = reinterpret_cast<const uint32_t**>(base);
  uint32_t* ratio = base + PTR_WORDS;
  uint32_t* tmp = ratio + RATIO_WORDS;
  if (scan_tmp_words((int)n, 1) > SCRATCH_WORDS - PTR_WORDS - RATIO_WORDS)
    throw std::runtime_error("lg_grand_product_noalloc: scan scratch too small");
  hipLaunchKernelGGL(k_put_ptrs, dim3(1), dim3(1), 0, s, table, num, den, (const uint32_t*)ratio, (const uint32_t*)gp);
  LG_HIP(hipGetLastError());
  launch_grand_product(table + 0, table + 1, const_cast<uint32_t* const*>(table + 2),
                       const_cast<uint32_t* const*>(table + 3), (int)n, 1, scan_tmp_words((int)n, 1) ? tmp : nullptr,
                       s);
}

}  // namespace lg
