// Fixed HIP kernels of the prove path (gfx950).  Every field value is a
// KoalaBear element in Montgomery form; E6 columns are stored as 6 limb
// arrays (structure of arrays) so that each limb streams like a base column.
#include "kernels.h"

#include <stdexcept>

#include "gpu.h"
#include "ntt16.h"
#include "p2.h"

namespace lg {

static thread_local PreLaunchHook g_hook = nullptr;
static thread_local void* g_hook_user = nullptr;

void set_prelaunch_hook(PreLaunchHook fn, void* user) {
  g_hook = fn;
  g_hook_user = user;
}

static thread_local bool g_suppress = false;

void set_launch_suppression(bool on) { g_suppress = on; }
bool launch_suppressed() { return g_suppress; }

bool prelaunch() {
  if (g_suppress) return false;
  if (g_hook) g_hook(g_hook_user);
  return true;
}

__device__ __forceinline__ uint32_t brev(uint32_t x, int bits) { return __brev(x) >> (32 - bits); }

// A pointer read from a job table is a generic pointer to the compiler, and
// accesses through it become flat_* instructions: they count against the LDS
// counter too, so every LDS wait also waits for them.  Every buffer here is
// device global memory; accessing it through this view gives global_*
// instructions (with the SGPR-base form where the pointer is uniform).
template <class T>
using gmem = __attribute__((address_space(1))) T*;
template <class T>
__device__ __forceinline__ gmem<T> gbl(T* p) {
  return (gmem<T>)p;
}
// Element i of a global buffer addressed by a 32-bit byte offset (i * sizeof
// T < 2^32): with a uniform base this is the SGPR-base form, one VGPR per
// address instead of a 64-bit pair.
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T& gat(gmem<T> base, uint32_t i) {
  using C = __attribute__((address_space(1))) char*;
  return *(gmem<T>)((C)base + i * (uint32_t)sizeof(T));
}
// A uniform pointer made opaque in an SGPR pair: the compiler cannot
// re-associate base + k * stride with the per-thread offset, so an access
// gat(sreg(base + k * stride), t) stays a global access with that SGPR base
// and the one shared VGPR offset (instead of a 64-bit VGPR address per k).
template <class T>
__device__ __forceinline__ gmem<T> sreg(gmem<T> p) {
  asm("" : "+s"(p));
  return p;
}
// 16-byte stores through gat (a builtin vector type: HIP's uint4 has no
// assignment for an object in the global address space)
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------ NTT
// Radix-2 decimation-in-time on LDS data already in bit-reversed order;
// leaves natural order.  tw[k] = w_n^k for k < n/2.
__device__ __forceinline__ void lds_dit(uint32_t* sh, int logn, const uint32_t* __restrict__ tw) {
  const int n = 1 << logn;
  const int T = blockDim.x;
  for (int s = 0; s < logn; ++s) {
    const int half = 1 << s;
    const int tsh = logn - 1 - s;
    for (int t = threadIdx.x; t < (n >> 1); t += T) {
      int j = t & (half - 1);
      int i0 = ((t >> s) << (s + 1)) | j;
      int i1 = i0 + half;
      uint32_t w = tw[j << tsh];
      uint32_t u = sh[i0];
      uint32_t v = kb_mul(sh[i1], w);
      sh[i0] = kb_add(u, v);
      sh[i1] = kb_sub(u, v);
    }
    __syncthreads();
  }
}

constexpr int NTT_MAX_PER_THREAD = 16;

struct NttArgs {
  int logn, inverse_first, n_out, bitrev_mask;
};

// KIND only names the phase (one kernel symbol per phase, so traces separate
// them); the code is the same.
template <int KIND>
__global__ __launch_bounds__(1024) void k_ntt_multi(const NttJob* __restrict__ jobs, NttArgs cfg,
                                                    const uint32_t* const* __restrict__ pres,
                                                    const uint32_t* __restrict__ tw_fwd,
                                                    const uint32_t* __restrict__ tw_inv) {
  extern __shared__ uint32_t sh[];
  const NttJob* __restrict__ jp = jobs + blockIdx.x;
  const uint32_t* __restrict__ jin = jp->in;
  const uint32_t* __restrict__ jpost = jp->post;
  const int logn = cfg.logn;
  const int n = 1 << logn;
  const int T = blockDim.x;
  const int per = n / T;
  const int tid = threadIdx.x;
  uint32_t reg[NTT_MAX_PER_THREAD];
#pragma unroll
  for (int r = 0; r < NTT_MAX_PER_THREAD; ++r)
    if (r < per) reg[r] = jin[tid + r * T];
  if (cfg.inverse_first) {
#pragma unroll
    for (int r = 0; r < NTT_MAX_PER_THREAD; ++r)
      if (r < per) sh[brev(tid + r * T, logn)] = reg[r];
    __syncthreads();
    lds_dit(sh, logn, tw_inv);
#pragma unroll
    for (int r = 0; r < NTT_MAX_PER_THREAD; ++r)
      if (r < per) {
        int k = tid + r * T;
        uint32_t v = sh[k];
        if (jpost) v = kb_mul(v, jpost[k]);
        reg[r] = v;
      }
    __syncthreads();
    if (cfg.n_out == 0) {
#pragma unroll
      for (int r = 0; r < NTT_MAX_PER_THREAD; ++r)
        if (r < per) jp->out[0][tid + r * T] = reg[r];
      return;
    }
  }
  for (int o = 0; o < cfg.n_out; ++o) {
    const uint32_t* pre = pres[o];
#pragma unroll
    for (int r = 0; r < NTT_MAX_PER_THREAD; ++r)
      if (r < per) {
        int k = tid + r * T;
        uint32_t v = reg[r];
        if (pre) v = kb_mul(v, pre[k]);
        sh[brev(k, logn)] = v;
      }
    __syncthreads();
    lds_dit(sh, logn, tw_fwd);
    uint32_t* out = jp->out[o];
    const bool br = (cfg.bitrev_mask >> o) & 1;
#pragma unroll
    for (int r = 0; r < NTT_MAX_PER_THREAD; ++r)
      if (r < per) {
        int k = tid + r * T;
        out[k] = sh[br ? brev(k, logn) : k];
      }
    __syncthreads();
  }
}

// Radix-16 form of k_ntt_multi for 2^8 <= n <= 2^14 (same job and argument
// semantics): n/16 threads, 16 values per thread, four radix-2 stages per
// pass in registers (ntt16.h), one LDS exchange and barrier per pass instead
// of per stage, twiddles from one table load per pass.  The inverse runs
// decimation-in-time from a bit-reversed LDS copy and ends with the values of
// thread t at natural positions t + k n/16, exactly the layout the first
// forward (decimation-in-frequency) pass reads, so the two transforms chain
// in registers.  Forward outputs come out in bit-reversed order, 16
// contiguous values per thread; a natural-order output goes through LDS.
// The pass twiddles of sub-size 2^logm at b (= t mod m/16), from the table
// the host builds with ntt16_ladder: four 16-byte loads instead of 14 field
// products per pass.
__device__ __forceinline__ Ladder load_ladder(const uint32_t* __restrict__ lad, int logm, uint32_t b) {
  // uniform base of the size's block, per-thread 32-bit offset: global loads
  // with an SGPR base and one VGPR offset
  const gmem<const u32x4> p = (gmem<const u32x4>)gbl(lad + 16 * ((1u << (logm - 4)) - 1u));
  const u32x4 v0 = gat(p, 4 * b), v1 = gat(p, 4 * b + 1), v2 = gat(p, 4 * b + 2), v3 = gat(p, 4 * b + 3);
  Ladder L;
  L.a[0] = v0.x; L.a[1] = v0.y; L.a[2] = v0.z; L.a[3] = v0.w;
  L.a[4] = v1.x; L.a[5] = v1.y; L.a[6] = v1.z; L.a[7] = v1.w;
  L.b[0] = v2.x; L.b[1] = v2.y; L.b[2] = v2.z; L.b[3] = v2.w;
  L.c[0] = v3.x; L.c[1] = v3.y;
  L.d = v3.z;
  return L;
}

// 4-bit reversal of a value index (compile-time after unrolling).
__device__ __forceinline__ constexpr uint32_t brev4(int k) {
  return (uint32_t)(((k & 1) << 3) | ((k & 2) << 1) | ((k & 4) >> 1) | ((k & 8) >> 3));
}

// The size L is a template parameter so that every pass unrolls with its
// stride q = 2^lq known: the LDS index of value k of a pass is then
// lds_swz(base) ^ lds_swz(k << lq) (ntt16.h), one XOR with a constant per
// access, and the scattered loads and stores of the bit-reversed copies are
// lds_swz(brev(t) << 4) ^ brev4(k) likewise.
template <int KIND, int L>
__global__ __launch_bounds__(1024) void k_ntt16(const NttJob* __restrict__ jobs, NttArgs cfg,
                                                const uint32_t* const* __restrict__ pres,
                                                const uint32_t* __restrict__ tw_fwd,
                                                const uint32_t* __restrict__ tw_inv,
                                                const uint32_t* __restrict__ lad_fwd,
                                                const uint32_t* __restrict__ lad_inv) {
  static_assert(L >= 8 && L <= 14, "k_ntt16 covers 2^8 .. 2^14");
  extern __shared__ uint32_t sh[];
  const NttJob* __restrict__ jp = jobs + blockIdx.x;
  constexpr uint32_t n = 1u << L;
  constexpr uint32_t q0 = n >> 4;
  constexpr int P = L >> 2, r = L & 3;
  const uint32_t t = threadIdx.x;
  // Global accesses of value k go through the uniform pointer base + k q0
  // (scalar arithmetic) with the one per-thread offset t, so the 16 offsets
  // take no vector registers.
  // LDS index bases: the bit-reversed scatter (value k of thread t at
  // brev(t + k q0) = brev_{L-4}(t) << 4 | brev4(k)) and the contiguous
  // 16-value row of the r-stage pass (16 t + k)
  const uint32_t s_br = lds_swz(brev(t, L - 4) << 4);
  const uint32_t s_row = lds_swz(16u * t);
  uint32_t x[16];
  if (cfg.inverse_first) {
    const NttRoots Ri = ntt16_roots(tw_inv, L);
    const gmem<const uint32_t> __restrict__ in = gbl(jp->in);
#pragma unroll
    for (int k = 0; k < 16; ++k) sh[s_br ^ brev4(k)] = gat(sreg(in + k * q0), t);
    __syncthreads();
    if constexpr (r != 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = sh[s_row ^ (uint32_t)k];
      dit_small<r>(x, Ri);
#pragma unroll
      for (int k = 0; k < 16; ++k) sh[s_row ^ (uint32_t)k] = x[k];
      __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int lq = r + 4 * p;  // pass sub-size 2^(lq + 4)
      const uint32_t sb = lds_swz(ntt16_base(t, lq));
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = sh[sb ^ lds_swz((uint32_t)k << lq)];
      const Ladder Ld = load_ladder(lad_inv, lq + 4, t & ((1u << lq) - 1u));
      dit16(x, Ld);
      if (p < P - 1) {
#pragma unroll
        for (int k = 0; k < 16; ++k) sh[sb ^ lds_swz((uint32_t)k << lq)] = x[k];
        __syncthreads();
      }
    }
    const gmem<const uint32_t> __restrict__ post = gbl(jp->post);
    if (post) {
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = kb_mul(x[k], gat(sreg(post + k * q0), t));
    }
    if (cfg.n_out == 0) {
      const gmem<uint32_t> __restrict__ out = gbl(jp->out[0]);
#pragma unroll
      for (int k = 0; k < 16; ++k) gat(sreg(out + k * q0), t) = x[k];
      return;
    }
    __syncthreads();  // the last inverse pass read LDS; the forward passes write it
  } else {
    const gmem<const uint32_t> __restrict__ in = gbl(jp->in);
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = gat(sreg(in + k * q0), t);
  }
  const NttRoots Rf = ntt16_roots(tw_fwd, L);
  for (int o = 0; o < cfg.n_out; ++o) {
    uint32_t y[16];
    const gmem<const uint32_t> __restrict__ pre = gbl(pres[o]);
    if (pre) {
#pragma unroll
      for (int k = 0; k < 16; ++k) y[k] = kb_mul(x[k], gat(sreg(pre + k * q0), t));
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) y[k] = x[k];
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int lq = L - 4 * p - 4;
      const uint32_t sb = lds_swz(ntt16_base(t, lq));
      if (p > 0) {
#pragma unroll
        for (int k = 0; k < 16; ++k) y[k] = sh[sb ^ lds_swz((uint32_t)k << lq)];
      }
      // the ladders do not depend on the output; the empty asm keeps the
      // compiler from hoisting all passes' ladders (45 registers) out of the
      // output loop, which spills
      uint32_t lb = t & ((1u << lq) - 1u);
      asm volatile("" : "+v"(lb));
      const Ladder Ld = load_ladder(lad_fwd, lq + 4, lb);
      dif16(y, Ld);
      if (p < P - 1 || r) {
#pragma unroll
        for (int k = 0; k < 16; ++k) sh[sb ^ lds_swz((uint32_t)k << lq)] = y[k];
        __syncthreads();
      }
    }
    if constexpr (r != 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) y[k] = sh[s_row ^ (uint32_t)k];
      dif_small<r>(y, Rf);
    }
    // y[k] is now the bit-reversed-order value at position 16 t + k
    const gmem<uint32_t> __restrict__ out = gbl(jp->out[o]);
    if (((cfg.bitrev_mask >> o) & 1) && !(reinterpret_cast<uintptr_t>(jp->out[o]) & 15)) {
      const gmem<u32x4> o4 = (gmem<u32x4>)out;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        gat(o4, 4 * t + (uint32_t)k) = u32x4{y[4 * k], y[4 * k + 1], y[4 * k + 2], y[4 * k + 3]};
    } else if ((cfg.bitrev_mask >> o) & 1) {
#pragma unroll
      for (int k = 0; k < 16; ++k) gat(out, 16 * t + (uint32_t)k) = y[k];
    } else {
      __syncthreads();  // other threads may still read LDS of this output's last pass
#pragma unroll
      for (int k = 0; k < 16; ++k) sh[s_row ^ (uint32_t)k] = y[k];
      __syncthreads();
      // the scatter base again, from an opaque copy of t: hoisted out of the
      // output loop, its 16 indices spill
      uint32_t tt = t;
      asm volatile("" : "+v"(tt));
      const uint32_t s_out = lds_swz(brev(tt, L - 4) << 4);
#pragma unroll
      for (int k = 0; k < 16; ++k) gat(sreg(out + k * q0), t) = sh[s_out ^ brev4(k)];
    }
    __syncthreads();  // before the next output writes LDS
  }
}

template <int KIND>
static void launch_ntt16(int logn, dim3 g, dim3 b, size_t lds, hipStream_t s, const NttJob* jobs, NttArgs a,
                         const uint32_t* const* pres, const uint32_t* tw_fwd, const uint32_t* tw_inv,
                         const uint32_t* lad_fwd, const uint32_t* lad_inv) {
  switch (logn) {
#define LG_NTT16_CASE(LL)                                                                                    \
  case LL:                                                                                                   \
    hipLaunchKernelGGL((k_ntt16<KIND, LL>), g, b, lds, s, jobs, a, pres, tw_fwd, tw_inv, lad_fwd, lad_inv); \
    break;
    LG_NTT16_CASE(8)
    LG_NTT16_CASE(9)
    LG_NTT16_CASE(10)
    LG_NTT16_CASE(11)
    LG_NTT16_CASE(12)
    LG_NTT16_CASE(13)
    LG_NTT16_CASE(14)
#undef LG_NTT16_CASE
    default: throw std::runtime_error("launch_ntt: unsupported radix-16 size");
  }
}

void launch_ntt(const NttJob* jobs_dev, int n_jobs, const NttCfg& cfg, const uint32_t* const* pres_dev,
                const uint32_t* tw_fwd, const uint32_t* tw_inv, const uint32_t* lad_fwd, const uint32_t* lad_inv,
                hipStream_t s) {
  if (n_jobs == 0) return;
  if (cfg.logn > NTT_MAX_LOG_LDS || cfg.logn < 1) throw std::runtime_error("launch_ntt: unsupported size");
  const int n = 1 << cfg.logn;
  int T = n / 2;
  if (T > 1024) T = 1024;
  if (T < 1) T = 1;
  if (n / T > NTT_MAX_PER_THREAD) throw std::runtime_error("launch_ntt: too many values per thread");
  NttArgs a{cfg.logn, cfg.inverse_first, cfg.n_out, 0};
  for (int o = 0; o < cfg.n_out; ++o) a.bitrev_mask |= (cfg.bitrev_out[o] ? 1 : 0) << o;
  if (!prelaunch()) return;
  const size_t lds = n * sizeof(uint32_t);
  if (cfg.logn >= 8) {
    const dim3 g(n_jobs), b(n / 16);
    const int L = cfg.logn;
    switch (cfg.kind) {
      case NTT_KIND_LDE: launch_ntt16<NTT_KIND_LDE>(L, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv, lad_fwd, lad_inv); break;
      case NTT_KIND_QUOT_COSET: launch_ntt16<NTT_KIND_QUOT_COSET>(L, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv, lad_fwd, lad_inv); break;
      case NTT_KIND_QUOT_BACK: launch_ntt16<NTT_KIND_QUOT_BACK>(L, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv, lad_fwd, lad_inv); break;
      case NTT_KIND_CHUNKS: launch_ntt16<NTT_KIND_CHUNKS>(L, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv, lad_fwd, lad_inv); break;
      case NTT_KIND_DEEP_LDE: launch_ntt16<NTT_KIND_DEEP_LDE>(L, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv, lad_fwd, lad_inv); break;
      default: throw std::runtime_error("launch_ntt: unknown kind");
    }
    LG_HIP(hipGetLastError());
    return;
  }
  const dim3 g(n_jobs), b(T);
  switch (cfg.kind) {
    case NTT_KIND_LDE: hipLaunchKernelGGL(k_ntt_multi<NTT_KIND_LDE>, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv); break;
    case NTT_KIND_QUOT_COSET: hipLaunchKernelGGL(k_ntt_multi<NTT_KIND_QUOT_COSET>, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv); break;
    case NTT_KIND_QUOT_BACK: hipLaunchKernelGGL(k_ntt_multi<NTT_KIND_QUOT_BACK>, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv); break;
    case NTT_KIND_CHUNKS: hipLaunchKernelGGL(k_ntt_multi<NTT_KIND_CHUNKS>, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv); break;
    case NTT_KIND_DEEP_LDE: hipLaunchKernelGGL(k_ntt_multi<NTT_KIND_DEEP_LDE>, g, b, lds, s, jobs_dev, a, pres_dev, tw_fwd, tw_inv); break;
    default: throw std::runtime_error("launch_ntt: unknown kind");
  }
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ two-pass NTT
// Size n = n1 * n2 (n1 = 2^a, n2 = 2^b), four-step:
//   pass A: for each j1 < n1, a size-n2 DFT of x[j1 + n1 j2] (over j2), times
//           w_n^{j1 k2}; written transposed to tmp[k2 * n1 + j1];
//   pass B: for each k2 < n2, a size-n1 DFT of tmp[k2 * n1 + .] (over j1),
//           giving X[k2 + n2 k1].
// A workgroup handles C = 8 consecutive j1 (pass A) or k2 (pass B), so global
// accesses come in runs of C words; the C transforms live in LDS.
constexpr int BIG_C = 8;


// Forward radix-16 DIF over one column of size m = 2^lg (lg >= 4) held by
// m/16 threads (index tt); on entry y[k] is the column's natural position
// tt + k m/16, on exit y[k] is its bit-reversed position 16 tt + k.  Column
// position `pos` lives at col[pos * cs] (cs: the column interleave stride).
// Every thread of the block calls it with the same lg (barriers inside).
__device__ __forceinline__ void col_dif(uint32_t* y, uint32_t* col, int cs, int lg, uint32_t tt,
                                        const uint32_t* __restrict__ tw, const NttRoots& R) {
  const int P = lg >> 2, r = lg & 3;
  // a contiguous column (cs == 1, 2^lg >= 256 words) takes the bank swizzle of
  // k_ntt16, whose access patterns it shares; 8 interleaved columns (cs == 8,
  // 2^lg >= 128) flip the low 3 index bits by bits 4-6, which leaves every
  // pattern of the passes conflict-free (checked exhaustively for 2^7..2^11)
  auto at = [&](uint32_t i) -> uint32_t& {
    return col[(cs == 1 ? lds_swz(i) : (cs == 8 ? i ^ ((i >> 4) & 7u) : i)) * cs];
  };
  for (int p = 0; p < P; ++p) {
    const int logm = lg - 4 * p;
    const uint32_t q = 1u << (logm - 4);
    if (p > 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) y[k] = at(ntt16_pos(tt, q, k));
    }
    const Ladder Ld = ntt16_ladder(tw[(tt % q) << (lg - logm)], R);
    dif16(y, Ld);
    if (p < P - 1 || r) {
#pragma unroll
      for (int k = 0; k < 16; ++k) at(ntt16_pos(tt, q, k)) = y[k];
      __syncthreads();
    }
  }
  if (r) {
#pragma unroll
    for (int k = 0; k < 16; ++k) y[k] = at(16 * tt + k);
    dif_small_r(y, r, R);
  }
}

// Column group of a four-step workgroup.  Workgroups go to the 8 XCDs in
// turn (x mod 8), and each XCD has its own L2; a group's 8 adjacent columns
// are 32 bytes of every row, so adjacent groups share cache lines.  Giving
// the workgroups of one XCD a contiguous range of groups keeps those
// neighbours in one L2.  A bijection on [0, G) for G a multiple of 8 (the
// group counts here are 2^(a-3) and 2^(b-3) >= 16).
__device__ __forceinline__ uint32_t xcd_group(uint32_t x, uint32_t G) {
  return (G & 7u) ? x : (x & 7u) * (G >> 3) + (x >> 3);
}

// Pass A: the n2-point transforms of 8 adjacent columns j1 per workgroup
// (radix-16 DIF, columns interleaved in LDS), times w_n^{j1 k2}, transposed
// into tmp[k2 n1 + j1].  tw_a: w_{n2}^k (k < n2/2); the pass-A twiddle
// w_n^e (e = j1 k2 < n) is wlo[e mod 2048] * whi[e / 2048].
__global__ __launch_bounds__(1024) void k_ntt_pass_a(const BigJob* __restrict__ jobs, int a, int b,
                                                     const uint32_t* __restrict__ tw_a,
                                                     const uint32_t* __restrict__ wlo,
                                                     const uint32_t* __restrict__ whi) {
  extern __shared__ uint32_t sh[];
  const BigJob* job = jobs + blockIdx.y;
  const uint32_t n1 = 1u << a, n2 = 1u << b;
  const uint32_t c = threadIdx.x & (BIG_C - 1), tt = threadIdx.x >> 3;
  const uint32_t j1 = xcd_group(blockIdx.x, gridDim.x) * BIG_C + c;
  const uint32_t qc = n2 >> 4;
  const gmem<const uint32_t> __restrict__ in = gbl(job->in);
  const gmem<const uint32_t> __restrict__ pre = gbl(job->pre);
  uint32_t y[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t gi = j1 + n1 * (tt + (uint32_t)k * qc);  // < n <= 2^22
    const uint32_t v = gat(in, gi);
    y[k] = pre ? kb_mul(v, gat(pre, gi)) : v;
  }
  const NttRoots R = ntt16_roots(tw_a, b);
  col_dif(y, sh + c, BIG_C, b, tt, tw_a, R);
  const gmem<uint32_t> __restrict__ tmp = gbl(job->tmp);
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    const uint32_t k2 = brev(16 * tt + k, b);
    const uint32_t e = j1 * k2;
    gat(tmp, k2 * n1 + j1) = kb_mul(y[k], kb_mul(wlo[e & 2047u], whi[e >> 11]));
  }
}

// Pass B: for 8 adjacent k2 per workgroup, the n1-point transform of the row
// tmp[k2 n1 + .] (radix-16 DIF), giving X[k2 + n2 k1] (times post, natural
// index), stored at that natural index or at its bit reversal
// brev_a(k1) + n1 brev_b(k2).
// 60 VGPRs at 8 waves per SIMD (no spills): two 1024-thread workgroups per
// CU (LDS 2 x 64 KB at n1 = 2^11), so one workgroup's LDS phases overlap the
// other's arithmetic; at the default 66 VGPRs only one fits
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_ntt_pass_b(const BigJob* __restrict__ jobs, int a, int b,
                                                     const uint32_t* __restrict__ tw_b, int bitrev_out) {
  extern __shared__ uint32_t sh[];
  const BigJob* job = jobs + blockIdx.y;
  const uint32_t n1 = 1u << a, n2 = 1u << b;
  const uint32_t tpc = n1 >> 4;  // threads per row
  // Bit-reversed output: contiguous rows (a thread's 16 outputs are
  // contiguous).  Natural output: the 8 rows interleaved, row index fastest
  // across threads, so that each store instruction writes 32-byte runs (8
  // adjacent k2) instead of one dword per lane.
  const bool il = !bitrev_out;
  const uint32_t tt = il ? threadIdx.x >> 3 : threadIdx.x % tpc;
  const uint32_t c = il ? threadIdx.x & (BIG_C - 1) : threadIdx.x / tpc;
  const uint32_t k2 = xcd_group(blockIdx.x, gridDim.x) * BIG_C + c;
  const gmem<const uint32_t> __restrict__ tmp = gbl(job->tmp);
  uint32_t y[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) y[k] = gat(tmp, k2 * n1 + tt + (uint32_t)k * tpc);
  const NttRoots R = ntt16_roots(tw_b, a);
  if (il) col_dif(y, sh + c, BIG_C, a, tt, tw_b, R);
  else col_dif(y, sh + c * n1, 1, a, tt, tw_b, R);
  const gmem<uint32_t> __restrict__ out = gbl(job->out);
  const gmem<const uint32_t> __restrict__ post = gbl(job->post);
  const uint32_t k2r = brev(k2, b);
  if (post) {
#pragma unroll
    for (int k = 0; k < 16; ++k) y[k] = kb_mul(y[k], gat(post, k2 + n2 * brev(16 * tt + k, a)));
  }
  if (bitrev_out) {
    // the 16 values of a thread are contiguous there (positions 16 tt + k =
    // brev_a(k1)): four 16-byte stores per thread instead of sixteen strided
    // dword stores
    const uint32_t o = 16 * tt + n1 * k2r;
    if (!(reinterpret_cast<uintptr_t>(job->out) & 15)) {
      const gmem<u32x4> o4 = (gmem<u32x4>)out;
#pragma unroll
      for (int k = 0; k < 4; ++k)
        gat(o4, o / 4 + (uint32_t)k) = u32x4{y[4 * k], y[4 * k + 1], y[4 * k + 2], y[4 * k + 3]};
    } else {
#pragma unroll
      for (int k = 0; k < 16; ++k) gat(out, o + (uint32_t)k) = y[k];
    }
  } else {
#pragma unroll
    for (int k = 0; k < 16; ++k) gat(out, k2 + n2 * brev(16 * tt + k, a)) = y[k];
  }
}

void launch_ntt_pass_a(const BigJob* jobs, int n_jobs, int a, int b, const uint32_t* tw_a, const uint32_t* wlo,
                       const uint32_t* whi, hipStream_t s) {
  if (n_jobs == 0) return;
  if (a < 7 || a > 11 || b < 4 || b > 11) throw std::runtime_error("launch_ntt_pass_a: unsupported size");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_ntt_pass_a, dim3((1 << a) / BIG_C, n_jobs), dim3(BIG_C * ((1 << b) >> 4)),
                     (size_t)BIG_C * (1 << b) * 4, s, jobs, a, b, tw_a, wlo, whi);
  LG_HIP(hipGetLastError());
}

void launch_ntt_pass_b(const BigJob* jobs, int n_jobs, int a, int b, const uint32_t* tw_b, int bitrev_out,
                       hipStream_t s) {
  if (n_jobs == 0) return;
  if (a < 4 || a > 11 || b < 7 || b > 11) throw std::runtime_error("launch_ntt_pass_b: unsupported size");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_ntt_pass_b, dim3((1 << b) / BIG_C, n_jobs), dim3(BIG_C * ((1 << a) >> 4)),
                     (size_t)BIG_C * (1 << a) * 4, s, jobs, a, b, tw_b, bitrev_out);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ fused LDE middle pass
// LDS index of position i of row c (8 interleaved rows): the low 3 position
// bits flipped by bits 4-6 (conflict-free for the pass patterns, as in pass
// A).  The map is linear over GF(2), so a pass's index is a per-thread base
// XOR a constant per value.
__device__ __forceinline__ constexpr uint32_t idx8(uint32_t i) { return i ^ ((i >> 4) & 7u); }

template <int A>
__global__ __launch_bounds__(1024) void k_lde_mid(const LdeMidJob* __restrict__ jobs, LdeMidArgs g) {
  static_assert(A >= 8 && A <= 11, "row sizes 2^8 .. 2^11");
  extern __shared__ uint32_t sh[];
  constexpr uint32_t n1 = 1u << A, q0 = n1 >> 4;
  constexpr int P = A >> 2, R = A & 3;
  const LdeMidJob* job = jobs + blockIdx.y;
  const uint32_t c = threadIdx.x & (BIG_C - 1), tt = threadIdx.x >> 3;
  const uint32_t n2 = 1u << g.b;
  const uint32_t k2 = xcd_group(blockIdx.x, gridDim.x) * BIG_C + c;
  const uint32_t s_row = (idx8(16u * tt) << 3) | c;  // position 16 tt + k: s_row ^ (k << 3)
  uint32_t x[16];
  {
    const gmem<const uint32_t> tmp = gbl(job->tmp);
#pragma unroll
    for (int k = 0; k < 16; ++k) x[k] = gat(sreg(tmp + k * q0), k2 * n1 + tt);
  }
  // inverse DFT over j1 (decimation in frequency): natural positions
  // tt + k q0 in, bit-reversed positions 16 tt + k out, i.e. x[k] =
  // c[k2 + n2 brev_A(16 tt + k)] (times n: 1/n is in the pre-scale)
  {
    const NttRoots Ri = ntt16_roots(g.tw_inv, A);
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int lq = A - 4 * p - 4;
      const uint32_t sb = (idx8(ntt16_base(tt, lq)) << 3) | c;
      if (p > 0) {
#pragma unroll
        for (int k = 0; k < 16; ++k) x[k] = sh[sb ^ (idx8((uint32_t)k << lq) << 3)];
      }
      const Ladder Ld = load_ladder(g.lad_inv, lq + 4, tt & ((1u << lq) - 1u));
      dif16(x, Ld);
      if (p < P - 1 || R) {
#pragma unroll
        for (int k = 0; k < 16; ++k) sh[sb ^ (idx8((uint32_t)k << lq) << 3)] = x[k];
        __syncthreads();
      }
    }
    if constexpr (R != 0) {
#pragma unroll
      for (int k = 0; k < 16; ++k) x[k] = sh[s_row ^ ((uint32_t)k << 3)];
      dif_small<R>(x, Ri);
    }
  }
  // twiddles w_n^(k2 i1) for i1 = tt + k q0: w_n^(k2 tt) times (w_n^(k2 q0))^k
  const uint32_t e0 = k2 * tt, e1 = k2 * q0;  // < n
  const uint32_t tw0 = kb_mul(g.wlo[e0 & 2047u], g.whi[e0 >> 11]);
  const uint32_t stp = kb_mul(g.wlo[e1 & 2047u], g.whi[e1 >> 11]);
  const NttRoots Rf = ntt16_roots(g.tw_fwd, A);
  for (int o = 0; o < g.n_out; ++o) {
    uint32_t y[16];
    {
      // pre-scale of coefficient k2 + n2 k1: blk_o at the value's position
      // (16 contiguous words per thread), the row factor pre_o[k2] later
      const gmem<const u32x4> b4 = (gmem<const u32x4>)gbl(g.blk[o]);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const u32x4 v = gat(b4, 4 * tt + (uint32_t)k);
        y[4 * k] = kb_mul(x[4 * k], v.x);
        y[4 * k + 1] = kb_mul(x[4 * k + 1], v.y);
        y[4 * k + 2] = kb_mul(x[4 * k + 2], v.z);
        y[4 * k + 3] = kb_mul(x[4 * k + 3], v.w);
      }
    }
    __syncthreads();  // the previous LDS reads (inverse or previous output) are done
    // forward DFT over k1 (decimation in time): bit-reversed positions in,
    // natural positions tt + k q0 out
    if constexpr (R != 0) {
      dit_small<R>(y, Rf);
#pragma unroll
      for (int k = 0; k < 16; ++k) sh[s_row ^ ((uint32_t)k << 3)] = y[k];
      __syncthreads();
    }
#pragma unroll
    for (int p = 0; p < P; ++p) {
      const int lq = R + 4 * p;
      const uint32_t sb = (idx8(ntt16_base(tt, lq)) << 3) | c;
      if (p > 0 || R != 0) {
#pragma unroll
        for (int k = 0; k < 16; ++k) y[k] = sh[sb ^ (idx8((uint32_t)k << lq) << 3)];
      }
      // not hoisted out of the output loop (the ladders of all passes would
      // stay live)
      uint32_t lb = tt & ((1u << lq) - 1u);
      asm volatile("" : "+v"(lb));
      const Ladder Ld = load_ladder(g.lad_fwd, lq + 4, lb);
      dit16(y, Ld);
      if (p < P - 1) {
#pragma unroll
        for (int k = 0; k < 16; ++k) sh[sb ^ (idx8((uint32_t)k << lq) << 3)] = y[k];
        __syncthreads();
      }
    }
    // times pre_o[k2] w_n^(k2 i1), transposed: u_o[i1 n2 + k2]
    uint32_t t = kb_mul(tw0, g.pre[o][k2]);
    const gmem<uint32_t> u = gbl(job->u[o]);
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      gat(sreg(u + k * q0 * n2), tt * n2 + k2) = kb_mul(y[k], t);
      t = kb_mul(t, stp);
    }
  }
}

void launch_lde_mid(const LdeMidJob* jobs, int n_jobs, const LdeMidArgs& args, hipStream_t s) {
  if (n_jobs == 0) return;
  if (args.a < 8 || args.a > 11 || args.b < 7 || args.b > args.a || args.n_out < 1 || args.n_out > NTT_MAX_OUT)
    throw std::runtime_error("launch_lde_mid: unsupported shape");
  if (!prelaunch()) return;
  const dim3 grid((1u << args.b) / BIG_C, n_jobs), block(BIG_C * ((1u << args.a) >> 4));
  const size_t lds = (size_t)BIG_C * (1u << args.a) * 4;
  switch (args.a) {
    case 8: hipLaunchKernelGGL(k_lde_mid<8>, grid, block, lds, s, jobs, args); break;
    case 9: hipLaunchKernelGGL(k_lde_mid<9>, grid, block, lds, s, jobs, args); break;
    case 10: hipLaunchKernelGGL(k_lde_mid<10>, grid, block, lds, s, jobs, args); break;
    default: hipLaunchKernelGGL(k_lde_mid<11>, grid, block, lds, s, jobs, args); break;
  }
  LG_HIP(hipGetLastError());
}

void launch_ntt_big(const BigJob* jobs_a, const BigJob* jobs_b, int n_jobs, int logn, const uint32_t* tw_a,
                    const uint32_t* tw_b, const uint32_t* wlo, const uint32_t* whi, int bitrev_out, hipStream_t s) {
  if (n_jobs == 0) return;
  const int a = (logn + 1) / 2, b = logn - a;  // n1 = 2^a >= n2 = 2^b
  if (a > 11 || b < 7 || logn < 15) throw std::runtime_error("launch_ntt_big: unsupported size");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_ntt_pass_a, dim3((1 << a) / BIG_C, n_jobs), dim3(BIG_C * ((1 << b) >> 4)),
                     (size_t)BIG_C * (1 << b) * 4, s, jobs_a, a, b, tw_a, wlo, whi);
  LG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_ntt_pass_b, dim3((1 << b) / BIG_C, n_jobs), dim3(BIG_C * ((1 << a) >> 4)),
                     (size_t)BIG_C * (1 << a) * 4, s, jobs_b, a, b, tw_b, bitrev_out);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ hashing
__global__ __launch_bounds__(256) void k_leaf_hash(const uint32_t* const* __restrict__ colptrs, int ncols,
                                                   const int32_t* __restrict__ msgmap, int msglen, uint32_t hdr1,
                                                   uint32_t hdr2, uint32_t* const* __restrict__ outs, int nleaves) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= nleaves) return;
  const uint32_t* const* cp = colptrs + (size_t)b * ncols;
  uint32_t s[24];
#pragma unroll
  for (int i = 0; i < 24; ++i) s[i] = 0;
  const int row0 = 2 * k;
  for (int off = 0; off < msglen; off += 16) {
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int m = off + i;
      if (m < msglen) {
        uint32_t v;
        if (m == 0) v = TAG_LEAF_MONT;
        else if (m == 1) v = hdr1;
        else if (m == 2) v = hdr2;
        else {
          const int code = msgmap[m];
          v = cp[code >> 1][row0 + (code & 1)];
        }
        s[i] = v;
      }
    }
    p2_permute<24>(s);
  }
  uint32_t* o = outs[b] + (size_t)8 * k;
#pragma unroll
  for (int i = 0; i < 8; ++i) o[i] = s[i];
}

void launch_leaf_hash(const uint32_t* const* colptrs, int ncols, const int32_t* msgmap, int msglen, uint32_t hdr1,
                      uint32_t hdr2, uint32_t* const* outs, int nleaves, int batch, hipStream_t s, int threads) {
  if (nleaves == 0 || batch == 0) return;
  if (threads < 64 || threads > 256 || (threads & 63)) throw std::runtime_error("leaf hash: threads must be 64..256");
  const int T = threads;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_leaf_hash, dim3((nleaves + T - 1) / T, batch), dim3(T), 0, s, colptrs, ncols, msgmap, msglen,
                     hdr1, hdr2, outs, nleaves);
  LG_HIP(hipGetLastError());
}

__device__ __forceinline__ void load8(const uint32_t* p, uint32_t* d) {
  const uint4* q = reinterpret_cast<const uint4*>(p);
  uint4 a = q[0], c = q[1];
  d[0] = a.x; d[1] = a.y; d[2] = a.z; d[3] = a.w; d[4] = c.x; d[5] = c.y; d[6] = c.z; d[7] = c.w;
}

__device__ __forceinline__ void store8(uint32_t* p, const uint32_t* d) {
  uint4* q = reinterpret_cast<uint4*>(p);
  q[0] = make_uint4(d[0], d[1], d[2], d[3]);
  q[1] = make_uint4(d[4], d[5], d[6], d[7]);
}

__global__ __launch_bounds__(256) void k_merkle_level(uint32_t* const* __restrict__ trees, int width) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= width) return;
  uint32_t* nodes = trees[blockIdx.y];
  const size_t node = (size_t)width + i;
  uint32_t l[8], r[8], d[8];
  load8(nodes + 16 * node, l);
  load8(nodes + 16 * node + 8, r);
  p2_node(l, r, d);
  store8(nodes + 8 * node, d);
}

void launch_merkle_level(uint32_t* const* trees, int width, int batch, hipStream_t s, int threads) {
  if (threads < 64 || threads > 256 || (threads & 63)) throw std::runtime_error("merkle level: threads must be 64..256");
  const int T = width < threads ? width : threads;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_merkle_level, dim3((width + T - 1) / T, batch), dim3(T), 0, s, trees, width);
  LG_HIP(hipGetLastError());
}

// Each workgroup owns G consecutive nodes of level width_hi and reduces them
// to one node of level width_hi / G, writing every level it computes.
__global__ __launch_bounds__(256) void k_merkle_fused(uint32_t* const* __restrict__ trees, int width_hi, int levels) {
  __shared__ uint32_t sh[256 * 8];
  uint32_t* nodes = trees[blockIdx.y];
  const int G = blockDim.x;
  const int tid = threadIdx.x;
  const int base = blockIdx.x * G;
  uint32_t l[8], r[8], d[8];
  {
    const size_t node = (size_t)width_hi + base + tid;
    load8(nodes + 16 * node, l);
    load8(nodes + 16 * node + 8, r);
    p2_node(l, r, d);
    store8(nodes + 8 * node, d);
#pragma unroll
    for (int k = 0; k < 8; ++k) sh[8 * tid + k] = d[k];
  }
  int cnt = G;
  for (int lev = 1; lev < levels; ++lev) {
    __syncthreads();
    cnt >>= 1;
    const int w = width_hi >> lev;
    if (tid < cnt) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        l[k] = sh[16 * tid + k];
        r[k] = sh[16 * tid + 8 + k];
      }
      p2_node(l, r, d);
      store8(nodes + 8 * ((size_t)w + (base >> lev) + tid), d);
    }
    __syncthreads();
    if (tid < cnt) {
#pragma unroll
      for (int k = 0; k < 8; ++k) sh[8 * tid + k] = d[k];
    }
  }
}

// Prefix words of a Fiat-Shamir challenge message, passed by value.
constexpr int CHAL_MAX_PREFIX = 32;
struct ChalPrefix {
  uint32_t w[CHAL_MAX_PREFIX];
};

// Sponge of one FRI fold challenge on a 16-lane row (see k_fri_challenge):
// prefix || prev (8 words) || root (8 words); returns the digest word of
// lane l < 8 in a (the state elements l and l + 12 in a, c).
__device__ __forceinline__ uint32_t fri_challenge_row(const P2Coop16& P, int l, const ChalPrefix& pre, int n_pre,
                                                      const uint32_t* prev, const uint32_t* root) {
  auto msg = [&](int i) -> uint32_t {
    return i < n_pre ? pre.w[i] : (i < n_pre + 8 ? prev[i - n_pre] : root[i - n_pre - 8]);
  };
  uint32_t a = 0u, c = 0u;
  const int len = n_pre + 16;
  for (int off = 0; off < len; off += 16) {
    const int k = len - off < 16 ? len - off : 16;
    if (l < 12 && l < k) a = msg(off + l);
    if (l < 12 && l + 12 < k) c = msg(off + l + 12);
    P.permute(a, c);
  }
  return a;
}

// Cooperative variant: a row of 16 lanes per node (P2Coop16), up to 32 rows
// per workgroup; a workgroup owns G <= 32 consecutive nodes of level width_hi
// and reduces them to one node, keeping the running level in LDS.  Node state:
// element 0 = NODE, 1..7 = 0, 8..15 = left child, 16..23 = right child; lane l
// holds elements l and l + 12.  With CHAL (one workgroup per instance, last
// level = the root) the first row then derives the FRI fold challenge of the
// root, as k_fri_challenge does.
struct FriChalArgs {
  ChalPrefix pre;
  int n_pre;
  const uint32_t* prev;
  uint32_t* dig;
  uint32_t* alpha;
  uint32_t* roots;
};

template <bool CHAL>
__global__ __launch_bounds__(512) void k_merkle_coop(uint32_t* const* __restrict__ trees, int width_hi, int levels,
                                                     FriChalArgs ch) {
  __shared__ uint32_t sh[2][32 * 8];
  const int bi = blockIdx.y;
  uint32_t* nodes = trees[bi];
  const int l = threadIdx.x & 15;
  const int grp = threadIdx.x >> 4;
  const int G = width_hi < 32 ? width_hi : 32;
  const int base = blockIdx.x * G;
  P2Coop16 P;
  P.init(l);
  int cnt = G;
  int cur = 0;
  for (int lev = 0; lev < levels; ++lev) {
    const int w = width_hi >> lev;
    if (grp < cnt) {
      const size_t node = (size_t)w + (base >> lev) + grp;
      // word cw (0..15) of the two children: child cw >> 3, word cw & 7
      auto child = [&](int cw) -> uint32_t {
        return lev == 0 ? nodes[8 * (2 * node + (cw >> 3)) + (cw & 7)] : sh[cur][8 * (2 * grp + (cw >> 3)) + (cw & 7)];
      };
      uint32_t a = 0u, b = 0u;
      if (l == 0) a = TAG_NODE_MONT;
      else if (l >= 8 && l < 12) a = child(l - 8);
      if (l < 12) b = child(l + 4);
      P.permute(a, b);
      if (l < 8) {
        nodes[8 * node + l] = a;
        sh[cur ^ 1][8 * grp + l] = a;
      }
    }
    __syncthreads();
    cur ^= 1;
    cnt >>= 1;
  }
  if constexpr (CHAL) {
    if (grp == 0) {
      const uint32_t* root = sh[cur];
      const uint32_t d = fri_challenge_row(P, l, ch.pre, ch.n_pre, ch.prev + 8 * bi, root);
      if (l < 8) {
        ch.dig[8 * bi + l] = d;
        ch.roots[8 * bi + l] = root[l];
        if (l < 6) ch.alpha[6 * bi + l] = d;
      }
    }
  }
}

void launch_merkle_coop(uint32_t* const* trees, int width_hi, int width_lo, int batch, hipStream_t s) {
  int levels = 1;
  while ((width_hi >> (levels - 1)) > width_lo) ++levels;
  const int G = width_hi < 32 ? width_hi : 32;
  if ((width_hi / width_lo) > G) throw std::runtime_error("merkle coop: span exceeds workgroup");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_merkle_coop<false>, dim3(width_hi / G, batch), dim3(16 * G), 0, s, trees, width_hi, levels,
                     FriChalArgs{});
  LG_HIP(hipGetLastError());
}

void launch_merkle_coop_challenge(uint32_t* const* trees, int width_hi, const uint32_t* prefix, int n_prefix,
                                  const uint32_t* prev, uint32_t* dig, uint32_t* alpha, uint32_t* roots, int batch,
                                  hipStream_t s) {
  if (width_hi < 1 || width_hi > 32 || (width_hi & (width_hi - 1)))
    throw std::runtime_error("merkle coop challenge: the top launch must hold the whole remaining tree");
  if (n_prefix < 1 || n_prefix > CHAL_MAX_PREFIX) throw std::runtime_error("fri challenge: bad prefix length");
  int levels = 1;
  while ((width_hi >> (levels - 1)) > 1) ++levels;
  FriChalArgs ch{};
  for (int i = 0; i < n_prefix; ++i) ch.pre.w[i] = prefix[i];
  ch.n_pre = n_prefix;
  ch.prev = prev;
  ch.dig = dig;
  ch.alpha = alpha;
  ch.roots = roots;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_merkle_coop<true>, dim3(1, batch), dim3(16 * width_hi), 0, s, trees, width_hi, levels, ch);
  LG_HIP(hipGetLastError());
}

void launch_merkle_levels_fused(uint32_t* const* trees, int width_hi, int width_lo, int batch, hipStream_t s) {
  int levels = 1;
  while ((width_hi >> (levels - 1)) > width_lo) ++levels;
  const int G = width_hi < 256 ? width_hi : 256;
  if ((width_hi / width_lo) > G) throw std::runtime_error("merkle fused: span exceeds workgroup");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_merkle_fused, dim3(width_hi / G, batch), dim3(G), 0, s, trees, width_hi, levels);
  LG_HIP(hipGetLastError());
}

__global__ void k_copy_roots(const uint32_t* const* trees, uint32_t* roots) {
  const int b = blockIdx.x;
  if (threadIdx.x < 8) roots[8 * b + threadIdx.x] = trees[b][8 + threadIdx.x];
}

void launch_copy_roots(const uint32_t* const* trees, uint32_t* roots, int batch, hipStream_t s) {
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_copy_roots, dim3(batch), dim3(64), 0, s, trees, roots);
  LG_HIP(hipGetLastError());
}

__global__ __launch_bounds__(256) void k_inject(uint32_t* const* __restrict__ trees,
                                                const uint32_t* const* __restrict__ inj, int width) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= width) return;
  uint32_t* nodes = trees[blockIdx.y];
  const uint32_t* in = inj[blockIdx.y];
  uint32_t l[8], r[8], d[8];
  load8(nodes + 8 * ((size_t)width + j), l);
  load8(in + 8 * (size_t)j, r);
  p2_node(l, r, d);
  store8(nodes + 8 * ((size_t)width + j), d);
}

// count injections at one level width: node j of tree b becomes
// H(...H(H(node, inj[0][b][j]), inj[1][b][j])..., inj[count-1][b][j]) (the
// schedule order of the level's injections); inj is [k * batch + b].
__global__ __launch_bounds__(256) void k_inject_seq(uint32_t* const* __restrict__ trees,
                                                    const uint32_t* const* __restrict__ inj, int count, int width,
                                                    int batch) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= width) return;
  uint32_t* nodes = trees[blockIdx.y];
  uint32_t l[8], r[8];
  load8(nodes + 8 * ((size_t)width + j), l);
  for (int k = 0; k < count; ++k) {
    load8(inj[(size_t)k * batch + blockIdx.y] + 8 * (size_t)j, r);
    p2_node(l, r, l);
  }
  store8(nodes + 8 * ((size_t)width + j), l);
}

void launch_inject_seq(uint32_t* const* trees, const uint32_t* const* inj, int count, int width, int batch,
                       hipStream_t s) {
  if (count <= 0) return;
  const int T = width < 256 ? width : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_inject_seq, dim3((width + T - 1) / T, batch), dim3(T), 0, s, trees, inj, count, width, batch);
  LG_HIP(hipGetLastError());
}

void launch_inject(uint32_t* const* trees, const uint32_t* const* inj, int width, int batch, hipStream_t s) {
  const int T = width < 256 ? width : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_inject, dim3((width + T - 1) / T, batch), dim3(T), 0, s, trees, inj, width);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ input scatter
__global__ void k_scatter(const uint32_t* __restrict__ src, const uint64_t* __restrict__ off,
                          uint32_t* const* __restrict__ dst, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int k = blockIdx.y;
  dst[k][j] = src[off[k] + j];
}

void launch_scatter(const uint32_t* src, const uint64_t* off, uint32_t* const* dst, int n, int n_jobs, hipStream_t s) {
  if (n_jobs == 0) return;
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_scatter, dim3((n + T - 1) / T, n_jobs), dim3(T), 0, s, src, off, dst, n);
  LG_HIP(hipGetLastError());
}

__global__ void k_copy_cols(const uint32_t* const* __restrict__ src, uint32_t* const* __restrict__ dst, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int k = blockIdx.y;
  dst[k][j] = src[k][j];
}

void launch_copy_cols(const uint32_t* const* src, uint32_t* const* dst, int n, int n_jobs, hipStream_t s) {
  if (n_jobs == 0) return;
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_copy_cols, dim3((n + T - 1) / T, n_jobs), dim3(T), 0, s, src, dst, n);
  LG_HIP(hipGetLastError());
}

// dst[b][rows[i]] = src[b][i] for i < k (public-input entries into their
// zero-filled columns).
__global__ void k_scatter_rows(const uint32_t* const* __restrict__ src, const uint32_t* __restrict__ rows, int k,
                               uint32_t* const* __restrict__ dst) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  dst[blockIdx.y][rows[i]] = src[blockIdx.y][i];
}

void launch_scatter_rows(const uint32_t* const* src, const uint32_t* rows, int k, uint32_t* const* dst, int batch,
                         hipStream_t s) {
  if (k <= 0) return;
  const int T = 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_scatter_rows, dim3((k + T - 1) / T, batch), dim3(T), 0, s, src, rows, k, dst);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ GenCol
// All generated columns of a prove in one launch: block x of job k (the last
// job with block0 <= x) covers rows [(x - block0) 256, +256) of every instance
// (blockIdx.y); the per-kind values are those of k_range_col, k_lagrange_col,
// k_id_support and k_sigma.
__global__ __launch_bounds__(256) void k_gencol(const GenColJob* __restrict__ jobs, int njobs) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (jobs[mid].block0 <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  const GenColJob& J = jobs[lo];
  const int j = ((int)blockIdx.x - J.block0) * 256 + (int)threadIdx.x;
  if (j >= J.n) return;
  const int b = blockIdx.y;
  uint32_t v;
  if (J.kind == GENCOL_RANGE) {
    v = (uint32_t)j < J.a0 ? kb_to_mont((uint32_t)j) : 0u;
  } else if (J.kind == GENCOL_LAGRANGE) {
    v = (uint32_t)j == J.a0 ? KB_R : 0u;
  } else if (J.kind == GENCOL_ID) {
    v = kb_mul(J.gk, J.wpow[j]);
  } else {
    // an entry outside the permutation's cells raises the instance's error
    // flag and reads no table (generatePermutation fails on it)
    const uint32_t t = J.S[b][(size_t)J.a0 * J.n + j];
    if (t < J.len) {
      v = kb_mul(J.gpow[t >> J.logn], J.wpow[t & (uint32_t)(J.n - 1)]);
    } else {
      J.err[b] = 1u;
      v = 0u;
    }
  }
  J.outs[b][j] = v;
}

void launch_gencol(const GenColJob* jobs_dev, int njobs, int total_blocks, int batch, hipStream_t s) {
  if (njobs == 0 || total_blocks == 0) return;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_gencol, dim3(total_blocks, batch), dim3(256), 0, s, jobs_dev, njobs);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ E6 helpers
__device__ __forceinline__ E6 ld6(const uint32_t* p, size_t stride, size_t i) {
  E6 r;
#pragma unroll
  for (int l = 0; l < 6; ++l) r.v[l] = p[l * stride + i];
  return r;
}

__device__ __forceinline__ void st6(uint32_t* p, size_t stride, size_t i, const E6& x) {
#pragma unroll
  for (int l = 0; l < 6; ++l) p[l * stride + i] = x.v[l];
}

__device__ __forceinline__ E6 ldc6(const uint32_t* p) {
  E6 r;
#pragma unroll
  for (int l = 0; l < 6; ++l) r.v[l] = p[l];
  return r;
}

// ------------------------------------------------------------------ grand product
// ratio[i] = num[i] / den[i], one element per thread (grid over all rows).
__global__ __launch_bounds__(256) void k_ratio(const uint32_t* const* num, const uint32_t* const* den,
                                               uint32_t* const* ratio, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = blockIdx.y;
  st6(ratio[b], n, i, e6_mul(ld6(num[b], n, i), e6_inv(ld6(den[b], n, i))));
}

// One workgroup per instance.  Thread t owns the contiguous chunk
// [t*c, (t+1)*c) of the ratio vector; chunk products are scanned in LDS;
// gp[i] = prod_{k < i} ratio[k].
__global__ __launch_bounds__(1024) void k_prefix_product(const uint32_t* const* ratio, uint32_t* const* gp, int n) {
  __shared__ uint32_t sh[1024 * 6];
  const int b = blockIdx.x;
  const uint32_t* Rt = ratio[b];
  uint32_t* G = gp[b];
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int c = n / T;
  const int lo = tid * c;
  E6 acc = e6_one();
  for (int i = lo; i < lo + c; ++i) acc = e6_mul(acc, ld6(Rt, n, i));
#pragma unroll
  for (int l = 0; l < 6; ++l) sh[6 * tid + l] = acc.v[l];
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    E6 v = ldc6(sh + 6 * tid);
    E6 u = e6_one();
    if (tid >= off) u = ldc6(sh + 6 * (tid - off));
    __syncthreads();
    if (tid >= off) {
      E6 w = e6_mul(u, v);
#pragma unroll
      for (int l = 0; l < 6; ++l) sh[6 * tid + l] = w.v[l];
    }
    __syncthreads();
  }
  E6 run = tid == 0 ? e6_one() : ldc6(sh + 6 * (tid - 1));
  for (int i = lo; i < lo + c; ++i) {
    st6(G, n, i, run);
    run = e6_mul(run, ld6(Rt, n, i));
  }
}

// Three-pass scan of E6 columns for n >= 2^13 (tiles of 2048 rows): tile
// aggregates, an exclusive scan of the aggregates per instance, then each tile
// rescanned from its prefix.  MUL: products (grand product), else sums
// (logup); the field operations commute, so a tile's aggregate can be formed
// in any order, while the rescan keeps the row order.
constexpr int SCAN_T = 256, SCAN_PER = 8, SCAN_TILE = SCAN_T * SCAN_PER;

template <bool MUL>
__device__ __forceinline__ E6 scan_op(const E6& a, const E6& b) {
  if constexpr (MUL) return e6_mul(a, b);
  else return e6_add(a, b);
}
template <bool MUL>
__device__ __forceinline__ E6 scan_id() {
  if constexpr (MUL) return e6_one();
  else return e6_zero();
}

template <bool MUL>
__device__ __forceinline__ E6 block_reduce(E6 v, uint32_t* sh) {
#pragma unroll
  for (int l = 0; l < 6; ++l) sh[6 * threadIdx.x + l] = v.v[l];
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
      const E6 w = scan_op<MUL>(ldc6(sh + 6 * threadIdx.x), ldc6(sh + 6 * (threadIdx.x + off)));
#pragma unroll
      for (int l = 0; l < 6; ++l) sh[6 * threadIdx.x + l] = w.v[l];
    }
    __syncthreads();
  }
  return ldc6(sh);
}

// Inclusive scan of one E6 per thread over the block (Hillis-Steele in LDS);
// returns the exclusive prefix of this thread.
template <bool MUL>
__device__ __forceinline__ E6 block_exclusive(E6 v, uint32_t* sh) {
#pragma unroll
  for (int l = 0; l < 6; ++l) sh[6 * threadIdx.x + l] = v.v[l];
  __syncthreads();
  for (int off = 1; off < (int)blockDim.x; off <<= 1) {
    E6 me = ldc6(sh + 6 * threadIdx.x);
    const bool has = (int)threadIdx.x >= off;
    E6 other = has ? ldc6(sh + 6 * (threadIdx.x - off)) : scan_id<MUL>();
    __syncthreads();
    if (has) {
      me = scan_op<MUL>(other, me);
#pragma unroll
      for (int l = 0; l < 6; ++l) sh[6 * threadIdx.x + l] = me.v[l];
    }
    __syncthreads();
  }
  E6 r = threadIdx.x == 0 ? scan_id<MUL>() : ldc6(sh + 6 * (threadIdx.x - 1));
  __syncthreads();
  return r;
}

template <bool MUL>
__global__ __launch_bounds__(256) void k_scan_reduce(const uint32_t* const* in, uint32_t* agg, int n) {
  __shared__ uint32_t sh[SCAN_T * 6];
  const int b = blockIdx.y, tile = blockIdx.x, nt = gridDim.x;
  const uint32_t* X = in[b];
  E6 acc = scan_id<MUL>();
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) acc = scan_op<MUL>(acc, ld6(X, n, (size_t)tile * SCAN_TILE + k * SCAN_T + threadIdx.x));
  const E6 r = block_reduce<MUL>(acc, sh);
  if (threadIdx.x < 6) agg[((size_t)b * nt + tile) * 6 + threadIdx.x] = r.v[threadIdx.x];
}

// agg[b][tile] -> exclusive prefix of the tile aggregates; tot[b] (optional)
// = the product/sum of all of them.
template <bool MUL>
__global__ __launch_bounds__(1024) void k_scan_tiles(uint32_t* agg, uint32_t* tot, int nt) {
  __shared__ uint32_t sh[1024 * 6];
  const int b = blockIdx.x;
  uint32_t* A = agg + (size_t)b * nt * 6;
  const int c = nt / blockDim.x;
  const int lo = threadIdx.x * c;
  E6 acc = scan_id<MUL>();
  for (int i = lo; i < lo + c; ++i) acc = scan_op<MUL>(acc, ldc6(A + 6 * i));
  E6 run = block_exclusive<MUL>(acc, sh);
  for (int i = lo; i < lo + c; ++i) {
    const E6 a = ldc6(A + 6 * i);
#pragma unroll
    for (int l = 0; l < 6; ++l) A[6 * i + l] = run.v[l];
    run = scan_op<MUL>(run, a);
  }
  if (tot && threadIdx.x == blockDim.x - 1) {
#pragma unroll
    for (int l = 0; l < 6; ++l) tot[6 * b + l] = run.v[l];
  }
}

template <bool MUL, bool INCL>
__global__ __launch_bounds__(256) void k_scan_apply(const uint32_t* const* in, uint32_t* const* out, const uint32_t* pre,
                                                    int n) {
  __shared__ uint32_t tl[6 * SCAN_TILE];
  __shared__ uint32_t sh[SCAN_T * 6];
  const int b = blockIdx.y, tile = blockIdx.x, nt = gridDim.x;
  const uint32_t* X = in[b];
  const size_t base = (size_t)tile * SCAN_TILE;
#pragma unroll
  for (int l = 0; l < 6; ++l)
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
      tl[l * SCAN_TILE + k * SCAN_T + threadIdx.x] = X[(size_t)l * n + base + k * SCAN_T + threadIdx.x];
  __syncthreads();
  E6 v[SCAN_PER];
  E6 acc = scan_id<MUL>();
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
#pragma unroll
    for (int l = 0; l < 6; ++l) v[k].v[l] = tl[l * SCAN_TILE + SCAN_PER * threadIdx.x + k];
    acc = scan_op<MUL>(acc, v[k]);
  }
  E6 run = scan_op<MUL>(ldc6(pre + ((size_t)b * nt + tile) * 6), block_exclusive<MUL>(acc, sh));
#pragma unroll
  for (int k = 0; k < SCAN_PER; ++k) {
    if constexpr (INCL) run = scan_op<MUL>(run, v[k]);
#pragma unroll
    for (int l = 0; l < 6; ++l) tl[l * SCAN_TILE + SCAN_PER * threadIdx.x + k] = run.v[l];
    if constexpr (!INCL) run = scan_op<MUL>(run, v[k]);
  }
  __syncthreads();
  uint32_t* Y = out[b];
#pragma unroll
  for (int l = 0; l < 6; ++l)
#pragma unroll
    for (int k = 0; k < SCAN_PER; ++k)
      Y[(size_t)l * n + base + k * SCAN_T + threadIdx.x] = tl[l * SCAN_TILE + k * SCAN_T + threadIdx.x];
}

size_t scan_tmp_words(int n, int batch) { return n >= SCAN_MIN_TILED ? (size_t)batch * (n / SCAN_TILE) * 6 : 0; }

template <bool MUL, bool INCL>
static void launch_scan_tiled(const uint32_t* const* in, uint32_t* const* out, uint32_t* tmp, uint32_t* tot, int n,
                              int batch, hipStream_t s) {
  if (!tmp) throw std::runtime_error("scan: scratch missing");
  const int nt = n / SCAN_TILE;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_scan_reduce<MUL>, dim3(nt, batch), dim3(SCAN_T), 0, s, in, tmp, n);
  LG_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_scan_tiles<MUL>, dim3(batch), dim3(nt < 1024 ? nt : 1024), 0, s, tmp, tot, nt);
  LG_HIP(hipGetLastError());
  hipLaunchKernelGGL((k_scan_apply<MUL, INCL>), dim3(nt, batch), dim3(SCAN_T), 0, s, in, out, (const uint32_t*)tmp, n);
  LG_HIP(hipGetLastError());
}

void launch_grand_product(const uint32_t* const* num, const uint32_t* const* den, uint32_t* const* ratio,
                          uint32_t* const* gp, int n, int batch, uint32_t* scan_tmp, hipStream_t s) {
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_ratio, dim3((n + T - 1) / T, batch), dim3(T), 0, s, num, den, ratio, n);
  LG_HIP(hipGetLastError());
  if (n >= SCAN_MIN_TILED) {
    launch_scan_tiled<true, false>((const uint32_t* const*)ratio, gp, scan_tmp, nullptr, n, batch, s);
    return;
  }
  int T2 = n < 1024 ? n : 1024;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_prefix_product, dim3(batch), dim3(T2), 0, s, (const uint32_t* const*)ratio, gp, n);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ lookups and logup
// Range column (board/module.go:56-71): row i holds i for i < bound.
__global__ void k_fill(uint32_t* const* p, int n, uint32_t v) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  p[blockIdx.y][j] = v;
}

void launch_fill(uint32_t* const* p, int n, uint32_t v, int batch, hipStream_t s) {
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_fill, dim3((n + T - 1) / T, batch), dim3(T), 0, s, p, n, v);
  LG_HIP(hipGetLastError());
}

// Counts the canonical values of src[b][0..n) (Montgomery words) below bound
// into counts[b]; values at or above bound count into *err[b].
__global__ void k_histogram(const uint32_t* const* src, int n, uint32_t* const* counts, uint32_t bound,
                            uint32_t* const* err) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int b = blockIdx.y;
  const uint32_t v = kb_from_mont(src[b][j]);
  if (v < bound) atomicAdd(&counts[b][v], 1u);
  else atomicAdd(err[b], 1u);
}

// Tables up to HIST_LDS_BINS entries: each workgroup counts HIST_ROWS rows into
// an LDS copy of the table and adds its non-zero bins once (the global
// table sees one atomic per bin and workgroup instead of one per row).
constexpr int HIST_LDS_BINS = 4096;
constexpr int HIST_ROWS = 4096;

__global__ __launch_bounds__(256) void k_histogram_lds(const uint32_t* const* src, int n, uint32_t* const* counts,
                                                       uint32_t bound, uint32_t* const* err) {
  __shared__ uint32_t h[HIST_LDS_BINS];
  __shared__ uint32_t e;
  const int b = blockIdx.y;
  for (uint32_t i = threadIdx.x; i < bound; i += blockDim.x) h[i] = 0u;
  if (threadIdx.x == 0) e = 0u;
  __syncthreads();
  const int j0 = blockIdx.x * HIST_ROWS;
  const int j1 = min(n, j0 + HIST_ROWS);
  const uint32_t* in = src[b];
  for (int j = j0 + (int)threadIdx.x; j < j1; j += blockDim.x) {
    const uint32_t v = kb_from_mont(in[j]);
    if (v < bound) atomicAdd(&h[v], 1u);
    else atomicAdd(&e, 1u);
  }
  __syncthreads();
  uint32_t* c = counts[b];
  for (uint32_t i = threadIdx.x; i < bound; i += blockDim.x)
    if (h[i]) atomicAdd(&c[i], h[i]);
  if (threadIdx.x == 0 && e) atomicAdd(err[b], e);
}

void launch_histogram(const uint32_t* const* src, int n, uint32_t* const* counts, uint32_t bound,
                      uint32_t* const* err, int batch, hipStream_t s) {
  if (n <= 0 || batch <= 0) return;
  if (!prelaunch()) return;
  if (bound <= (uint32_t)HIST_LDS_BINS) {
    hipLaunchKernelGGL(k_histogram_lds, dim3((n + HIST_ROWS - 1) / HIST_ROWS, batch), dim3(256), 0, s, src, n, counts,
                       bound, err);
  } else {
    const int T = n < 256 ? n : 256;
    hipLaunchKernelGGL(k_histogram, dim3((n + T - 1) / T, batch), dim3(T), 0, s, src, n, counts, bound, err);
  }
  LG_HIP(hipGetLastError());
}

// ---- general lookup multiplicities
// Tuples are read where they lie (row i of table t, one word per column), so
// their width is not bounded by registers.
__device__ __forceinline__ uint32_t lk_hash(const LkTable& t, uint32_t i, int w) {
  uint32_t h = 0x9e3779b9u;
  for (int c = 0; c < w; ++c) {
    h = (h ^ t.col[c][i]) * 0x85ebca6bu;
    h ^= h >> 13;
  }
  h *= 0xc2b2ae35u;
  return h ^ (h >> 16);
}

// the table of global row g among n tables (offsets ascending)
__device__ __forceinline__ int lk_table_of(const LkTable* T, int n, uint32_t g) {
  int k = 0;
  while (k + 1 < n && g >= T[k + 1].off) ++k;
  return k;
}

// row g of tables T (n of them) equals row qi of table q
__device__ __forceinline__ bool lk_row_equal(const LkTable* T, int n, uint32_t g, const LkTable& q, uint32_t qi, int w) {
  const LkTable& t = T[lk_table_of(T, n, g)];
  const uint32_t i = g - t.off;
  for (int c = 0; c < w; ++c)
    if (t.col[c][i] != q.col[c][qi]) return false;
  return true;
}

// Finds (or with claim, claims) the slot of the tuple at row qi of table q
// among rows of tables T: slot_row holds a representative row + 1, immutable
// once set, whose words stay readable in its table.  The tables hold at most
// half as many rows as slots, so a probe meets an empty slot or its tuple
// within cap steps.
__device__ __forceinline__ int lk_probe(uint32_t* slot_row, uint32_t cap_log, const LkTable* T, int n, const LkTable& q,
                                        uint32_t qi, int w, uint32_t g_claim, bool claim) {
  const uint32_t mask = (1u << cap_log) - 1u;
  uint32_t slot = lk_hash(q, qi, w) & mask;
  for (uint32_t it = 0; it <= mask; ++it) {
    uint32_t cur = __hip_atomic_load(&slot_row[slot], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (cur == 0u) {
      if (!claim) return -1;
      cur = atomicCAS(&slot_row[slot], 0u, g_claim + 1u);
      if (cur == 0u) return (int)slot;
    }
    if (lk_row_equal(T, n, cur - 1u, q, qi, w)) return (int)slot;
    slot = (slot + 1u) & mask;
  }
  return -2;  // not reached while the table is at most half full
}

__global__ __launch_bounds__(256) void k_lookup_targets(const LkTable* __restrict__ tabs, const LkJob* __restrict__ jobs,
                                                        int nt, int ns, int w, uint32_t total_t) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_t) return;
  const LkTable* T = tabs + (size_t)blockIdx.y * (nt + ns);
  const LkJob& J = jobs[blockIdx.y];
  const LkTable& t = T[lk_table_of(T, nt, g)];
  const uint32_t i = g - t.off;
  if (i >= t.n || (t.sel && t.sel[i] == 0u)) return;
  const int slot = lk_probe(J.slot_row, J.cap_log, T, nt, t, i, w, g, true);
  if (slot < 0) {
    atomicOr(J.err, 2u);
    return;
  }
  atomicMin(&J.slot_first[slot], g);
}

__global__ __launch_bounds__(256) void k_lookup_sources(const LkTable* __restrict__ tabs, const LkJob* __restrict__ jobs,
                                                        int nt, int ns, int w, uint32_t total_s) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_s) return;
  const LkTable* T = tabs + (size_t)blockIdx.y * (nt + ns);
  const LkTable* S = T + nt;
  const LkJob& J = jobs[blockIdx.y];
  const LkTable& src = S[lk_table_of(S, ns, g)];
  const uint32_t i = g - src.off;
  if (i >= src.n) return;
  const uint32_t wt = src.sel ? kb_from_mont(src.sel[i]) : 1u;
  if (wt == 0u) return;
  const int slot = lk_probe(J.slot_row, J.cap_log, T, nt, src, i, w, 0u, false);
  if (slot >= 0) {
    atomicAdd(&J.slot_w[slot], (unsigned long long)wt);
    return;
  }
  if (!J.miss_row) {  // weights of ones: a missing tuple's sum is not 0 mod p
    atomicOr(J.err, 1u);
    return;
  }
  const int ms = lk_probe(J.miss_row, J.miss_cap_log, S, ns, src, i, w, g, true);
  if (ms < 0) {
    atomicOr(J.err, 2u);
    return;
  }
  atomicAdd(&J.miss_w[ms], (unsigned long long)wt);
}

__global__ __launch_bounds__(256) void k_lookup_assign(const LkTable* __restrict__ tabs, const LkJob* __restrict__ jobs,
                                                       int nt, int ns, int w, uint32_t total_t) {
  const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= total_t) return;
  const LkTable* T = tabs + (size_t)blockIdx.y * (nt + ns);
  const LkJob& J = jobs[blockIdx.y];
  const LkTable& t = T[lk_table_of(T, nt, g)];
  const uint32_t i = g - t.off;
  if (i >= t.n) return;
  uint32_t m = 0u;
  if (!t.sel || t.sel[i] != 0u) {
    const int slot = lk_probe(J.slot_row, J.cap_log, T, nt, t, i, w, 0u, false);
    if (slot >= 0 && J.slot_first[slot] == g) {
      m = kb_to_mont((uint32_t)(J.slot_w[slot] % KB_P));
      if (t.sel) m = kb_mul(m, kb_inv(t.sel[i]));
    }
  }
  t.mult[i] = m;
}

__global__ __launch_bounds__(256) void k_lookup_missing(const LkJob* __restrict__ jobs, uint32_t max_cap) {
  const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
  const LkJob& J = jobs[blockIdx.y];
  if (!J.miss_row || s >= (1u << J.miss_cap_log) || s >= max_cap) return;
  if (J.miss_row[s] != 0u && J.miss_w[s] % KB_P != 0u) atomicOr(J.err, 1u);
}

void launch_lookup_general(const LkTable* tabs, const LkJob* jobs, int nt, int ns, int w, uint32_t total_t,
                           uint32_t total_s, uint32_t max_cap, uint32_t max_miss_cap, int batch, hipStream_t s) {
  if (batch <= 0) return;
  if (w < 1 || nt < 1 || ns < 1) throw std::runtime_error("launch_lookup_general: bad shape");
  if (!prelaunch()) return;
  const int T = 256;
  if (total_t) {
    hipLaunchKernelGGL(k_lookup_targets, dim3((total_t + T - 1) / T, batch), dim3(T), 0, s, tabs, jobs, nt, ns, w,
                       total_t);
    LG_HIP(hipGetLastError());
  }
  if (total_s) {
    hipLaunchKernelGGL(k_lookup_sources, dim3((total_s + T - 1) / T, batch), dim3(T), 0, s, tabs, jobs, nt, ns, w,
                       total_s);
    LG_HIP(hipGetLastError());
  }
  if (total_t) {
    hipLaunchKernelGGL(k_lookup_assign, dim3((total_t + T - 1) / T, batch), dim3(T), 0, s, tabs, jobs, nt, ns, w,
                       total_t);
    LG_HIP(hipGetLastError());
  }
  if (max_miss_cap) {
    hipLaunchKernelGGL(k_lookup_missing, dim3((max_miss_cap + T - 1) / T, batch), dim3(T), 0, s, jobs, max_miss_cap);
    LG_HIP(hipGetLastError());
  }
}

// In place: counts (below p) to field elements, Montgomery.
__global__ void k_counts_to_field(uint32_t* const* c, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  uint32_t* p = c[blockIdx.y];
  p[j] = kb_to_mont(p[j]);
}

void launch_counts_to_field(uint32_t* const* c, int n, int batch, hipStream_t s) {
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_counts_to_field, dim3((n + T - 1) / T, batch), dim3(T), 0, s, c, n);
  LG_HIP(hipGetLastError());
}

// acc[i] = (first ? 0 : acc[i]) + m[i] / e[i] over E6 columns (limb stride n).
__global__ __launch_bounds__(256) void k_logup_acc(const uint32_t* const* e, const uint32_t* const* m,
                                                   uint32_t* const* acc, int n, int first) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = blockIdx.y;
  E6 v = e6_mul(ld6(m[b], n, i), e6_inv(ld6(e[b], n, i)));
  if (!first) v = e6_add(v, ld6(acc[b], n, i));
  st6(acc[b], n, i, v);
}

void launch_logup_acc(const uint32_t* const* e, const uint32_t* const* m, uint32_t* const* acc, int n, int first,
                      int batch, hipStream_t s) {
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_logup_acc, dim3((n + T - 1) / T, batch), dim3(T), 0, s, e, m, acc, n, first);
  LG_HIP(hipGetLastError());
}

// Inclusive running sum out[i] = sum_{r <= i} in[r] of an E6 column, one
// workgroup per instance (chunk per thread, LDS scan of the chunk sums);
// the total goes to tot[6 b .. 6 b + 6).
__global__ __launch_bounds__(1024) void k_prefix_sum(const uint32_t* const* in, uint32_t* const* out, uint32_t* tot,
                                                     int n) {
  __shared__ uint32_t sh[1024 * 6];
  const int b = blockIdx.x;
  const uint32_t* X = in[b];
  uint32_t* Y = out[b];
  const int T = blockDim.x;
  const int tid = threadIdx.x;
  const int c = n / T;
  const int lo = tid * c;
  E6 acc = e6_zero();
  for (int i = lo; i < lo + c; ++i) acc = e6_add(acc, ld6(X, n, i));
#pragma unroll
  for (int l = 0; l < 6; ++l) sh[6 * tid + l] = acc.v[l];
  __syncthreads();
  for (int off = 1; off < T; off <<= 1) {
    E6 v = ldc6(sh + 6 * tid);
    E6 u = e6_zero();
    if (tid >= off) u = ldc6(sh + 6 * (tid - off));
    __syncthreads();
    if (tid >= off) {
      E6 w = e6_add(u, v);
#pragma unroll
      for (int l = 0; l < 6; ++l) sh[6 * tid + l] = w.v[l];
    }
    __syncthreads();
  }
  E6 run = tid == 0 ? e6_zero() : ldc6(sh + 6 * (tid - 1));
  for (int i = lo; i < lo + c; ++i) {
    run = e6_add(run, ld6(X, n, i));
    st6(Y, n, i, run);
  }
  if (tid == T - 1) {
#pragma unroll
    for (int l = 0; l < 6; ++l) tot[6 * b + l] = run.v[l];
  }
}

// L_i = S_i - (i + 1) * tot / n in place (board/logup.go:160-171).
__global__ void k_logup_finish(uint32_t* const* col, const uint32_t* tot, int n, uint32_t ninv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = blockIdx.y;
  const E6 avg = e6_scale(ldc6(tot + 6 * b), ninv);
  const E6 v = e6_sub(ld6(col[b], n, i), e6_scale(avg, kb_to_mont((uint32_t)i + 1u)));
  st6(col[b], n, i, v);
}

void launch_cyclic_logup_finish(const uint32_t* const* acc, uint32_t* const* col, uint32_t* tot, int n,
                                uint32_t ninv, int batch, uint32_t* scan_tmp, hipStream_t s) {
  if (n >= SCAN_MIN_TILED) {
    launch_scan_tiled<false, true>(acc, col, scan_tmp, tot, n, batch, s);
  } else {
    int T2 = n < 1024 ? n : 1024;
    if (!prelaunch()) return;
    hipLaunchKernelGGL(k_prefix_sum, dim3(batch), dim3(T2), 0, s, acc, col, tot, n);
    LG_HIP(hipGetLastError());
  }
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_logup_finish, dim3((n + T - 1) / T, batch), dim3(T), 0, s, col, (const uint32_t*)tot, n, ninv);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ quotient
__global__ void k_quot_combine(const uint32_t* Y, uint32_t* coef, const uint32_t* ginv, const uint32_t* wr, int rho,
                               int n) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= n) return;
  const int l = blockIdx.y;
  const int b = blockIdx.z;
  const size_t rn = (size_t)rho * n;
  const uint32_t* y = Y + ((size_t)b * 6 + l) * rn;
  uint32_t yi[8];
#pragma unroll
  for (int i = 0; i < 8; ++i)
    if (i < rho) yi[i] = y[(size_t)i * n + a];
  for (int t = 0; t < rho; ++t) {
    uint32_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i)
      if (i < rho) acc = kb_add(acc, kb_mul(wr[(t * i) % rho], yi[i]));
    coef[(((size_t)b * rho + t) * 6 + l) * n + a] = kb_mul(acc, ginv[(size_t)t * n + a]);
  }
}

void launch_quot_combine(const uint32_t* Y, uint32_t* coef, const uint32_t* ginv, const uint32_t* wr, int rho, int n,
                         int batch, hipStream_t s) {
  if (rho > 8) throw std::runtime_error("quotient: more than 8 cosets");
  const int T = n < 256 ? n : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_quot_combine, dim3((n + T - 1) / T, 6, batch), dim3(T), 0, s, Y, coef, ginv, wr, rho, n);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ claimed values
__global__ void k_lagrange_weights(const uint32_t* zeta, const uint32_t* common, const uint32_t* wpow,
                                   uint32_t* const* W, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int b = blockIdx.y;
  E6 z = ldc6(zeta + 6 * b);
  E6 c = ldc6(common + 6 * b);
  const uint32_t wi = wpow[i];
  E6 d = z;
  d.v[0] = kb_sub(d.v[0], wi);
  E6 r = e6_mul(e6_scale(c, wi), e6_inv(d));
  st6(W[b], n, i, r);
}

void launch_lagrange_weights(const uint32_t* zeta, const uint32_t* common, const uint32_t* wpow, uint32_t* const* W,
                             int n, int batch, hipStream_t s) {
  const int T = 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_lagrange_weights, dim3((n + T - 1) / T, batch), dim3(T), 0, s, zeta, common, wpow, W, n);
  LG_HIP(hipGetLastError());
}

// Sum of a field word over the 256 threads of a block: wave reduction by
// lane exchanges, then the 4 wave sums through LDS (sh: 4 words per value).
__device__ __forceinline__ uint32_t block_sum4(uint32_t v, uint32_t* sh) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = kb_add(v, (uint32_t)__shfl_xor((int)v, o, 64));
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  return kb_add(kb_add(sh[0], sh[1]), kb_add(sh[2], sh[3]));
}

// Block (set, chunk): rows j in [chunk n/S, (chunk+1) n/S) of one claim set.
// For each group of up to CLAIMS_GROUP claims, every thread loads W[j] once
// per row and accumulates W[j] P_c[(j + s_c) mod n] for every claim c of the
// group; the block sums are the claimed values (S == 1) or partial sums.
__global__ __launch_bounds__(256) void k_claims(const ClaimSet* __restrict__ sets, const ClaimItem* __restrict__ items,
                                                int n, uint32_t* part) {
  __shared__ uint32_t sh[CLAIMS_GROUP * 6][4];
  const ClaimSet set = sets[blockIdx.x];
  const int S = gridDim.y;
  const int len = n / S;
  const int lo = blockIdx.y * len;
  const uint32_t mask = (uint32_t)(n - 1);
  for (int g0 = 0; g0 < set.count; g0 += CLAIMS_GROUP) {
    const int cnt = min(CLAIMS_GROUP, set.count - g0);
    E6 acc[CLAIMS_GROUP];
#pragma unroll
    for (int c = 0; c < CLAIMS_GROUP; ++c) acc[c] = e6_zero();
    for (int j = lo + (int)threadIdx.x; j < lo + len; j += blockDim.x) {
      const E6 w = ld6(set.W, n, j);
#pragma unroll
      for (int c = 0; c < CLAIMS_GROUP; ++c) {
        if (c < cnt) {
          const ClaimItem it = items[set.first + g0 + c];
          const uint32_t i = ((uint32_t)j + (uint32_t)it.shift) & mask;
          if (it.ext) acc[c] = e6_add(acc[c], e6_mul(w, ld6(it.poly, n, i)));
          else acc[c] = e6_add(acc[c], e6_scale(w, it.poly[i]));
        }
      }
    }
    // the wave sums of all 6 cnt words of the group (independent lane
    // exchanges, so their latencies overlap), one barrier, then thread k
    // adds the four wave sums of word k
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < CLAIMS_GROUP; ++c) {
      if (c < cnt) {
#pragma unroll
        for (int l = 0; l < 6; ++l) {
          uint32_t v = acc[c].v[l];
#pragma unroll
          for (int o = 32; o > 0; o >>= 1) v = kb_add(v, (uint32_t)__shfl_xor((int)v, o, 64));
          if (lane == 0) sh[c * 6 + l][wv] = v;
        }
      }
    }
    __syncthreads();
    const int k = threadIdx.x;
    if (k < cnt * 6) {
      const int item = set.first + g0 + k / 6, l = k % 6;
      const uint32_t v = kb_add(kb_add(sh[k][0], sh[k][1]), kb_add(sh[k][2], sh[k][3]));
      if (S == 1) items[item].out[l] = v;
      else part[((size_t)item * S + blockIdx.y) * 6 + l] = v;
    }
    __syncthreads();  // sh is reused by the next group
  }
}

// One block per claim: the sum of its S partial sums.
__global__ __launch_bounds__(256) void k_claims_sum(const ClaimItem* items, int S, const uint32_t* part) {
  __shared__ uint32_t sh[6][4];
  const int j = blockIdx.x;
  E6 acc = e6_zero();
  for (int c = threadIdx.x; c < S; c += blockDim.x) acc = e6_add(acc, ldc6(part + ((size_t)j * S + c) * 6));
#pragma unroll
  for (int l = 0; l < 6; ++l) {
    const uint32_t v = block_sum4(acc.v[l], sh[l]);
    if (threadIdx.x == 0) items[j].out[l] = v;
  }
}

// 1024 rows per block (a 40-instance batch of 2^14-row modules: 640 blocks;
// one 2^22-row module: 4096 blocks)
int claims_split(int n, int rows_per_block) {
  return std::max(1, std::min(CLAIMS_MAX_SPLIT, n / std::max(256, rows_per_block)));
}

void launch_claims(const ClaimSet* sets, int n_sets, const ClaimItem* items, int n_items, int n, uint32_t* part,
                   hipStream_t s, int rows_per_block) {
  if (n_sets == 0 || n_items == 0) return;
  const int S = claims_split(n, rows_per_block);
  if (S > 1 && !part) throw std::runtime_error("launch_claims: partial-sum buffer missing");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_claims, dim3(n_sets, S), dim3(256), 0, s, sets, items, n, part);
  LG_HIP(hipGetLastError());
  if (S > 1) {
    hipLaunchKernelGGL(k_claims_sum, dim3(n_items), dim3(256), 0, s, items, S, (const uint32_t*)part);
    LG_HIP(hipGetLastError());
  }
}

// ------------------------------------------------------------------ DEEP
constexpr int DEEP_MAX_POINTS = 8;

// acc += sum_{p in [C0, C0 + C)} num[p] / (z_p - wx) with one E6 inversion
// (prefix products, one inverse, back substitution); C <= 4 keeps the
// products in registers.
template <int NP, int C0, int C>
__device__ __forceinline__ void deep_points(const E6* num, const DeepDesc& d, int b, uint32_t wx, E6& acc) {
  E6 den[C], pre[C];
  const int pb = d.points_mod ? b % d.points_mod : b;
#pragma unroll
  for (int k = 0; k < C; ++k) {
    den[k] = ldc6(d.points + ((size_t)pb * NP + C0 + k) * 6);
    den[k].v[0] = kb_sub(den[k].v[0], wx);
    pre[k] = k ? e6_mul(pre[k - 1], den[k]) : den[k];
  }
  E6 inv = e6_inv(pre[C - 1]);
#pragma unroll
  for (int k = C - 1; k >= 0; --k) {
    const E6 ik = k ? e6_mul(inv, pre[k - 1]) : inv;  // 1 / den_k
    if (k) inv = e6_mul(inv, den[k]);
    acc = e6_add(acc, e6_mul(num[C0 + k], ik));
  }
}

// One row x per thread.  NP (the class's distinct opening points) is a
// compile-time constant, so num[] and the denominators stay in registers
// (every index below is static after unrolling), and the NP inverses
// 1/(z_p - w^x) come from one E6 inversion (prefix products, one inverse,
// back substitution).
template <int NP>
__global__ __launch_bounds__(256) void k_deep(DeepDesc d) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (x >= d.n) return;
  E6 num[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) num[p] = ldc6(d.consts + ((size_t)b * NP + p) * 6);
  const uint32_t* const* cols = d.cols + (size_t)b * d.n_cols;
  for (int i = 0; i < d.n_bundles; ++i) {
    const DeepBundle bd = d.bundles[i];
    const E6 a = ldc6(d.scale + ((size_t)b * d.n_bundles + i) * 6);
    const E6 t = bd.ext ? e6_mul(a, ld6(cols[bd.col], d.n, x)) : e6_scale(a, cols[bd.col][x]);
    for (int k = 0; k < bd.n_pts; ++k) {
      const int p = d.pt_list[bd.first_pt + k];
#pragma unroll
      for (int q = 0; q < NP; ++q)
        if (q == p) num[q] = e6_sub(num[q], t);
    }
  }
  const uint32_t wx = d.wpow[x];
  E6 acc = e6_zero();
  deep_points<NP, 0, (NP < 4 ? NP : 4)>(num, d, b, wx, acc);
  if constexpr (NP > 4) deep_points<NP, 4, NP - 4>(num, d, b, wx, acc);
  st6(d.out[b], d.n, x, acc);
}

void launch_deep(const DeepDesc& d, int batch, hipStream_t s) {
  if (d.n_points < 1 || d.n_points > DEEP_MAX_POINTS) throw std::runtime_error("deep: 1..8 distinct opening points");
  const int T = 256;
  const dim3 g((d.n + T - 1) / T, batch);
  if (!prelaunch()) return;
  switch (d.n_points) {
    case 1: hipLaunchKernelGGL(k_deep<1>, g, dim3(T), 0, s, d); break;
    case 2: hipLaunchKernelGGL(k_deep<2>, g, dim3(T), 0, s, d); break;
    case 3: hipLaunchKernelGGL(k_deep<3>, g, dim3(T), 0, s, d); break;
    case 4: hipLaunchKernelGGL(k_deep<4>, g, dim3(T), 0, s, d); break;
    case 5: hipLaunchKernelGGL(k_deep<5>, g, dim3(T), 0, s, d); break;
    case 6: hipLaunchKernelGGL(k_deep<6>, g, dim3(T), 0, s, d); break;
    case 7: hipLaunchKernelGGL(k_deep<7>, g, dim3(T), 0, s, d); break;
    default: hipLaunchKernelGGL(k_deep<8>, g, dim3(T), 0, s, d); break;
  }
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ FRI
__global__ void k_fri_fold(const uint32_t* const* in, uint32_t* const* out, const uint32_t* alpha,
                           const uint32_t* xinv2, int half) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= half) return;
  const int b = blockIdx.y;
  const uint32_t* L = in[b];
  const size_t m = 2 * (size_t)half;
  E6 p = ld6(L, m, 2 * i), q = ld6(L, m, 2 * i + 1);
  // xinv2[i] = xinv_i / 2:  next = (p + q) / 2 + alpha * (p - q) * xinv_i / 2
  constexpr uint32_t INV2_MONT = 0x00ffffffu;  // 1/2 in Montgomery form (2^31 mod p)
  E6 s = e6_scale(e6_add(p, q), INV2_MONT);
  E6 dlt = e6_mul(e6_scale(e6_sub(p, q), xinv2[i]), ldc6(alpha + 6 * b));
  st6(out[b], half, i, e6_add(s, dlt));
}

// Layers of at most this many leaf pairs take the cooperative form below.

// One FRI fold fused with the leaf hashing of the folded layer's tree: thread
// k forms next[2k] and next[2k+1] (from in[4k .. 4k+3]), stores them, and
// hashes the row pair as the leaf sponge [LEAF, 0, hdr2, next[2k] limbs,
// next[2k+1] limbs] (nb = 0, ne = 1: 15 elements, one permutation;
// internal/fri/hash.go:39-56) into leaves[b] + 8k.
__global__ __launch_bounds__(256) void k_fri_fold_leaf(const uint32_t* const* in, uint32_t* const* out,
                                                       const uint32_t* alpha, const uint32_t* xinv2, int half,
                                                       uint32_t hdr2, uint32_t* const* leaves) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (2 * k >= half) return;
  const int b = blockIdx.y;
  const uint32_t* L = in[b];
  const size_t m = 2 * (size_t)half;
  constexpr uint32_t INV2_MONT = 0x00ffffffu;  // 1/2 in Montgomery form (2^31 mod p)
  const E6 a = ldc6(alpha + 6 * b);
  uint32_t s[24];
#pragma unroll
  for (int i = 0; i < 24; ++i) s[i] = 0;
  s[0] = TAG_LEAF_MONT;
  s[2] = hdr2;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int i = 2 * k + t;
    const E6 p = ld6(L, m, 2 * i), q = ld6(L, m, 2 * i + 1);
    const E6 r = e6_add(e6_scale(e6_add(p, q), INV2_MONT), e6_mul(e6_scale(e6_sub(p, q), xinv2[i]), a));
    st6(out[b], half, i, r);
#pragma unroll
    for (int l = 0; l < 6; ++l) s[3 + 6 * t + l] = r.v[l];
  }
  p2_permute<24>(s);
  uint32_t* o = leaves[b] + (size_t)8 * k;
#pragma unroll
  for (int i = 0; i < 8; ++i) o[i] = s[i];
}

// The same for small layers, latency-bound: one 16-lane row per leaf
// (P2Coop16).  Every lane of the row forms both fold outputs; lanes 0..5
// store the limbs of next[2k], lanes 6..11 those of next[2k+1], and the row
// hashes the leaf with one cooperative permutation.
__global__ __launch_bounds__(256) void k_fri_fold_leaf_coop(const uint32_t* const* in, uint32_t* const* out,
                                                            const uint32_t* alpha, const uint32_t* xinv2, int half,
                                                            uint32_t hdr2, uint32_t* const* leaves) {
  const int l = threadIdx.x & 15;
  const int k = blockIdx.x * 16 + (threadIdx.x >> 4);
  if (2 * k >= half) return;  // whole rows only: 16 divides the block
  const int b = blockIdx.y;
  const uint32_t* L = in[b];
  const size_t m = 2 * (size_t)half;
  constexpr uint32_t INV2_MONT = 0x00ffffffu;
  const E6 a6 = ldc6(alpha + 6 * b);
  E6 r[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int i = 2 * k + t;
    const E6 p = ld6(L, m, 2 * i), q = ld6(L, m, 2 * i + 1);
    r[t] = e6_add(e6_scale(e6_add(p, q), INV2_MONT), e6_mul(e6_scale(e6_sub(p, q), xinv2[i]), a6));
  }
  if (l < 12) {
    const int t = l < 6 ? 0 : 1, limb = l < 6 ? l : l - 6;
    out[b][(size_t)limb * half + 2 * k + t] = r[t].v[limb];
  }
  // state elements: 0 LEAF, 1 zero, 2 hdr2, 3..8 next[2k], 9..14 next[2k+1]
  uint32_t sa = 0u, sb = 0u;
  if (l == 0) sa = TAG_LEAF_MONT;
  else if (l == 2) sa = hdr2;
#pragma unroll
  for (int i = 3; i < 9; ++i)
    if (l == i) sa = r[0].v[i - 3];
#pragma unroll
  for (int i = 9; i < 12; ++i)
    if (l == i) sa = r[1].v[i - 9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    if (l == i) sb = r[1].v[3 + i];
  P2Coop16 P;
  P.init(l);
  P.permute(sa, sb);
  if (l < 8) leaves[b][(size_t)8 * k + l] = sa;
}

void launch_fri_fold_leaf(const uint32_t* const* in, uint32_t* const* out, const uint32_t* alpha,
                          const uint32_t* xinv2, int half, uint32_t hdr2, uint32_t* const* leaves, int batch,
                          hipStream_t s, int coop_pairs) {
  if (half < 2 || (half & 1)) throw std::runtime_error("fri fold+leaf: layer length must be even");
  const int pairs = half / 2;
  if (!prelaunch()) return;
  if (pairs <= coop_pairs) {
    // rows of 16 lanes, 16 rows per block; rows past the layer return whole
    hipLaunchKernelGGL(k_fri_fold_leaf_coop, dim3((pairs + 15) / 16, batch), dim3(256), 0, s, in, out, alpha, xinv2,
                       half, hdr2, leaves);
  } else {
    const int T = pairs < 256 ? pairs : 256;
    hipLaunchKernelGGL(k_fri_fold_leaf, dim3((pairs + T - 1) / T, batch), dim3(T), 0, s, in, out, alpha, xinv2,
                       half, hdr2, leaves);
  }
  LG_HIP(hipGetLastError());
}

void launch_fri_fold(const uint32_t* const* in, uint32_t* const* out, const uint32_t* alpha, const uint32_t* xinv2,
                     int half, int batch, hipStream_t s) {
  const int T = half < 256 ? half : 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_fri_fold, dim3((half + T - 1) / T, batch), dim3(T), 0, s, in, out, alpha, xinv2, half);
  LG_HIP(hipGetLastError());
}

// Fiat-Shamir challenge of one FRI fold round on the device, so the commit
// phase of FRI runs without a host round trip per round.  Per instance the
// message is prefix (StringToElements(FSID, name), shared) || prev[b] (the
// previous challenge's digest) || the root of trees[b] (node 1), absorbed by
// the rate-16 overwrite sponge (internal/fiat-shamir/transcript.go:285-315;
// internal/hash/poseidon2.go:165-255).  One row of 16 lanes per instance
// (P2Coop16); the digest goes to dig[b] (8 words), the E6 challenge, its first
// 6 limbs, to alpha[b], and the root itself to roots[b] (8 words).

__global__ __launch_bounds__(64) void k_fri_challenge(ChalPrefix pre, int n_pre, const uint32_t* __restrict__ prev,
                                                      uint32_t* const* __restrict__ trees, uint32_t* __restrict__ dig,
                                                      uint32_t* __restrict__ alpha, uint32_t* __restrict__ roots,
                                                      int batch) {
  const int l = threadIdx.x & 15;
  const int b_raw = blockIdx.x * 4 + (threadIdx.x >> 4);
  const bool live = b_raw < batch;
  const int b = live ? b_raw : batch - 1;
  const uint32_t* root = trees[b] + 8;
  P2Coop16 P;
  P.init(l);
  const uint32_t d = fri_challenge_row(P, l, pre, n_pre, prev + 8 * b, root);
  if (live && l < 8) {
    dig[8 * b + l] = d;
    roots[8 * b + l] = root[l];
    if (l < 6) alpha[6 * b + l] = d;
  }
}

// alpha_DEEP on the device, with the DEEP constants of one class
// (deep.go:71-266): per instance (one 16-lane row) the sponge of
// prefix || prev[b] || the claimed values in binding order (value order[j] of
// cvals[b]), then, on lane 0, scale[b][i] = alpha^i for the bundles and
// consts[b][p] = sum over the terms (i, k, p) of alpha^i cvals[b][k].
__global__ __launch_bounds__(64) void k_deep_alpha(DeepAlphaArgs a, ChalPrefix pre, int batch) {
  __shared__ uint32_t alpha_sh[4][8];
  const int l = threadIdx.x & 15, row = threadIdx.x >> 4;
  const int b_raw = blockIdx.x * 4 + row;
  const bool live = b_raw < batch;
  const int b = live ? b_raw : batch - 1;
  const uint32_t* cv = a.cvals + (size_t)b * a.ncv * 6;
  auto msg = [&](int i) -> uint32_t {
    if (i < a.n_pre) return pre.w[i];
    if (i < a.n_pre + 8) return a.prev[8 * b + (i - a.n_pre)];
    const int j = i - a.n_pre - 8;
    return cv[6 * a.order[j / 6] + j % 6];
  };
  P2Coop16 P;
  P.init(l);
  uint32_t x = 0u, y = 0u;
  const int len = a.n_pre + 8 + 6 * a.ncv;
  for (int off = 0; off < len; off += 16) {
    const int k = len - off < 16 ? len - off : 16;
    if (l < 12 && l < k) x = msg(off + l);
    if (l < 12 && l + 12 < k) y = msg(off + l + 12);
    P.permute(x, y);
  }
  if (l < 8) {
    alpha_sh[row][l] = x;
    if (live) a.dig[8 * b + l] = x;
  }
  __syncthreads();
  if (!live || l != 0) return;
  E6 al;
#pragma unroll
  for (int i = 0; i < 6; ++i) al.v[i] = alpha_sh[row][i];
  E6 pw = e6_one();
  for (int i = 0; i < a.n_bundles; ++i) {
#pragma unroll
    for (int q = 0; q < 6; ++q) a.scale[((size_t)b * a.n_bundles + i) * 6 + q] = pw.v[q];
    pw = e6_mul(pw, al);
  }
  for (int p = 0; p < a.n_points; ++p) {
    E6 c = e6_zero();
    for (int t = 0; t < a.n_terms; ++t) {
      const int3 tm = a.terms[t];  // (bundle, claimed value, point)
      if (tm.z != p) continue;
      c = e6_add(c, e6_mul(ldc6(a.scale + ((size_t)b * a.n_bundles + tm.x) * 6), ldc6(cv + 6 * tm.y)));
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) a.consts[((size_t)b * a.n_points + p) * 6 + q] = c.v[q];
  }
}

void launch_deep_alpha(const DeepAlphaArgs& a, const uint32_t* prefix, int n_prefix, int batch, hipStream_t s) {
  if (n_prefix < 1 || n_prefix > CHAL_MAX_PREFIX) throw std::runtime_error("deep alpha: bad prefix length");
  ChalPrefix pre{};
  for (int i = 0; i < n_prefix; ++i) pre.w[i] = prefix[i];
  DeepAlphaArgs args = a;
  args.n_pre = n_prefix;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_deep_alpha, dim3((batch + 3) / 4), dim3(64), 0, s, args, pre, batch);
  LG_HIP(hipGetLastError());
}

// Fiat-Shamir challenge per instance from a message of segments (transcript.go:
// 285-315): the sponge runs on a 16-lane row (four instances per 64-thread
// block), word i of segment k of instance b being seg[k].p[b stride + i].
// A lane's word indexes only grow, so each of its two word streams keeps a
// segment cursor.
struct FsCursor {
  int k = 0, base = 0;  // current segment and its first word index
  __device__ uint32_t word(const FsSeg* __restrict__ seg, int nseg, int b, int i) {
    while (k < nseg && i >= base + (int)seg[k].len) {
      base += (int)seg[k].len;
      ++k;
    }
    if (k >= nseg) return 0u;
    return seg[k].p[(size_t)b * seg[k].stride + (i - base)];
  }
};

__global__ __launch_bounds__(64) void k_fs_challenge(const FsSeg* __restrict__ seg, int nseg, int len, uint32_t* dig,
                                                     uint32_t* e6, int e6_stride, int batch) {
  const int l = threadIdx.x & 15, row = threadIdx.x >> 4;
  const int b_raw = blockIdx.x * 4 + row;
  const bool live = b_raw < batch;
  const int b = live ? b_raw : batch - 1;
  FsCursor ca, cb;
  P2Coop16 P;
  P.init(l);
  uint32_t x = 0u, y = 0u;
  for (int off = 0; off < len; off += 16) {
    const int k = len - off < 16 ? len - off : 16;
    if (l < 12 && l < k) x = ca.word(seg, nseg, b, off + l);
    if (l < 12 && l + 12 < k) y = cb.word(seg, nseg, b, off + l + 12);
    P.permute(x, y);
  }
  if (!live) return;
  if (l < 8) dig[8 * b + l] = x;
  if (e6 && l < 6) e6[(size_t)b * e6_stride + l] = x;
}

void launch_fs_challenge(const FsSeg* seg_dev, int nseg, int len, uint32_t* dig, uint32_t* e6, int e6_stride,
                         int batch, hipStream_t s) {
  if (nseg < 1 || len < 1) throw std::runtime_error("fs challenge: empty message");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_fs_challenge, dim3((batch + 3) / 4), dim3(64), 0, s, seg_dev, nseg, len, dig, e6, e6_stride,
                     batch);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ proof of work
// The salt search of one transcript challenge per instance (grindChallenge,
// transcript.go:317-332): the message is the segments (len words) followed by
// [FSPW, bits, salt]; found[b] ends as the smallest salt whose digest has
// `bits` zero low bits (hasZeroGrindingBits, :349-365, one 31-bit element for
// bits <= 31), or keeps ~0 if no salt below p has.  Workgroup w of instance b
// tries the 256-salt blocks w, w + W, w + 2W, ... in increasing order, one
// salt per thread, and stops once its block starts above found[b]: every
// salt below the final found[b] has then been tried by its workgroup.  The
// sponge state after the message's full blocks is the same for every salt;
// each trial overwrites the final block (the salt last) and permutes once.
__global__ __launch_bounds__(256) void k_fs_grind(const FsSeg* __restrict__ seg, int nseg, int len, uint32_t bits,
                                                  uint32_t* found) {
  const int b = blockIdx.y;
  const uint64_t step = (uint64_t)gridDim.x * 256u;
  const int total = len + 3, nfull = (total - 1) / 16, f0 = 16 * nfull, nlast = total - f0;
  const uint32_t tagw = kb_to_mont(TAG_FSPW_WORD), bitsw = kb_to_mont(bits);
  FsCursor cur;
  auto word = [&](int i) -> uint32_t { return i < len ? cur.word(seg, nseg, b, i) : (i == len ? tagw : bitsw); };
  uint32_t pre[24];
#pragma unroll
  for (int i = 0; i < 24; ++i) pre[i] = 0u;
  for (int blk = 0; blk < nfull; ++blk) {
#pragma unroll
    for (int i = 0; i < 16; ++i) pre[i] = word(16 * blk + i);
    p2_permute<24>(pre);
  }
  // the final block's words before the salt (lanes nlast - 1 .. 15 keep the state)
#pragma unroll
  for (int i = 0; i < 15; ++i)
    if (i < nlast - 1) pre[i] = word(f0 + i);
  const uint32_t mask = bits >= 31 ? 0x7fffffffu : ((1u << bits) - 1u);
  for (uint64_t base = (uint64_t)blockIdx.x * 256u; base < KB_P; base += step) {
    if (base > *(volatile uint32_t*)&found[b]) break;
    const uint32_t salt = (uint32_t)base + threadIdx.x;
    if (salt >= KB_P) continue;
    const uint32_t sw = kb_to_mont(salt);
    uint32_t st[24];
#pragma unroll
    for (int i = 0; i < 24; ++i) st[i] = (i == nlast - 1) ? sw : pre[i];
    p2_permute<24>(st);
    if ((kb_from_mont(st[0]) & mask) == 0u) atomicMin(&found[b], salt);
  }
}

// The trailer [FSPW, bits, salt] of each instance's message (Montgomery).
__global__ void k_grind_trailer(const uint32_t* __restrict__ found, uint32_t bits, uint32_t* __restrict__ tr,
                                int batch) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  const uint32_t salt = found[b];
  tr[3 * b] = kb_to_mont(TAG_FSPW_WORD);
  tr[3 * b + 1] = kb_to_mont(bits);
  tr[3 * b + 2] = kb_to_mont(salt < KB_P ? salt : 0u);
}

void launch_fs_grind(const FsSeg* seg_dev, int nseg, int len, uint32_t bits, uint32_t* found, uint32_t* trailer,
                     int workgroups, int batch, hipStream_t s) {
  if (nseg < 1 || len < 1 || bits < 1 || bits > 31 || workgroups < 1 || batch < 1)
    throw std::runtime_error("proof of work: bad arguments");
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_fs_grind, dim3(workgroups, batch), dim3(256), 0, s, seg_dev, nseg, len, bits, found);
  LG_HIP(hipGetLastError());
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_grind_trailer, dim3((batch + 63) / 64), dim3(64), 0, s, found, bits, trailer, batch);
  LG_HIP(hipGetLastError());
}

// out[b][k] = chal[b][ids[k]] (E6 slots): the challenge operands of one DAG.
__global__ void k_chal_compact(const uint32_t* __restrict__ chal, int ncols, const uint32_t* __restrict__ ids, int nch,
                               uint32_t* __restrict__ out) {
  const int b = blockIdx.x;
  for (int i = threadIdx.x; i < nch * 6; i += blockDim.x)
    out[((size_t)b * nch) * 6 + i] = chal[((size_t)b * ncols + ids[i / 6]) * 6 + i % 6];
}

void launch_chal_compact(const uint32_t* chal, int ncols, const uint32_t* ids, int nch, uint32_t* out, int batch,
                         hipStream_t s) {
  if (nch == 0) return;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_chal_compact, dim3(batch), dim3(64), 0, s, chal, ncols, ids, nch, out);
  LG_HIP(hipGetLastError());
}

// chal[b][col] = tot[b] * s (an E6 times a base-field scalar).
__global__ void k_scale_into(const uint32_t* __restrict__ tot, uint32_t s, uint32_t* __restrict__ chal, int ncols,
                             int col, int batch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * 6) return;
  const int b = i / 6, l = i % 6;
  chal[((size_t)b * ncols + col) * 6 + l] = kb_mul(tot[i], s);
}

void launch_scale_into(const uint32_t* tot, uint32_t s_mont, uint32_t* chal, int ncols, int col, int batch,
                       hipStream_t s) {
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_scale_into, dim3((batch * 6 + 255) / 256), dim3(256), 0, s, tot, s_mont, chal, ncols, col,
                     batch);
  LG_HIP(hipGetLastError());
}

// From zeta (first six words of zdig[b]): z[b] = zeta and, per module j,
// cm[j][b] = (zeta^N_j - 1) / N_j (the Lagrange-weight constant).
__global__ void k_zeta_consts(const uint32_t* __restrict__ zdig, const uint32_t* __restrict__ sizes,
                              const uint32_t* __restrict__ ninv, int nmod, uint32_t* __restrict__ z,
                              uint32_t* __restrict__ cm, int batch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * nmod) return;
  const int b = i % batch, j = i / batch;
  E6 zeta;
  for (int l = 0; l < 6; ++l) zeta.v[l] = zdig[8 * b + l];
  if (j == 0)
    for (int l = 0; l < 6; ++l) z[6 * b + l] = zeta.v[l];
  const E6 c = e6_scale(e6_sub(e6_pow(zeta, sizes[j]), e6_one()), ninv[j]);
  for (int l = 0; l < 6; ++l) cm[((size_t)j * batch + b) * 6 + l] = c.v[l];
}

void launch_zeta_consts(const uint32_t* zdig, const uint32_t* sizes, const uint32_t* ninv, int nmod, uint32_t* z,
                        uint32_t* cm, int batch, hipStream_t s) {
  if (nmod == 0) return;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_zeta_consts, dim3((batch * nmod + 63) / 64), dim3(64), 0, s, zdig, sizes, ninv, nmod, z, cm,
                     batch);
  LG_HIP(hipGetLastError());
}

// out[b][p] = zeta_b * w[p] (the DEEP points zeta * omega^shift).
__global__ void k_zeta_points(const uint32_t* __restrict__ zdig, const uint32_t* __restrict__ w, int np,
                              uint32_t* __restrict__ out, int batch) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= batch * np) return;
  const int b = i / np, p = i % np;
  E6 zeta;
  for (int l = 0; l < 6; ++l) zeta.v[l] = zdig[8 * b + l];
  const E6 r = e6_scale(zeta, w[p]);
  for (int l = 0; l < 6; ++l) out[(size_t)i * 6 + l] = r.v[l];
}

void launch_zeta_points(const uint32_t* zdig, const uint32_t* w, int np, uint32_t* out, int batch, hipStream_t s) {
  if (np == 0) return;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_zeta_points, dim3((batch * np + 63) / 64), dim3(64), 0, s, zdig, w, np, out, batch);
  LG_HIP(hipGetLastError());
}

void launch_fri_challenge(const uint32_t* prefix, int n_prefix, const uint32_t* prev, uint32_t* const* trees,
                          uint32_t* dig, uint32_t* alpha, uint32_t* roots, int batch, hipStream_t s) {
  if (n_prefix < 1 || n_prefix > CHAL_MAX_PREFIX) throw std::runtime_error("fri challenge: bad prefix length");
  ChalPrefix pre{};
  for (int i = 0; i < n_prefix; ++i) pre.w[i] = prefix[i];
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_fri_challenge, dim3((batch + 3) / 4), dim3(64), 0, s, pre, n_prefix, prev, trees, dig, alpha,
                     roots, batch);
  LG_HIP(hipGetLastError());
}

__global__ void k_fri_combine(uint32_t* const* running, const uint32_t* const* level, const uint32_t* gamma, int len) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  const int b = blockIdx.y;
  E6 r = ld6(running[b], len, i);
  r = e6_add(r, e6_mul(ld6(level[b], len, i), ldc6(gamma + 6 * b)));
  st6(running[b], len, i, r);
}

void launch_fri_combine(uint32_t* const* running, const uint32_t* const* level, const uint32_t* gamma, int len,
                        int batch, hipStream_t s) {
  const int T = 256;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_fri_combine, dim3((len + T - 1) / T, batch), dim3(T), 0, s, running, level, gamma, len);
  LG_HIP(hipGetLastError());
}

// ------------------------------------------------------------------ gathers
constexpr int GATHER_T = 256;

__global__ __launch_bounds__(GATHER_T) void k_gather(const GatherOp* __restrict__ ops, int n_ops,
                                                     const uint32_t* __restrict__ pos, int nq, int b0, int nb,
                                                     uint32_t* __restrict__ out, int stride) {
  int k = 0;
  while (k + 1 < n_ops && (int)blockIdx.x >= ops[k + 1].block0) ++k;
  const GatherOp op = ops[k];
  const long u = (long)((int)blockIdx.x - op.block0) * GATHER_T + threadIdx.x;
  const int per = op.kind == 0 ? 2 * op.ncols : (op.kind == 1 ? op.ncols : 1);
  const long total = (long)nb * nq * per;
  if (u >= total) return;
  const int bq = b0 * nq + (int)(u / per);
  const int e = (int)(u % per);
  const int b = bq / nq;
  const uint32_t row = pos[bq] >> op.shift;
  const uint32_t lo = row & ~1u;
  uint32_t* o = out + (size_t)bq * stride + op.out_off;
  if (op.kind == 0) {
    const int c = e % op.ncols, hi = e / op.ncols;
    o[e] = op.ptrs[(size_t)b * op.ncols + c][lo + hi];
  } else if (op.kind == 1) {
    const uint32_t node = (((uint32_t)op.nleaves + (lo >> 1)) >> e) ^ 1u;
    const uint4* src = reinterpret_cast<const uint4*>(op.ptrs[b] + 8 * (size_t)node);
    uint4* dst = reinterpret_cast<uint4*>(o + 8 * e);
    uint4 a = src[0], c = src[1];
    dst[0] = a;
    dst[1] = c;
  } else {
    const uint4* src = reinterpret_cast<const uint4*>(op.ptrs[b] + 8 * (size_t)((lo >> 1) >> op.d));
    uint4* dst = reinterpret_cast<uint4*>(o);
    uint4 a = src[0], c = src[1];
    dst[0] = a;
    dst[1] = c;
  }
}

void launch_gather(const GatherOp* ops_dev, int n_ops, int total_blocks, const uint32_t* pos, int nq, int b0, int nb,
                   uint32_t* out, int stride, hipStream_t s) {
  if (n_ops == 0 || total_blocks == 0 || nb == 0) return;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_gather, dim3(total_blocks), dim3(GATHER_T), 0, s, ops_dev, n_ops, pos, nq, b0, nb, out, stride);
  LG_HIP(hipGetLastError());
}

__global__ void k_gather_rows(const uint32_t* pos, int nq, int shift, const uint32_t* const* cols, int ncols,
                              uint32_t* out, int stride, int off) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int b = blockIdx.y;
  const uint32_t row = pos[(size_t)b * nq + q] >> shift;
  const uint32_t lo = row & ~1u;
  uint32_t* o = out + ((size_t)b * nq + q) * stride + off;
  const uint32_t* const* cp = cols + (size_t)b * ncols;
  for (int c = 0; c < ncols; ++c) {
    o[c] = cp[c][lo];
    o[ncols + c] = cp[c][lo + 1];
  }
}

void launch_gather_rows(const uint32_t* pos, int nq, int shift, const uint32_t* const* cols, int ncols, uint32_t* out,
                        int stride, int off, int batch, hipStream_t s) {
  const int T = 64;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_gather_rows, dim3((nq + T - 1) / T, batch), dim3(T), 0, s, pos, nq, shift, cols, ncols, out,
                     stride, off);
  LG_HIP(hipGetLastError());
}

__global__ void k_gather_path(const uint32_t* pos, int nq, int shift, const uint32_t* const* trees, int nleaves,
                              uint32_t* out, int stride, int off) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int b = blockIdx.y;
  const uint32_t row = pos[(size_t)b * nq + q] >> shift;
  uint32_t node = (uint32_t)nleaves + ((row & ~1u) >> 1);
  const uint32_t* nodes = trees[b];
  uint32_t* o = out + ((size_t)b * nq + q) * stride + off;
  for (int k = 0; node > 1; ++k, node >>= 1) {
    const uint32_t* src = nodes + 8 * (size_t)(node ^ 1u);
#pragma unroll
    for (int l = 0; l < 8; ++l) o[8 * k + l] = src[l];
  }
}

void launch_gather_path(const uint32_t* pos, int nq, int shift, const uint32_t* const* trees, int nleaves,
                        uint32_t* out, int stride, int off, int batch, hipStream_t s) {
  const int T = 64;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_gather_path, dim3((nq + T - 1) / T, batch), dim3(T), 0, s, pos, nq, shift, trees, nleaves, out,
                     stride, off);
  LG_HIP(hipGetLastError());
}

__global__ void k_gather_digest(const uint32_t* pos, int nq, int shift, int d, const uint32_t* const* inj,
                                uint32_t* out, int stride, int off) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int b = blockIdx.y;
  const uint32_t row = pos[(size_t)b * nq + q] >> shift;
  const uint32_t leaf = ((row & ~1u) >> 1) >> d;
  const uint32_t* src = inj[b] + 8 * (size_t)leaf;
  uint32_t* o = out + ((size_t)b * nq + q) * stride + off;
#pragma unroll
  for (int l = 0; l < 8; ++l) o[l] = src[l];
}

void launch_gather_digest(const uint32_t* pos, int nq, int shift, int d, const uint32_t* const* inj, uint32_t* out,
                          int stride, int off, int batch, hipStream_t s) {
  const int T = 64;
  if (!prelaunch()) return;
  hipLaunchKernelGGL(k_gather_digest, dim3((nq + T - 1) / T, batch), dim3(T), 0, s, pos, nq, shift, d, inj, out,
                     stride, off);
  LG_HIP(hipGetLastError());
}

}  // namespace lg

namespace lg {

// ------------------------------------------------------------------ issue-rate probe
// Eight independent dependency chains per lane of one instruction class, so
// that issue (not latency) bounds the loop: KIND 0 v_add_u32, 1 v_mul_lo_u32,
// 2 v_mad_u64_u32 (the high word of x*y + x feeds the next step), 3 the full
// Montgomery product kb_mul.  Each lane stores its chains' xor so that no
// step is dead.
template <int KIND>
__global__ __launch_bounds__(256) void k_issue_probe(const uint32_t* __restrict__ seed, uint32_t* __restrict__ out,
                                                     int iters) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t x[8];
  uint64_t acc[8];
  const uint32_t y = seed[8] | 1u;
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    x[i] = seed[i] ^ (uint32_t)t;
    acc[i] = x[i];
  }
  for (int k = 0; k < iters; ++k) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      if constexpr (KIND == 0) {
        x[i] = x[i] + y;
      } else if constexpr (KIND == 1) {
        x[i] = x[i] * y;
      } else if constexpr (KIND == 2) {
        // acc = x * y + acc, one v_mad_u64_u32 per step; each chain's carry-out
        // goes to an SGPR pair of its own (no write-after-write waits)
#define LG_PROBE_MAD(I, S0, S1)                                                                                 \
  if (i == I)                                                                                                   \
    asm volatile("v_mad_u64_u32 %0, s[" #S0 ":" #S1 "], %1, %2, %0" : "+v"(acc[I]) : "v"(x[I]), "v"(y)     \
                 : "s" #S0, "s" #S1);
        LG_PROBE_MAD(0, 64, 65) LG_PROBE_MAD(1, 66, 67) LG_PROBE_MAD(2, 68, 69) LG_PROBE_MAD(3, 70, 71)
        LG_PROBE_MAD(4, 72, 73) LG_PROBE_MAD(5, 74, 75) LG_PROBE_MAD(6, 76, 77) LG_PROBE_MAD(7, 78, 79)
#undef LG_PROBE_MAD
      } else {
        x[i] = kb_mul(x[i], y);
      }
    }
  }
  uint32_t r = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) r ^= x[i] ^ (uint32_t)acc[i] ^ (uint32_t)(acc[i] >> 32);
  out[t] = r;
}

void issue_probe(int device, double* gops, double* clock_mhz) {
  int clk_khz = 0, cus = 0;
  LG_HIP(hipSetDevice(device));
  LG_HIP(hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, device));
  LG_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  *clock_mhz = clk_khz / 1000.0;
  const int T = 256, blocks = cus * 16, iters = 4096;
  const size_t lanes = (size_t)T * blocks;
  DevBuf seed(64), out(lanes * 4);
  const uint32_t sv[16] = {0x1234567u, 0x2345678u, 0x3456789u, 0x456789au, 0x56789abu, 0x6789abcu,
                           0x789abcdu, 0x89abcdeu, 0x7654321u};
  LG_HIP(hipMemcpy(seed.as<void>(), sv, 64, hipMemcpyHostToDevice));
  hipEvent_t a, b;
  LG_HIP(hipEventCreate(&a));
  LG_HIP(hipEventCreate(&b));
  for (int kind = 0; kind < 4; ++kind) {
    auto run = [&] {
      switch (kind) {
        case 0: hipLaunchKernelGGL(k_issue_probe<0>, dim3(blocks), dim3(T), 0, 0, seed.as<uint32_t>(), out.as<uint32_t>(), iters); break;
        case 1: hipLaunchKernelGGL(k_issue_probe<1>, dim3(blocks), dim3(T), 0, 0, seed.as<uint32_t>(), out.as<uint32_t>(), iters); break;
        case 2: hipLaunchKernelGGL(k_issue_probe<2>, dim3(blocks), dim3(T), 0, 0, seed.as<uint32_t>(), out.as<uint32_t>(), iters); break;
        default: hipLaunchKernelGGL(k_issue_probe<3>, dim3(blocks), dim3(T), 0, 0, seed.as<uint32_t>(), out.as<uint32_t>(), iters); break;
      }
      LG_HIP(hipGetLastError());
    };
    run();  // warm-up
    LG_HIP(hipEventRecord(a, 0));
    for (int r = 0; r < 5; ++r) run();
    LG_HIP(hipEventRecord(b, 0));
    LG_HIP(hipEventSynchronize(b));
    float ms = 0;
    LG_HIP(hipEventElapsedTime(&ms, a, b));
    gops[kind] = 5.0 * lanes * iters * 8 / (ms * 1e-3) / 1e9;  // lane operations per ns
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
}

}  // namespace lg
