// HIP helpers shared by the prover's host code.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace lg {

#define LG_HIP(call)                                                                        \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " at " + \
                               __FILE__ + ":" + std::to_string(__LINE__) + ": " #call);     \
  } while (0)

// Owning device allocation.
class DevBuf {
 public:
  DevBuf() = default;
  explicit DevBuf(size_t bytes) { alloc(bytes); }
  ~DevBuf() { release(); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), owned_(o.owned_) { o.p_ = nullptr; o.n_ = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) {
      release();
      p_ = o.p_; n_ = o.n_; owned_ = o.owned_;
      o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  // A non-owning handle of memory owned elsewhere (an arena lends its buffers
  // this way); releasing it frees nothing.
  static DevBuf view(void* p, size_t bytes) {
    DevBuf b;
    b.p_ = p;
    b.n_ = bytes;
    b.owned_ = false;
    return b;
  }
  void alloc(size_t bytes) {
    release();
    if (bytes) LG_HIP(hipMalloc(&p_, bytes));
    n_ = bytes;
    owned_ = true;
  }
  void release() {
    if (p_ && owned_) (void)hipFree(p_);
    p_ = nullptr;
    n_ = 0;
  }
  template <class T = uint32_t>
  T* as() const { return reinterpret_cast<T*>(p_); }
  size_t bytes() const { return n_; }
  bool owned() const { return owned_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
  bool owned_ = true;
};

// Device copy of a small host vector (kernel argument tables).  The copy is
// synchronous, so the host vector may be released as soon as this returns.
template <class T>
inline DevBuf upload(const std::vector<T>& v) {
  DevBuf b(v.size() * sizeof(T));
  if (!v.empty()) LG_HIP(hipMemcpy(b.as<void>(), v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return b;
}

}  // namespace lg
