// Device memory helpers shared by the host modules.
#pragma once
#include <vector>

#include "chm_internal.hpp"

namespace chm {

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw Error{ERR_HIP, std::string(what) + ": " + hipGetErrorString(e)};
}

// Owning device allocation.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) return;
    hip_check(hipMalloc((void**)&p, count * sizeof(T)), "hipMalloc");
    n = count;
  }
  // Synchronous copy: the host vector may be released right after the call.
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) {
      hip_check(hipStreamSynchronize(s), "upload");
      hip_check(hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    }
  }
};

}  // namespace chm
