// engine_buf.hpp — error type and owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <mutex>
#include <stdexcept>
#include <string>

#include "common.hpp"

namespace topohip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Device allocation, freeing, pinned allocation and stream creation may
// synchronise with other streams; several problems (the serving loop) capture
// PCG graphs on their own threads. Those calls and every graph capture hold
// this process-wide lock, so none of them lands inside another thread's capture.
inline std::recursive_mutex& device_api_mutex() {
  static std::recursive_mutex m;
  return m;
}

template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) {
      std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
      (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  void ensure(size_t m) {
    if (m <= n && p) return;
    std::lock_guard<std::recursive_mutex> lk(device_api_mutex());
    release();
    TOPOHIP_CHECK(hipMalloc(&p, (m > 0 ? m : 1) * sizeof(T)));
    n = m;
  }
  size_t bytes() const { return n * sizeof(T); }
};

}  // namespace topohip
