// engine_buf.hpp — error type and owning device buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "common.hpp"

namespace topohip {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

template <typename T>
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
  void ensure(size_t m) {
    if (m <= n && p) return;
    release();
    TOPOHIP_CHECK(hipMalloc(&p, (m > 0 ? m : 1) * sizeof(T)));
    n = m;
  }
  size_t bytes() const { return n * sizeof(T); }
};

}  // namespace topohip
