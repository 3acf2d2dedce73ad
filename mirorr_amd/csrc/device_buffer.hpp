// RAII device allocations charged against a context's HBM budget.
#pragma once

#include <atomic>
#include <cstddef>
#include <utility>

#include "common.hpp"

namespace mirorr {

// Bytes charged against a limit. A budget with a parent charges the parent
// too, so several contexts can share one limit (a server's worker slots).
struct Budget {
  size_t limit = 0;  // 0 = unlimited
  std::atomic<size_t> used{0};
  std::atomic<size_t> peak{0};
  Budget* parent = nullptr;
  void charge(size_t n) {
    if (parent) parent->charge(n);
    const size_t now = used.fetch_add(n) + n;
    if (limit != 0 && now > limit) {
      used.fetch_sub(n);
      if (parent) parent->release(n);
      throw Error(kErrNoMem, "HBM budget exceeded");
    }
    size_t p = peak.load();
    while (now > p && !peak.compare_exchange_weak(p, now)) {
    }
  }
  void release(size_t n) {
    used.fetch_sub(n);
    if (parent) parent->release(n);
  }
};

class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept { *this = std::move(o); }
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    if (this != &o) {
      reset();
      ptr_ = o.ptr_;
      bytes_ = o.bytes_;
      budget_ = o.budget_;
      o.ptr_ = nullptr;
      o.bytes_ = 0;
      o.budget_ = nullptr;
    }
    return *this;
  }
  ~DeviceBuffer() { reset(); }

  // Ensures at least `bytes` bytes; keeps the allocation when large enough
  // (buffers are reused while sizes do not grow).
  void ensure(size_t bytes, Budget* budget) {
    if (bytes <= bytes_) return;
    reset();
    if (bytes == 0) return;
    if (budget) budget->charge(bytes);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      if (budget) budget->release(bytes);
      throw Error(e == hipErrorOutOfMemory ? kErrNoMem : kErrPlatform,
                  std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    ptr_ = p;
    bytes_ = bytes;
    budget_ = budget;
  }
  void reset() {
    if (ptr_) {
      (void)hipFree(ptr_);
      if (budget_) budget_->release(bytes_);
    }
    ptr_ = nullptr;
    bytes_ = 0;
    budget_ = nullptr;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(ptr_);
  }
  size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
  Budget* budget_ = nullptr;
};

}  // namespace mirorr
