// Per-kernel timing with HIP events on the launching stream, plus a device
// counter per launch for the algorithmic work the kernel reports.
#pragma once

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "device_buffer.hpp"

namespace mirorr {

constexpr int kProfFast = 1;    // MIRORR_PROF_FAST
constexpr int kProfMasked = 2;  // MIRORR_PROF_MASKED
constexpr int kProfFinest = 4;  // MIRORR_PROF_FINEST_ONLY
// A launch's work counter is a stripe of kWorkStripe counters: workgroup b
// adds to entry b % kWorkStripe, so the updates do not contend on one address.
constexpr int kWorkStripe = 64;

struct ProfAgg {
  int64_t launches = 0;
  double ms = 0.0;
  double units = 0.0;  // algorithmic work units reported by the kernel
};

class Profiler {
 public:
  ~Profiler() { release(); }
  bool on() const { return (mask_ & (kProfFast | kProfMasked)) != 0; }
  // Registrations instrument only their last (full-resolution) level.
  bool finest_only() const { return (mask_ & kProfFinest) != 0; }
  // `mask`: the kernels (bits) whose launches are instrumented; 0 = off.
  void enable(int mask, Budget* bud) {
    if (mask && !counters_.bytes()) {
      counters_.ensure(sizeof(unsigned long long) * kSlots * kWorkStripe, bud);
      MIRORR_HIP(hipMemset(counters_.as<void>(), 0,
                           sizeof(unsigned long long) * kSlots * kWorkStripe));
    }
    mask_ = mask;
  }
  // Records the start event; returns the launch's work counter stripe
  // (device) or nullptr when this kernel is not instrumented.
  unsigned long long* begin(const char* name, int bit, hipStream_t s) {
    if (!(mask_ & bit)) return nullptr;
    if (used_ >= kSlots) collect(s);
    Slot sl;
    sl.name = name;
    sl.start = take_event();
    sl.stop = take_event();
    sl.counter = used_++;  // zero: slots are cleared together after each collection
    MIRORR_HIP(hipEventRecord(sl.start, s));
    pending_.push_back(sl);
    if (std::find(streams_.begin(), streams_.end(), s) == streams_.end()) streams_.push_back(s);
    return counters_.as<unsigned long long>() + (size_t)sl.counter * kWorkStripe;
  }
  void end(int bit, hipStream_t s) {
    if (!(mask_ & bit) || pending_.empty()) return;
    MIRORR_HIP(hipEventRecord(pending_.back().stop, s));
  }
  // Waits for pending launches (on every stream they were recorded on) and
  // folds them into the aggregates.
  void collect(hipStream_t s) {
    if (pending_.empty()) return;
    for (hipStream_t x : streams_) MIRORR_HIP(hipStreamSynchronize(x));
    std::vector<unsigned long long> c((size_t)used_ * kWorkStripe);
    MIRORR_HIP(hipMemcpy(c.data(), counters_.as<unsigned long long>(),
                         sizeof(unsigned long long) * c.size(), hipMemcpyDeviceToHost));
    for (Slot& sl : pending_) {
      float ms = 0.f;
      MIRORR_HIP(hipEventElapsedTime(&ms, sl.start, sl.stop));
      ProfAgg& a = agg_[sl.name];
      a.launches += 1;
      a.ms += ms;
      for (int k = 0; k < kWorkStripe; ++k)
        a.units += (double)c[(size_t)sl.counter * kWorkStripe + k];
      free_.push_back(sl.start);
      free_.push_back(sl.stop);
    }
    pending_.clear();
    // Cleared before any stream launches with these counters again.
    MIRORR_HIP(hipMemsetAsync(counters_.as<void>(), 0,
                              sizeof(unsigned long long) * kWorkStripe * used_, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    used_ = 0;
    streams_.clear();
  }
  const std::map<std::string, ProfAgg>& aggregates() const { return agg_; }
  void reset() { agg_.clear(); }
  void release() {
    for (hipEvent_t e : free_) (void)hipEventDestroy(e);
    for (Slot& s : pending_) {
      (void)hipEventDestroy(s.start);
      (void)hipEventDestroy(s.stop);
    }
    free_.clear();
    pending_.clear();
  }

 private:
  std::vector<hipStream_t> streams_;  // streams with pending events
  static constexpr int kSlots = 4096;
  struct Slot {
    std::string name;
    hipEvent_t start, stop;
    int counter;
  };
  hipEvent_t take_event() {
    if (!free_.empty()) {
      hipEvent_t e = free_.back();
      free_.pop_back();
      return e;
    }
    hipEvent_t e;
    MIRORR_HIP(hipEventCreate(&e));
    return e;
  }
  int mask_ = 0;
  int used_ = 0;
  DeviceBuffer counters_;
  std::vector<Slot> pending_;
  std::vector<hipEvent_t> free_;
  std::map<std::string, ProfAgg> agg_;
};

}  // namespace mirorr
