// Context object behind the C-ABI handle mirorr_ctx.
#pragma once

#include <memory>
#include <mutex>
#include <string>

#include "../../include/mirorr_gpu.h"
#include "device_buffer.hpp"
#include "profiler.hpp"

namespace mirorr {
struct RegCache;    // registration workspace, defined with the driver
struct BatchCache;  // batched registrations' workspaces, likewise

// Upper bound of the device bytes one mirorr_register call with these
// parameters and images (volumes of `storage`) charges to a context whose
// cache is empty, the two volumes (and masks) included. Throws on
// parameters whose pyramid schedule is empty.
size_t registration_bytes_bound(const mirorr_reg_params* p, const Geom& fixed, bool fixed_mask,
                                const Geom& moving, bool moving_mask, int storage = 0);
}

struct mirorr_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // Second stream and fork/join events of a registration: the two
  // block-matching passes of a symmetric iteration run on one each.
  // Created on first use, default priority.
  hipStream_t stream2 = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  // Whether registrations fork the second pass onto stream2. A server's
  // worker contexts do not: their requests already overlap, and more
  // streams than hardware queues measured slower.
  bool overlap_passes = true;
  // Whether a registration replays each level's outer iterations after the
  // first from a HIP graph captured at the second (mirorr_ctx_set_iteration_graphs).
  bool iteration_graphs = false;
  mirorr::Budget budget;
  std::string last_error;
  std::mutex mu;
  mirorr::Profiler prof;  // optional per-kernel event timing
  // Device workspace, pyramid levels and pinned flags of registrations run on
  // this context: grown to the largest request and reused, so a registration
  // after the first performs no device allocation. Guarded by `mu`.
  std::shared_ptr<mirorr::RegCache> reg;
  // Per-pair caches of mirorr_register_batch, likewise grown and reused.
  std::shared_ptr<mirorr::BatchCache> batch;
  // Pinned staging chunks of volume uploads (double-buffered), created on
  // first use and freed with the context.
  void* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
};

namespace mirorr {

// Makes the context's device current for the calling host thread.
inline void use_device(mirorr_ctx* c) { MIRORR_HIP(hipSetDevice(c->device)); }

// Runs f(), converting library errors into a status recorded on the context.
template <class F>
mirorr_status guarded(mirorr_ctx* c, F&& f) {
  try {
    if (c) use_device(c);
    f();
    return MIRORR_OK;
  } catch (const Error& e) {
    if (c) c->last_error = e.what();
    return static_cast<mirorr_status>(e.code);
  } catch (const std::bad_alloc&) {
    if (c) c->last_error = "host allocation failed";
    return MIRORR_ERR_NOMEM;
  } catch (const std::exception& e) {
    if (c) c->last_error = e.what();
    return MIRORR_ERR_PLATFORM;
  }
}

inline Geom to_geom(const mirorr_geom& g) {
  Geom o{};
  for (int i = 0; i < 3; ++i) {
    o.dims[i] = g.dims[i];
    o.spacing[i] = g.spacing[i];
    o.origin[i] = g.origin[i];
  }
  for (int i = 0; i < 9; ++i) o.dir[i] = g.direction[i];
  return o;
}

}  // namespace mirorr
