// Registration server: requests admitted onto worker slots of one GPU by an
// HBM budget, their work overlapping on the device.
//
// Each slot owns a mirorr_ctx (its own HIP stream and registration cache)
// and a host thread that runs mirorr_register for the request it is given.
// Kernels of different slots' streams run concurrently, so a request that
// arrives while others run starts at once if a slot is free and its memory
// fits; the host round trips of one request (LTS flag reads, uploads) are
// covered by the others' kernels. A slot's context keeps its workspace
// between requests (grown to the largest it has run), so repeated requests
// of similar size allocate nothing.
//
// Admission (under the server lock): a slot running a request may hold up to
// the bytes it held when admitted plus the request's bound (each workspace
// buffer only grows to the new size), an idle slot holds its kept workspace.
// A request is admitted onto an idle slot when the sum stays within the
// budget; if not, idle slots' workspaces are released first; if it still does
// not fit, it waits for a running request to finish, or fails with
// MIRORR_ERR_NOMEM when none is running and nothing waits before it. Requests
// are taken in arrival order, but later requests that fit may overtake one
// that waits for memory, at most kMaxOvertake times.
//
// Batching (mirorr_server_set_batch): a slot that takes a request also takes
// the later pending requests with the same parameters (the NCC tiled search,
// the same storage and initialisation kind) that fit, up to the batch size,
// and runs them together through mirorr_register_batch: their iterations
// share launches (one search launch per pass over every pair's tiles, one
// per LTS phase). Each result equals the request's alone.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "bm_kernels.hpp"
#include "context.hpp"

namespace {

using mirorr::Error;

thread_local std::string g_server_error;

constexpr int kMaxSlots = 64;
constexpr int kMaxOvertake = 8;  // admissions that may pass a request waiting for memory

struct Request {
  int64_t ticket = 0;
  mirorr_reg_params p{};
  mirorr_geom gF{}, gM{};
  std::vector<float> F, M;  // copies (empty when borrowed)
  std::vector<uint8_t> Fm, Mm;
  const float* pF = nullptr;  // what the worker uploads: the copies or the caller's arrays
  const float* pM = nullptr;
  const uint8_t* pFm = nullptr;
  const uint8_t* pMm = nullptr;
  bool has_init = false, has_center = false;
  double init[12] = {0}, center[3] = {0};
  int storage = MIRORR_STORAGE_F32;
  size_t bound = 0;
  double t_submit = 0.0;
  int overtaken = 0;  // later requests admitted while this one waited for memory
};

struct Slot {
  mirorr_ctx* ctx = nullptr;
  std::thread th;
  std::vector<std::unique_ptr<Request>> jobs;  // handed over by the scheduler (a batch)
  bool busy = false;
  size_t reserve = 0;  // bytes the slot may hold while its requests run
};

// Requests mirorr_register_batch can run together: the NCC metric with the
// tiled search, equal parameters, one storage, the same kind of start.
bool batchable(const Request& r) {
  const mirorr_reg_params& p = r.p;
  if (p.block_metric != MIRORR_METRIC_NC) return false;
  const int dims[3] = {64, 64, 64};
  return mirorr::tiled_supported(
      mirorr::make_grid(dims, p.block_width, p.block_gap, p.nhood_width, p.nhood_gap));
}
bool same_batch(const Request& a, const Request& b) {
  const mirorr_reg_params &x = a.p, &y = b.p;
  return x.symmetric == y.symmetric && x.affine == y.affine && x.block_width == y.block_width &&
         x.block_gap == y.block_gap && x.nhood_width == y.nhood_width &&
         x.nhood_gap == y.nhood_gap && x.iterations == y.iterations &&
         x.pyr_start == y.pyr_start && x.pyr_end == y.pyr_end && x.pyr_num == y.pyr_num &&
         x.pyr_min_size == y.pyr_min_size && x.portion_kept == y.portion_kept &&
         x.variance_select == y.variance_select && x.verbosity == y.verbosity &&
         x.centre_init == y.centre_init && x.stop_tolerance == y.stop_tolerance &&
         x.lts_tolerance == y.lts_tolerance && x.block_metric == y.block_metric &&
         x.resampling == y.resampling && a.storage == b.storage &&
         a.has_init == b.has_init && a.has_center == b.has_center;
}

size_t checked_nvox(const mirorr_geom& g) {
  size_t n = 1;
  for (int a = 0; a < 3; ++a) {
    if (g.dims[a] < 1 || !(g.spacing[a] > 0.0)) throw Error(mirorr::kErrArg, "bad geometry");
    n *= (size_t)g.dims[a];
  }
  return n;
}

}  // namespace

struct mirorr_server {
  int device = 0;
  mirorr::Budget budget;  // every slot's context charges it
  std::mutex mu;
  std::condition_variable cv_work;  // slots wait for a request
  std::condition_variable cv_done;  // pollers wait for results
  std::condition_variable cv_idle;  // destroy waits for the pollers to leave
  int pollers = 0;                  // threads inside mirorr_poll
  std::deque<std::unique_ptr<Request>> pending;
  std::map<int64_t, mirorr_request_result> done;
  std::map<int64_t, bool> live;  // tickets submitted and not yet handed out
  std::vector<Slot> slots;
  int64_t next_ticket = 1;
  int64_t completed = 0;
  int max_running = 0;
  size_t reserve_peak = 0;  // largest sum of the slots' reservations at an admission
  int batch_max = 1;        // requests a slot may take at once
  int max_batch = 0;        // largest batch run so far
  bool stopping = false;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  size_t held(const Slot& s) const { return s.busy ? s.reserve : s.ctx->budget.used.load(); }

  void finish_unadmitted(std::unique_ptr<Request> r, mirorr_status st, const std::string& msg) {
    mirorr_request_result res{};
    res.status = st;
    res.t_submit = r->t_submit;
    res.t_start = res.t_end = now();
    res.slot = -1;
    res.bytes_bound = r->bound;
    std::snprintf(res.error, sizeof(res.error), "%s", msg.c_str());
    done[r->ticket] = res;
    ++completed;
    cv_done.notify_all();
  }

  // Admits pending requests (caller holds mu). The queue is scanned in
  // arrival order; a request that does not fit in memory waits while later
  // ones that fit are admitted, until it has been overtaken kMaxOvertake
  // times: from then on nothing behind it is admitted before it.
  void schedule() {
    const size_t limit = budget.limit;
    auto running = [&] {
      for (const Slot& x : slots)
        if (x.busy) return true;
      return false;
    };
    auto fits = [&](const Request& r) {
      if (limit == 0) return true;
      size_t sum = 0;
      for (const Slot& x : slots) sum += held(x);
      return sum + r.bound <= limit;  // the chosen slot: its kept workspace plus the bound
    };
    // Releases the workspace of the idle slot keeping the largest one (on
    // this server's device; the calling thread's device is restored).
    // Returns false when no idle slot keeps a workspace.
    auto release_one_idle = [&] {
      Slot* big = nullptr;
      for (Slot& x : slots)
        if (!x.busy && x.jobs.empty() && (x.ctx->reg || x.ctx->batch) &&
            (!big || x.ctx->budget.used.load() > big->ctx->budget.used.load()))
          big = &x;
      if (!big) return false;
      int prev = device;
      (void)hipGetDevice(&prev);
      (void)hipSetDevice(device);
      {
        std::lock_guard<std::mutex> lk(big->ctx->mu);
        big->ctx->reg.reset();
        big->ctx->batch.reset();
      }
      (void)hipSetDevice(prev);
      return true;
    };
    std::vector<Request*> waiting;  // requests passed over in this scan
    for (auto it = pending.begin(); it != pending.end() && !stopping;) {
      int w = -1;
      for (size_t i = 0; i < slots.size(); ++i)
        if (!slots[i].busy && slots[i].jobs.empty()) {
          // prefer the idle slot keeping the largest workspace
          if (w < 0 || slots[i].ctx->budget.used.load() > slots[w].ctx->budget.used.load())
            w = (int)i;
        }
      if (w < 0) return;
      Request& r = **it;
      // Idle workspaces are released one at a time, largest first, only
      // until the request fits.
      while (!fits(r) && release_one_idle()) {
      }
      if (!fits(r)) {
        if (!running() && waiting.empty()) {
          const std::string msg = "request needs up to " + std::to_string(r.bound) +
                                  " device bytes, beyond the server's HBM budget of " +
                                  std::to_string(limit);
          std::unique_ptr<Request> rr = std::move(*it);
          it = pending.erase(it);
          finish_unadmitted(std::move(rr), MIRORR_ERR_NOMEM, msg);
          continue;
        }
        if (r.overtaken >= kMaxOvertake) return;  // keep its place: admit nothing behind it
        waiting.push_back(&r);
        ++it;
        continue;
      }
      for (Request* q : waiting) ++q->overtaken;
      Slot& s = slots[w];
      // The slot keeps its workspace and each buffer only grows, so while
      // the request runs the slot holds at most its kept bytes plus the
      // request's bound (the sum of per-buffer maxima can exceed the larger
      // of the two).
      s.reserve = s.ctx->budget.used.load() + r.bound;
      s.busy = true;
      s.jobs.clear();
      s.jobs.push_back(std::move(*it));
      it = pending.erase(it);
      if (batch_max > 1 && batchable(*s.jobs[0])) {
        // Later requests that run with it and fit join the batch (indices:
        // erasing from a deque invalidates its iterators).
        const size_t pos = (size_t)(it - pending.begin());
        size_t sum = 0;
        for (const Slot& x : slots) sum += held(x);
        for (size_t k = pos; k < pending.size() && (int)s.jobs.size() < batch_max;) {
          Request& q = *pending[k];
          if (same_batch(*s.jobs[0], q) && (limit == 0 || sum + q.bound <= limit)) {
            sum += q.bound;
            s.reserve += q.bound;
            for (Request* wq : waiting) ++wq->overtaken;
            s.jobs.push_back(std::move(pending[k]));
            pending.erase(pending.begin() + (std::ptrdiff_t)k);
          } else {
            ++k;
          }
        }
        it = pending.begin() + (std::ptrdiff_t)pos;
      }
      int n = 0;
      size_t sum = 0;
      for (const Slot& x : slots) {
        n += x.busy ? 1 : 0;
        sum += held(x);
      }
      if (n > max_running) max_running = n;
      if (sum > reserve_peak) reserve_peak = sum;
      cv_work.notify_all();
    }
  }

  void run(Slot& slot, int w, Request& r, mirorr_request_result& res) {
    res.t_submit = r.t_submit;
    res.t_start = now();
    res.slot = w;
    res.bytes_bound = r.bound;
    mirorr_ctx* c = slot.ctx;
    mirorr_volume* vf = nullptr;
    mirorr_volume* vm = nullptr;
    mirorr_status st = mirorr_volume_create_storage(c, r.pF, r.pFm, &r.gF, r.storage, &vf);
    if (st == MIRORR_OK) st = mirorr_volume_create_storage(c, r.pM, r.pMm, &r.gM, r.storage, &vm);
    // The host copies are not needed once uploaded.
    std::vector<float>().swap(r.F);
    std::vector<float>().swap(r.M);
    std::vector<uint8_t>().swap(r.Fm);
    std::vector<uint8_t>().swap(r.Mm);
    if (st == MIRORR_OK)
      st = mirorr_register(c, &r.p, vf, vm, r.has_init ? r.init : nullptr,
                           r.has_center ? r.center : nullptr, res.params, res.center, &res.stats);
    if (st != MIRORR_OK) std::snprintf(res.error, sizeof(res.error), "%s", mirorr_last_error(c));
    mirorr_volume_destroy(vf);
    mirorr_volume_destroy(vm);
    res.status = st;
    res.t_end = now();
  }

  // Several requests through mirorr_register_batch; their results carry the
  // batch's start and end times.
  void run_batch(Slot& slot, int w, std::vector<std::unique_ptr<Request>>& jobs,
                 std::vector<mirorr_request_result>& res) {
    if (jobs.size() == 1) {
      run(slot, w, *jobs[0], res[0]);
      return;
    }
    mirorr_ctx* c = slot.ctx;
    const int n = (int)jobs.size();
    const double t_start = now();
    std::vector<mirorr_volume*> vf(n, nullptr), vm(n, nullptr);
    std::vector<int> ok;
    for (int i = 0; i < n; ++i) {
      Request& r = *jobs[i];
      res[i].t_submit = r.t_submit;
      res[i].t_start = t_start;
      res[i].slot = w;
      res[i].bytes_bound = r.bound;
      mirorr_status st = mirorr_volume_create_storage(c, r.pF, r.pFm, &r.gF, r.storage, &vf[i]);
      if (st == MIRORR_OK)
        st = mirorr_volume_create_storage(c, r.pM, r.pMm, &r.gM, r.storage, &vm[i]);
      std::vector<float>().swap(r.F);
      std::vector<float>().swap(r.M);
      std::vector<uint8_t>().swap(r.Fm);
      std::vector<uint8_t>().swap(r.Mm);
      res[i].status = st;
      if (st != MIRORR_OK)
        std::snprintf(res[i].error, sizeof(res[i].error), "%s", mirorr_last_error(c));
      else
        ok.push_back(i);
    }
    const int k = (int)ok.size();
    if (k > 0) {
      const Request& r0 = *jobs[ok[0]];
      std::vector<const mirorr_volume*> F(k), M(k);
      std::vector<double> init(r0.has_init ? 12 * k : 0), center(r0.has_center ? 3 * k : 0);
      std::vector<double> out(12 * k), oc(3 * k);
      std::vector<mirorr_reg_stats> st(k);
      for (int j = 0; j < k; ++j) {
        const Request& r = *jobs[ok[j]];
        F[j] = vf[ok[j]];
        M[j] = vm[ok[j]];
        if (r0.has_init) std::memcpy(&init[12 * j], r.init, sizeof(r.init));
        if (r0.has_center) std::memcpy(&center[3 * j], r.center, sizeof(r.center));
      }
      const mirorr_status bs = mirorr_register_batch(
          c, &r0.p, k, F.data(), M.data(), r0.has_init ? init.data() : nullptr,
          r0.has_center ? center.data() : nullptr, out.data(), oc.data(), st.data());
      for (int j = 0; j < k; ++j) {
        mirorr_request_result& q = res[ok[j]];
        q.status = bs;
        if (bs == MIRORR_OK) {
          std::memcpy(q.params, &out[12 * j], sizeof(q.params));
          std::memcpy(q.center, &oc[3 * j], sizeof(q.center));
          q.stats = st[j];
        } else {
          std::snprintf(q.error, sizeof(q.error), "%s", mirorr_last_error(c));
        }
      }
    }
    for (int i = 0; i < n; ++i) {
      mirorr_volume_destroy(vf[i]);
      mirorr_volume_destroy(vm[i]);
    }
    const double t_end = now();
    for (auto& q : res) q.t_end = t_end;
  }

  void worker(int w) {
    (void)hipSetDevice(device);
    for (;;) {
      std::vector<std::unique_ptr<Request>> jobs;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stopping || !slots[w].jobs.empty(); });
        if (slots[w].jobs.empty()) return;  // stopping with nothing to run
        jobs.swap(slots[w].jobs);
      }
      std::vector<mirorr_request_result> res(jobs.size());
      run_batch(slots[w], w, jobs, res);
      std::lock_guard<std::mutex> lk(mu);
      slots[w].busy = false;
      slots[w].reserve = 0;
      for (size_t i = 0; i < jobs.size(); ++i) {
        done[jobs[i]->ticket] = res[i];
        ++completed;
      }
      max_batch = std::max(max_batch, (int)jobs.size());
      schedule();
      cv_done.notify_all();
    }
  }
};

extern "C" {

const char* mirorr_server_last_error(void) { return g_server_error.c_str(); }

mirorr_status mirorr_server_create(int device, size_t hbm_budget, int nslots,
                                   mirorr_server** out) {
  if (!out) {
    g_server_error = "null argument";
    return MIRORR_ERR_ARG;
  }
  *out = nullptr;
  if (nslots > kMaxSlots) {
    g_server_error = "at most " + std::to_string(kMaxSlots) + " slots";
    return MIRORR_ERR_ARG;
  }
  auto S = std::make_unique<mirorr_server>();
  S->device = device;
  const int n = nslots > 0 ? nslots : 4;
  S->slots.resize(n);
  for (int i = 0; i < n; ++i) {
    mirorr_status st = mirorr_ctx_create(device, 0, &S->slots[i].ctx);
    if (st != MIRORR_OK) {
      g_server_error = "worker context: HIP device unavailable";
      for (Slot& s : S->slots) mirorr_ctx_destroy(s.ctx);
      return st;
    }
    S->slots[i].ctx->budget.parent = &S->budget;
    S->slots[i].ctx->overlap_passes = false;  // the slots overlap whole requests
  }
  if (hbm_budget == 0) {
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
      g_server_error = "hipMemGetInfo failed";
      for (Slot& s : S->slots) mirorr_ctx_destroy(s.ctx);
      return MIRORR_ERR_PLATFORM;
    }
    const size_t reserve = std::min<size_t>(size_t(8) << 30, free_b / 10);
    hbm_budget = free_b - reserve;
  }
  S->budget.limit = hbm_budget;
  mirorr_server* raw = S.release();
  try {
    for (int i = 0; i < n; ++i) raw->slots[i].th = std::thread([raw, i] { raw->worker(i); });
  } catch (const std::exception& e) {
    g_server_error = std::string("worker thread: ") + e.what();
    mirorr_server_destroy(raw);  // joins the threads already started
    return MIRORR_ERR_PLATFORM;
  }
  *out = raw;
  return MIRORR_OK;
}

mirorr_status mirorr_submit(mirorr_server* S, const mirorr_request* q, int64_t* ticket) {
  if (!S || !q || !ticket) {
    g_server_error = "null argument";
    return MIRORR_ERR_ARG;
  }
  try {
    if (!q->fixed || !q->moving) throw Error(mirorr::kErrArg, "null volume data");
    auto r = std::make_unique<Request>();
    r->p = q->params;
    r->gF = q->fixed_geom;
    r->gM = q->moving_geom;
    if (q->storage != MIRORR_STORAGE_F32 && q->storage != MIRORR_STORAGE_F16)
      throw Error(mirorr::kErrArg, "unknown volume storage");
    r->storage = q->storage;
    const size_t nF = checked_nvox(r->gF), nM = checked_nvox(r->gM);
    r->bound = mirorr::registration_bytes_bound(&r->p, mirorr::to_geom(r->gF),
                                                q->fixed_mask != nullptr, mirorr::to_geom(r->gM),
                                                q->moving_mask != nullptr, r->storage);
    if (q->borrow) {
      r->pF = q->fixed;
      r->pM = q->moving;
      r->pFm = q->fixed_mask;
      r->pMm = q->moving_mask;
    } else {
      r->F.assign(q->fixed, q->fixed + nF);
      r->M.assign(q->moving, q->moving + nM);
      if (q->fixed_mask) r->Fm.assign(q->fixed_mask, q->fixed_mask + nF);
      if (q->moving_mask) r->Mm.assign(q->moving_mask, q->moving_mask + nM);
      r->pF = r->F.data();
      r->pM = r->M.data();
      r->pFm = q->fixed_mask ? r->Fm.data() : nullptr;
      r->pMm = q->moving_mask ? r->Mm.data() : nullptr;
    }
    if (q->init_params) {
      r->has_init = true;
      std::memcpy(r->init, q->init_params, sizeof(double) * (q->params.affine ? 12 : 6));
    }
    if (q->init_center) {
      r->has_center = true;
      std::memcpy(r->center, q->init_center, sizeof(r->center));
    }
    std::lock_guard<std::mutex> lk(S->mu);
    if (S->stopping) throw Error(mirorr::kErrState, "server is shutting down");
    r->ticket = S->next_ticket++;
    r->t_submit = S->now();
    *ticket = r->ticket;
    S->live[r->ticket] = true;
    S->pending.push_back(std::move(r));
    S->schedule();
    return MIRORR_OK;
  } catch (const Error& e) {
    g_server_error = e.what();
    return static_cast<mirorr_status>(e.code);
  } catch (const std::bad_alloc&) {
    g_server_error = "host allocation failed";
    return MIRORR_ERR_NOMEM;
  }
}

mirorr_status mirorr_poll(mirorr_server* S, int64_t ticket, double timeout_s,
                          mirorr_request_result* out) {
  if (!S || !out) {
    g_server_error = "null argument";
    return MIRORR_ERR_ARG;
  }
  std::unique_lock<std::mutex> lk(S->mu);
  if (!S->live.count(ticket)) {
    g_server_error = "unknown ticket " + std::to_string(ticket);
    return MIRORR_ERR_ARG;
  }
  // Counted while inside, so that mirorr_server_destroy frees nothing a
  // poller still uses; the wait also ends when another poll of the same
  // ticket has collected it.
  ++S->pollers;
  struct Leave {
    mirorr_server* S;
    ~Leave() {
      if (--S->pollers == 0) S->cv_idle.notify_all();
    }
  } leave{S};
  auto ready = [&] { return S->done.count(ticket) != 0 || S->live.count(ticket) == 0; };
  if (!ready()) {
    if (timeout_s == 0.0) return MIRORR_PENDING;
    if (timeout_s < 0.0 || !(timeout_s < 1e6)) {  // negative, NaN or past any deadline: wait
      S->cv_done.wait(lk, ready);
    } else if (!S->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready)) {
      return MIRORR_PENDING;
    }
  }
  if (!S->done.count(ticket)) {
    g_server_error = "ticket " + std::to_string(ticket) + " was collected by another poll";
    return MIRORR_ERR_ARG;
  }
  *out = S->done[ticket];
  S->done.erase(ticket);
  S->live.erase(ticket);
  S->cv_done.notify_all();  // a concurrent poll of this ticket stops waiting
  return MIRORR_OK;
}

mirorr_status mirorr_server_info_get(mirorr_server* S, mirorr_server_info* out) {
  if (!S || !out) return MIRORR_ERR_ARG;
  std::lock_guard<std::mutex> lk(S->mu);
  std::memset(out, 0, sizeof(*out));
  out->slots = (int)S->slots.size();
  out->pending = (int)S->pending.size();
  for (const Slot& s : S->slots) out->running += s.busy ? 1 : 0;
  out->completed = S->completed;
  out->hbm_budget = S->budget.limit;
  out->hbm_used = S->budget.used.load();
  out->hbm_peak = S->budget.peak.load();
  out->max_running = S->max_running;
  out->reserve_peak = S->reserve_peak;
  out->batch_max = S->batch_max;
  out->max_batch = S->max_batch;
  return MIRORR_OK;
}

mirorr_status mirorr_server_set_batch(mirorr_server* S, int max_pairs) {
  if (!S || max_pairs < 1 || max_pairs > 256) {
    g_server_error = "batch size must be 1 to 256";
    return MIRORR_ERR_ARG;
  }
  std::lock_guard<std::mutex> lk(S->mu);
  S->batch_max = max_pairs;
  return MIRORR_OK;
}

void mirorr_server_destroy(mirorr_server* S) {
  if (!S) return;
  {
    // Requests not yet admitted complete with MIRORR_ERR_STATE, so that
    // their pollers return.
    std::lock_guard<std::mutex> lk(S->mu);
    S->stopping = true;
    while (!S->pending.empty()) {
      std::unique_ptr<Request> r = std::move(S->pending.front());
      S->pending.pop_front();
      S->finish_unadmitted(std::move(r), MIRORR_ERR_STATE, "server shut down");
    }
    S->cv_work.notify_all();
  }
  // Workers finish the request they were handed, then exit.
  for (Slot& s : S->slots)
    if (s.th.joinable()) s.th.join();
  {
    // Every live ticket now has its result; wait for the pollers inside
    // mirorr_poll to leave before freeing what they use.
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.notify_all();
    S->cv_idle.wait(lk, [&] { return S->pollers == 0; });
  }
  for (Slot& s : S->slots) mirorr_ctx_destroy(s.ctx);
  delete S;
}

}  // extern "C"
