// mirorr_lts_fit_noalloc: the device LTS of mirorr_lts_fit without any
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc).
//
// Every buffer of the fit (the pairs, the kept flags, residuals, candidate
// keys, moment partials, histograms and the LTS and outer states) is carved
// from one file-scope device array, `scratch`, sized for
// kNoAllocMaxPoints pairs. Each GPU holds its own copy of the array until
// the process ends; its address on a device is looked up once per device
// with hipGetSymbolAddress, with that device current, and only that address
// is handed to kernels and copies. Every call on a GPU shares that GPU's
// array: calls that overlap on one GPU (two contexts, two threads) race and
// must be serialised by the caller. What must start at zero or one is
// written by kernels at every call (no memset nodes).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "../../include/mirorr_gpu.h"
#include "context.hpp"
#include "lts.hpp"

using namespace mirorr;

namespace {

constexpr size_t kNoAllocMaxPoints = size_t(1) << 20;
constexpr size_t kPartPoints = 1024;  // fewest points per moment partial (lts.hip)
constexpr size_t kPartDoubles = 32;   // doubles per moment partial (lts.hip)
constexpr size_t kHistBins = 4096;

constexpr size_t al(size_t n) { return (n + 255) & ~size_t(255); }
constexpr size_t kOffX = 0;
constexpr size_t kOffY = kOffX + al(3 * sizeof(double) * kNoAllocMaxPoints);
constexpr size_t kOffKept = kOffY + al(3 * sizeof(double) * kNoAllocMaxPoints);
constexpr size_t kOffR = kOffKept + al(kNoAllocMaxPoints);
constexpr size_t kOffCand = kOffR + al(sizeof(double) * kNoAllocMaxPoints);
constexpr size_t kOffPart = kOffCand + al(sizeof(unsigned long long) * kNoAllocMaxPoints);
constexpr size_t kMaxParts = (kNoAllocMaxPoints + kPartPoints - 1) / kPartPoints;
constexpr size_t kOffHist1 = kOffPart + al(sizeof(double) * kPartDoubles * kMaxParts);
constexpr size_t kOffHist2 = kOffHist1 + al(sizeof(unsigned int) * kHistBins);
constexpr size_t kOffLts = kOffHist2 + al(sizeof(unsigned int) * kHistBins);
constexpr size_t kOffOuter = kOffLts + al(sizeof(LtsState));
constexpr size_t kOffFlags = kOffOuter + al(sizeof(OuterState));
constexpr size_t kScratchBytes = kOffFlags + al(3 * sizeof(int));  // lts_finish writes 3

// The file-scope scratch array (about 66 MB). In host code its name is a
// host-side stand-in, never a device address: only scratch_base() below
// yields the address kernels and copies may use.
__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
unsigned char* g_base[kMaxDevices] = {};

// The device address of `scratch` on the current device (`device`), looked
// up on the first call for that device.
unsigned char* scratch_base(int device) {
  if (device < 0 || device >= kMaxDevices) throw Error(kErrArg, "device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_base[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr)
      throw Error(kErrPlatform,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_base[device] = static_cast<unsigned char*>(p);
  }
  return g_base[device];
}

__global__ void fill_u8_kernel(uint8_t* p, int64_t n, uint8_t v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void zero_f64_kernel(double* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0.0;
}

__global__ void set_count_kernel(LtsState* st, unsigned long long n) { st->n_kept = n; }

}  // namespace

extern "C" mirorr_status mirorr_lts_fit_noalloc(mirorr_ctx* c, int symmetric, int affine,
                                                const double* x, const double* y, size_t n,
                                                double* params, double* center, double* fwd34,
                                                double* bwd34, int* iterations) {
  if (!c || !params || (n > 0 && (!x || !y))) return MIRORR_ERR_ARG;
  std::lock_guard<std::mutex> lk(c->mu);
  return guarded(c, [&] {
    if (n > kNoAllocMaxPoints)
      throw Error(kErrArg, "mirorr_lts_fit_noalloc: more than 1048576 pairs");
    unsigned char* base = scratch_base(c->device);
    hipStream_t s = c->stream;
    double* X = reinterpret_cast<double*>(base + kOffX);
    double* Y = reinterpret_cast<double*>(base + kOffY);
    uint8_t* kept = base + kOffKept;
    LtsBuffers lb;
    lb.r = reinterpret_cast<double*>(base + kOffR);
    lb.cand = reinterpret_cast<unsigned long long*>(base + kOffCand);
    lb.part = reinterpret_cast<double*>(base + kOffPart);
    lb.hist1 = reinterpret_cast<unsigned int*>(base + kOffHist1);
    lb.hist2 = reinterpret_cast<unsigned int*>(base + kOffHist2);
    LtsState* st = reinterpret_cast<LtsState*>(base + kOffLts);
    OuterState* outer = reinterpret_cast<OuterState*>(base + kOffOuter);
    int* dflags = reinterpret_cast<int*>(base + kOffFlags);

    const int kind = affine ? xf::kAffine : xf::kRigid;
    const int64_t np = (int64_t)n;
    const int64_t npts = std::max<int64_t>(np, 1);  // as mirorr_lts_fit
    lb.nparts_cap = lts_parts(npts);
    if (lb.nparts_cap > (int64_t)kMaxParts) throw Error(kErrState, "moment partials exceed the scratch");
    if (np > 0) {
      MIRORR_HIP(hipMemcpyAsync(X, x, sizeof(double) * 3 * np, hipMemcpyHostToDevice, s));
      MIRORR_HIP(hipMemcpyAsync(Y, y, sizeof(double) * 3 * np, hipMemcpyHostToDevice, s));
    } else {
      hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, s, X, 3);
      hipLaunchKernelGGL(zero_f64_kernel, dim3(1), dim3(64), 0, s, Y, 3);
    }
    const int64_t fill_grid = std::min<int64_t>((npts + 255) / 256, 1024);
    hipLaunchKernelGGL(fill_u8_kernel, dim3((unsigned)fill_grid), dim3(256), 0, s, kept, npts,
                       (uint8_t)1);
    MIRORR_CHECK_LAUNCH();
    // The outer state of a fit on its own, as mirorr_lts_fit sets it.
    OuterState o{};
    const Geom unit{{1, 1, 1}, {1, 1, 1}, {0, 0, 0}, {1, 0, 0, 0, 1, 0, 0, 0, 1}};
    o.gF = o.gM = o.gFfull = o.gMfull = o.gHF = o.gHM = unit;
    o.fixT = o.movT = xf::identity(xf::kRigid);
    o.t12F = o.t12B = xf::identity(kind);
    o.T = xf::identity(kind);
    if (!symmetric) {
      if (center) xf::set_center(o.T, center);
      xf::set_params(o.T, params);
    }
    xf::get_params(o.T, o.lastp);
    MIRORR_HIP(hipMemcpyAsync(outer, &o, sizeof(o), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(set_count_kernel, dim3(1), dim3(1), 0, s, st, (unsigned long long)np);
    MIRORR_CHECK_LAUNCH();
    mirorr_reg_params rp;
    mirorr_reg_default_params(&rp);
    lts_start(st, outer, kind, symmetric ? 1 : 0, rp.lts_tolerance, lb, s);
    // Chunks of 4 predicated inner iterations, each followed by the finish
    // and one read-back of its flags into host stack memory.
    int hflags[3] = {1, 0, 0};
    const int chunk = 4;
    for (int it = 0; it < 101; it += chunk) {
      const int it1 = std::min(it + chunk, 101);
      for (int j = it; j < it1; ++j) lts_enqueue_iteration(X, Y, kept, npts, st, outer, lb, j, s);
      lts_finish(st, outer, dflags, s);
      MIRORR_HIP(hipMemcpyAsync(hflags, dflags, sizeof(hflags), hipMemcpyDeviceToHost, s));
      MIRORR_HIP(hipStreamSynchronize(s));
      if (!hflags[0]) break;
    }
    MIRORR_HIP(hipMemcpyAsync(&o, outer, sizeof(o), hipMemcpyDeviceToHost, s));
    MIRORR_HIP(hipStreamSynchronize(s));
    xf::get_params(o.T, params);
    if (center)
      for (int k = 0; k < 3; ++k) center[k] = o.T.c[k];
    for (int r = 0; r < 3; ++r) {
      for (int k = 0; k < 3; ++k) {
        if (fwd34) fwd34[4 * r + k] = o.t12F.M.a[r][k];
        if (bwd34) bwd34[4 * r + k] = o.t12B.M.a[r][k];
      }
      if (fwd34) fwd34[4 * r + 3] = o.t12F.off[r];
      if (bwd34) bwd34[4 * r + 3] = o.t12B.off[r];
    }
    if (iterations) *iterations = o.inner_total;
  });
}
