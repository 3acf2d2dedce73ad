// Shared host/device definitions of the MI355X registration library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace mirorr {

// Failure raised inside the library; converted to a mirorr_status at the
// C-ABI boundary (never crosses it).
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

enum : int { kErrArg = 1, kErrPlatform = 2, kErrAlgo = 3, kErrNoMem = 4, kErrState = 5 };

#define MIRORR_HIP(call)                                                        \
  do {                                                                          \
    hipError_t e_ = (call);                                                     \
    if (e_ != hipSuccess)                                                       \
      throw ::mirorr::Error(e_ == hipErrorOutOfMemory ? ::mirorr::kErrNoMem     \
                                                      : ::mirorr::kErrPlatform, \
                            std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define MIRORR_CHECK_LAUNCH() MIRORR_HIP(hipGetLastError())

// Element type of the image buffers a search or registration reads: float,
// or binary16 holding an affine encoding of the intensities, (v - offset) *
// scale with a power-of-two scale (integers of up to 12 bits are exact).
// NCC^2 does not change under an affine map of either image, so kernels
// work on the encoded values directly.
enum Storage : int { kStoreF32 = 0, kStoreF16 = 1 };
typedef _Float16 half_t;
inline size_t storage_bytes(int storage) { return storage == kStoreF16 ? 2 : 4; }

// Device image geometry (index -> physical: origin + D * diag(spacing) * idx).
struct Geom {
  int dims[3];
  double spacing[3];
  double origin[3];
  double dir[9];
};

// Block grid of a matcher (itkBlockMatcher.txx:531-600) plus parameters.
struct BMGrid {
  int dims[3];
  int bw, bg, nw, ng;
  int nb[3];     // blocks per axis
  int start[3];  // first block corner per axis (centred grid)
  int64_t nblocks;
};

inline BMGrid make_grid(const int dims[3], int bw, int bg, int nw, int ng) {
  BMGrid g{};
  for (int a = 0; a < 3; ++a) g.dims[a] = dims[a];
  g.bw = bw;
  g.bg = bg;
  g.nw = nw;
  g.ng = ng;
  g.nblocks = 1;
  for (int a = 0; a < 3; ++a) {
    const int n = dims[a] >= bw ? (dims[a] - bw) / ng + 1 : 0;
    g.nb[a] = n;
    g.start[a] = n > 0 ? (dims[a] - (ng * (n - 1) + bw)) / 2 : 0;
    g.nblocks *= n;
  }
  return g;
}

// Candidate range on one axis for a block whose corner is p: first start and
// count, following itkBlockMatcher.txx:744-756 and the sparse iterator's end
// rule (a zero-size span with gap >= 2 still visits its start once).
__host__ __device__ inline void cand_range(int p, int L, int bw, int bg, int nw,
                                           int* first, int* count) {
  long start = (long)p - (long)(nw / 2) * bg;
  long end = start + (long)nw * bg;
  const long search_end = (long)L - bw + 1;
  if (start < 0) start += ((-start + bg - 1) / bg) * bg;
  if (end > search_end) end -= ((end - search_end + bg - 1) / bg) * bg;
  const long size = end - start;
  *first = (int)start;
  *count = size <= 0 ? 1 : (int)((size - 1) / bg) + 1;
}

}  // namespace mirorr
