// Device helpers for radix selection (workgroup scan and bin search), shared
// by the LTS and the block-variance selection. Workgroups have kSelT threads.
#pragma once

#include <hip/hip_runtime.h>

namespace mirorr {
namespace sel {

constexpr int kSelT = 256;

// Inclusive scan over the kSelT threads of a workgroup.
__device__ inline unsigned long long block_scan(unsigned long long v) {
  __shared__ unsigned long long wsum[kSelT / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const unsigned long long u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  if (lane == 63) wsum[w] = v;
  __syncthreads();
  for (int i = 0; i < w; ++i) v += wsum[i];
  __syncthreads();
  return v;
}

// Finds the bin holding rank k (0-based, ascending) of a histogram of
// kSelT * PER bins, and the rank left within that bin. Called by the whole
// workgroup; the thread owning the bin writes *bin and *krem.
template <int PER>
__device__ inline void find_bin(const unsigned int* hist, unsigned long long k, unsigned int* bin,
                                unsigned long long* krem) {
  const int t = threadIdx.x;
  unsigned int h[PER];
  unsigned long long s = 0;
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    h[j] = hist[PER * t + j];
    s += h[j];
  }
  const unsigned long long inc = block_scan(s);
  const unsigned long long before = inc - s;
  if (k >= before && k < inc) {
    unsigned long long c = before;
    for (int j = 0; j < PER; ++j) {
      if (k < c + h[j]) {
        *bin = PER * t + j;
        *krem = k - c;
        break;
      }
      c += h[j];
    }
  }
}

}  // namespace sel
}  // namespace mirorr
