// Ordered compaction and image pruning shared by the list builders.
#pragma once

#include "fb_device.h"

#ifndef LWG
#define LWG 256
#endif

// ordered block-wide exclusive scan of a predicate; returns this thread's
// position, writes the block total to *total (all threads)
__device__ __forceinline__ int block_rank(bool pred, int *s_wave, int *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  unsigned long long m = __ballot(pred);
  int before = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) s_wave[wave] = __popcll(m);
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < LWG / 64; w++) {
    int c = s_wave[w];
    if (w < wave) off += c;
    tot += c;
  }
  __syncthreads();
  *total = tot;
  return off + before;
}

__device__ __forceinline__ bool image_may_hit(const DStruct &S, const double *xlb, const double *r1,
                                              double reach) {
  double d2 = 0.0;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    double lo = S.bmin[k] + xlb[k], hi = S.bmax[k] + xlb[k];
    double d = fmax(0.0, fmax(lo - r1[k], r1[k] - hi));
    d2 += d * d;
  }
  return d2 <= reach * reach;
}

