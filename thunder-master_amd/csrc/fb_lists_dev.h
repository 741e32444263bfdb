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


// ordered block-wide exclusive scan of per-thread counts; returns this
// thread's offset, writes the block total to *total (all threads)
__device__ __forceinline__ int block_scan_count(int c, int *s_wave, int *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = c;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(x, d);
    if (lane >= d) x += y;
  }
  if (lane == 63) s_wave[wave] = x;
  __syncthreads();
  int off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < LWG / 64; w++) {
    const int t = s_wave[w];
    if (w < wave) off += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return off + x - c;
}

// cell of a home position along each axis, clamped to the grid
__device__ __forceinline__ int cell_of(const DStruct &S, const double *r) {
  int c[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double f = floor((r[a] - S.clo[a]) / S.cw[a]);
    c[a] = (int)fmin(fmax(f, 0.0), (double)(S.cn[a] - 1));
  }
  return (c[2] * S.cn[1] + c[1]) * S.cn[0] + c[0];
}

// The cells of structure S that overlap the box q +- S.crange: their first
// entries in s_cs, the prefix of their sizes in s_pre (s_pre[n] = candidates).
// Every atom within S.crange of q is in one of them (the cell of a position is
// monotone in the position, and the clamp keeps that).  All threads call it;
// returns the number of cells (0: none overlaps).
__device__ __forceinline__ int query_cells(const DView &v, const DStruct &S, const double *q, int64_t *s_cs,
                                           int *s_pre) {
  int lo[3], n[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const double fl = floor((q[a] - S.crange - S.clo[a]) / S.cw[a]);
    const double fh = floor((q[a] + S.crange - S.clo[a]) / S.cw[a]);
    if (fh < 0.0 || fl > (double)(S.cn[a] - 1)) return 0;
    lo[a] = (int)fmax(fl, 0.0);
    n[a] = min((int)fmin(fh, (double)(S.cn[a] - 1)) - lo[a] + 1, 3);   // cw > crange: three at most
  }
  const int nc = n[0] * n[1] * n[2];   // at most 27
  const int t = threadIdx.x;
  if (t < nc) {
    const int cx = lo[0] + t % n[0], cy = lo[1] + (t / n[0]) % n[1], cz = lo[2] + t / (n[0] * n[1]);
    const int64_t c = S.cell_off + (int64_t)(cz * S.cn[1] + cy) * S.cn[0] + cx;
    const int64_t a = v.cell_start[c];
    s_cs[t] = a;
    s_pre[t + 1] = (int)(v.cell_start[c + 1] - a);
  }
  __syncthreads();
  if (t == 0) {
    s_pre[0] = 0;
    for (int k = 1; k <= nc; k++) s_pre[k] += s_pre[k - 1];
  }
  __syncthreads();
  return nc;
}

// local atom of candidate t of the cells found by query_cells
__device__ __forceinline__ int cell_candidate(const DView &v, const int64_t *s_cs, const int *s_pre, int nc, int t) {
  int k = 0;
  while (k + 1 < nc && s_pre[k + 1] <= t) k++;
  return v.cell_atoms[s_cs[k] + (t - s_pre[k])];
}

// The images of structure S whose shifted atom bounding box may come within
// reach of r1 (image_may_hit), in increasing mbeta, into s_img
// [FB_NIMG]; the workgroup tests LWG images at a time instead of walking all
// of them.  All threads call it; returns the count.
__device__ __forceinline__ int hit_images(const DStruct &S, const double *xls, const double *r1, double reach,
                                          int *s_img, int *s_wave) {
  int n = 0;
  for (int m0 = 0; m0 <= S.mbeta_max; m0 += LWG) {
    const int mb = m0 + (int)threadIdx.x;
    const bool hit = mb <= S.mbeta_max && image_may_hit(S, xls + 3 * mb, r1, reach);
    int tot;
    const int pos = block_rank(hit, s_wave, &tot);
    if (hit) s_img[n + pos] = mb;
    n += tot;
  }
  __syncthreads();
  return n;
}
