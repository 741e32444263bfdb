// Binary64 three-centre interpolation shared by the 3c assemblers:
// getMEs_Fdata_3c (e.FDATA/M_Fdata_3c.f90:293-517) evaluated literally, and the
// segment ownership of a workgroup over triplets sorted by target block.
#pragma once

#include "fb_device.h"

// Legendre P0..P4 and sin(theta) (M_Fdata_3c.f90:347-354)
__device__ __forceinline__ void d_legendre(double cost, double *p, double *sint) {
  double cost2 = cost * cost;
  *sint = sqrt(fmax(1.0e-5, 1.0 - cost2));
  p[0] = 1.0;
  p[1] = cost;
  p[2] = 0.5 * (3.0 * cost2 - 1.0);
  p[3] = 0.5 * cost * (5.0 * cost2 - 3.0);
  p[4] = 0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2));
}

// getMEs_Fdata_3c in binary64 with the reference's method selection and
// evaluation order; writes the recovered molecular-frame block hm (nmu x nnu).
template <int NMAX>
__device__ inline int d_getMEs_3c(const DView &v, int gid, double x, double y, double cost, int nmu, int nnu,
                                  double *hm) {
  if (gid < 0) return ST_NO_TABLE;
  const DGroup3c g = v.grp3c[gid];
  const int nME = g.nME, nx = g.nx, ny = g.ny;
  if (nME > NMAX) return ST_CAPACITY;
  double p[5], sint;
  d_legendre(cost, p, &sint);
  const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  const double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
  const double *tab = v.tab3c + g.off;
  const int64_t tstride = (int64_t)ny * nx * nME;
#define T3D(th, i, j, m) tab[(th) * tstride + ((int64_t)((j) - 1) * nx + ((i) - 1)) * nME + (m)]
  const double gx = (T3D(0, ix + 1, iy, 0) - T3D(0, ix, iy, 0)) / dx;
  const double gy = (T3D(0, ix, iy + 1, 0) - T3D(0, ix, iy, 0)) / dy;
  const double gg = gx * gx + gy * gy;
  double F[NMAX];
  for (int m = 0; m < nME; m++) F[m] = 0.0;
  for (int th = 0; th < FB_NTHETA; th++) {
    for (int m = 0; m < nME; m++) {
      double interim;
      if (gg < 1.0e-6) {
        interim = ((1.0 - px - py) * T3D(th, ix, iy, m) + px * T3D(th, ix + 1, iy, m)) + py * T3D(th, ix, iy + 1, m);
      } else {
        double gk[4];
        for (int k = -1; k <= 2; k++) {
          double f1m1 = T3D(th, ix + k, iy - 1, m);
          double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
          double f0p3 = 3.0 * T3D(th, ix + k, iy, m), f0p6 = 2.0 * f0p3;
          double f1p3 = 3.0 * T3D(th, ix + k, iy + 1, m), f1p6 = 2.0 * f1p3;
          double f2p1 = T3D(th, ix + k, iy + 2, m);
          double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
          double bb2 = f1m3 - f0p6 + f1p3;
          double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
          double bb0 = f0p6;
          gk[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
        }
        double f1m1 = gk[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
        double f0p3 = 3.0 * gk[1], f0p6 = 2.0 * f0p3;
        double f1p3 = 3.0 * gk[2], f1p6 = 2.0 * f1p3;
        double f2p1 = gk[3];
        double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        double bb2 = f1m3 - f0p6 + f1p3;
        double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        double bb0 = f0p6;
        interim = (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
      }
      F[m] = F[m] + p[th] * interim;
    }
  }
#undef T3D
  for (int i = 0; i < nmu * nnu; i++) hm[i] = 0.0;
  for (int k = 0; k < nME; k++) {
    double val = (v.mv[g.munu + k] != 1) ? F[k] : F[k] * sint;
    hm[(v.nu[g.munu + k] - 1) * nmu + (v.mu[g.munu + k] - 1)] = val;
  }
  return 0;
}

// segment range owned by the workgroup of tile [ts, te)
__device__ __forceinline__ bool tile_segments(const DView &v, int64_t ts, int64_t te, int64_t &first, int64_t &last,
                                              int64_t &r0, int64_t &r1) {
  int64_t sgs = v.trip_seg[ts];
  first = (v.seg_off[sgs] == ts) ? sgs : sgs + 1;
  last = v.trip_seg[te - 1];
  if (first > last) return false;
  r0 = v.seg_off[first];
  r1 = v.seg_off[last + 1];
  return true;
}


// ---------------------------------------------------------------------------
// s/p species with canonical tables (make_munu_3c order): the same algorithm
// with a compile-time element count, F in table order.
// NM0: leading elements with mvalue 0; sin(theta) multiplies the others
template <int NME, int NM0 = (NME == 10) ? 6 : (NME == 3 ? 2 : 1)>
__device__ __forceinline__ int getMEs_3c_sp(const DView &v, int gid, double x, double y, double cost, double *F) {
  if (gid < 0) return ST_NO_TABLE;
  const DGroup3c g = v.grp3c[gid];
  const int nx = g.nx, ny = g.ny;
  double p[5], sint;
  d_legendre(cost, p, &sint);
  const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  const double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
  const double *tab = v.tab3c + g.off;
  const int64_t tstride = (int64_t)ny * nx * NME;
#define T3S(th, i, j, m) tab[(th) * tstride + ((int64_t)((j) - 1) * nx + ((i) - 1)) * NME + (m)]
  const double gx = (T3S(0, ix + 1, iy, 0) - T3S(0, ix, iy, 0)) / dx;
  const double gy = (T3S(0, ix, iy + 1, 0) - T3S(0, ix, iy, 0)) / dy;
  const double gg = gx * gx + gy * gy;
#pragma unroll
  for (int m = 0; m < NME; m++) F[m] = 0.0;
  for (int th = 0; th < FB_NTHETA; th++) {
#pragma unroll
    for (int m = 0; m < NME; m++) {
      double interim;
      if (gg < 1.0e-6) {
        interim = ((1.0 - px - py) * T3S(th, ix, iy, m) + px * T3S(th, ix + 1, iy, m)) + py * T3S(th, ix, iy + 1, m);
      } else {
        double gk[4];
#pragma unroll
        for (int k = -1; k <= 2; k++) {
          double f1m1 = T3S(th, ix + k, iy - 1, m);
          double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
          double f0p3 = 3.0 * T3S(th, ix + k, iy, m), f0p6 = 2.0 * f0p3;
          double f1p3 = 3.0 * T3S(th, ix + k, iy + 1, m), f1p6 = 2.0 * f1p3;
          double f2p1 = T3S(th, ix + k, iy + 2, m);
          double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
          double bb2 = f1m3 - f0p6 + f1p3;
          double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
          double bb0 = f0p6;
          gk[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
        }
        double f1m1 = gk[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
        double f0p3 = 3.0 * gk[1], f0p6 = 2.0 * f0p3;
        double f1p3 = 3.0 * gk[2], f1p6 = 2.0 * f1p3;
        double f2p1 = gk[3];
        double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        double bb2 = f1m3 - f0p6 + f1p3;
        double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        double bb0 = f0p6;
        interim = (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
      }
      F[m] = F[m] + p[th] * interim;
    }
  }
#undef T3S
  // sin(theta) on the mvalue = 1 entries (the last nME - nME(delta m = 0) of the canonical order)
#pragma unroll
  for (int m = NM0; m < NME; m++) F[m] = F[m] * sint;
  return 0;
}

