// Device-side geometry, rotation and interpolation routines for gfx950.
//
// The binary64 routines keep the reference's evaluation order (the library is
// compiled with -ffp-contract=off) so that list decisions are bit-identical
// and assembled blocks agree to rounding:
//   distance            b.FUNCTIONS/distance.f90:72-73
//   epsilon_function    b.FUNCTIONS/epsilon.f90:73-177
//   twister / rotate    c.SYSTEM/M_rotations.f90:98-313
//   getMEs_Fdata_2c     e.FDATA/M_Fdata_2c.f90:302-415
//   getMEs_Fdata_3c     e.FDATA/M_Fdata_3c.f90:293-517
//   smoother            b.FUNCTIONS/smoother.f90:102-145
#pragma once

#include "fb_internal.h"


// Tile of workgroup b in a grid of n.  The dispatcher hands workgroups to the
// eight XCDs round-robin (workgroup b to XCD b mod 8) and each XCD has its own
// L2; renumbering so that XCD x runs one contiguous range of tiles keeps the
// gathers of neighbouring tiles (the same atoms' blocks and positions) in one
// L2.  A bijection of [0, n): any tile order is valid for the kernels that use it.
__device__ __forceinline__ int64_t xcd_tile(unsigned b, unsigned n) {
  const unsigned x = b & 7u, q = b >> 3, per = n >> 3, rem = n & 7u;
  return x < rem ? (int64_t)x * (per + 1) + q : (int64_t)rem * (per + 1) + (int64_t)(x - rem) * per + q;
}

__device__ __forceinline__ double d_distance(const double *a, const double *b) {
  double d1 = b[0] - a[0], d2 = b[1] - a[1], d3 = b[2] - a[2];
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

__device__ __forceinline__ double d_mag(const double *a) {
  return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

__device__ __forceinline__ void d_cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

__device__ __forceinline__ double d_smoother(double r, double rend, double x) {
  double rbegin = x * rend;
  if (r > rend) return 0.0;
  if (r < rbegin) return 1.0;
  double ratio = (r - rbegin) / (rend - rbegin);
  double sm = 1.0 - ratio * ratio;
  return sm * sm;
}

// frame: spe column-major 3x3 (x', y', z'); returns 1 when a fallback branch was taken
__device__ inline int d_epsilon(const double *r1, const double *r2, double *spe) {
  double r1_mag = d_mag(r1), r2_mag = d_mag(r2);
  double zp[3], yp[3] = {0.0, 0.0, 0.0}, xp[3], r1h[3];
  int fb = 0;
  if (r2_mag < 1.0e-4) {
#pragma unroll
    for (int i = 0; i < 9; i++) spe[i] = 0.0;
    spe[0] = spe[4] = spe[8] = 1.0;
    return 1;
  }
#pragma unroll
  for (int i = 0; i < 3; i++) zp[i] = r2[i] / r2_mag;
  double yp_mag = 0.0;
  if (r1_mag > 1.0e-4) {
#pragma unroll
    for (int i = 0; i < 3; i++) r1h[i] = r1[i] / r1_mag;
    d_cross(zp, r1h, yp);
    yp_mag = d_mag(yp);
  }
  if (yp_mag < 1.0e-4) {
    fb = 1;
    if (fabs(zp[0]) > 1.0e-4) {
      yp[0] = -(zp[1] + zp[2]) / zp[0]; yp[1] = 1.0; yp[2] = 1.0;
    } else if (fabs(zp[1]) > 1.0e-4) {
      yp[0] = 1.0; yp[1] = -(zp[0] + zp[2]) / zp[1]; yp[2] = 1.0;
    } else {
      yp[0] = 1.0; yp[1] = 1.0; yp[2] = -(zp[0] + zp[1]) / zp[2];
    }
    yp_mag = d_mag(yp);
  }
#pragma unroll
  for (int i = 0; i < 3; i++) yp[i] = yp[i] / yp_mag;
  d_cross(yp, zp, xp);
#pragma unroll
  for (int i = 0; i < 3; i++) { spe[i] = xp[i]; spe[3 + i] = yp[i]; spe[6 + i] = zp[i]; }
  return fb;
}

// p-orbital rotation in (y, z, x) order from eps, c.SYSTEM/M_rotations.f90:260-270
template <typename T, typename E>
__device__ __forceinline__ void d_pmat(const E *eps, T *pm) {
#define EPS(i, j) ((T)eps[3 * ((j) - 1) + (i) - 1])
  pm[0] = EPS(2, 2); pm[3] = EPS(2, 3); pm[6] = EPS(2, 1);
  pm[1] = EPS(3, 2); pm[4] = EPS(3, 3); pm[7] = EPS(3, 1);
  pm[2] = EPS(1, 2); pm[5] = EPS(1, 3); pm[8] = EPS(1, 1);
#undef EPS
}

// The d-orbital rotation of twister (c.SYSTEM/M_rotations.f90:272-300):
// column-major dm[5 (m - 1) + (k - 1)] = dmat(k, m) from eps (column-major
// eps(i, j) = eps[3 (j - 1) + (i - 1)]), with the A matrices of
// include/constants.h:26-35 and the reference's sums in its order (terms with
// a zero A entry skipped, as the reference skips them).
__device__ inline void d_dmat(const double *eps, double *dm) {
  // amat(ix, jx, imu), column-major, imu = xy, yz, z^2, xz, x^2-y^2
  const double r3 = 1.0 / (2.0 * sqrt(3.0));
  const double amat[5][3][3] = {
      {{0.0, 0.5, 0.0}, {0.5, 0.0, 0.0}, {0.0, 0.0, 0.0}},
      {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.5}, {0.0, 0.5, 0.0}},
      {{-r3, 0.0, 0.0}, {0.0, -r3, 0.0}, {0.0, 0.0, 2.0 / (2.0 * sqrt(3.0))}},
      {{0.0, 0.0, 0.5}, {0.0, 0.0, 0.0}, {0.5, 0.0, 0.0}},
      {{0.5, 0.0, 0.0}, {0.0, -0.5, 0.0}, {0.0, 0.0, 0.0}}};
#define EPSD(i, j) eps[3 * ((j) - 1) + (i) - 1]
  for (int imu = 0; imu < 5; imu++) {
    double l11 = 0.0, l12 = 0.0, l13 = 0.0, l32 = 0.0, l33 = 0.0;
    for (int jx = 1; jx <= 3; jx++)
      for (int ix = 1; ix <= 3; ix++) {
        const double a = amat[imu][jx - 1][ix - 1];
        if (a != 0.0) {
          l11 = l11 + a * EPSD(jx, 1) * EPSD(ix, 1);
          l12 = l12 + a * EPSD(jx, 1) * EPSD(ix, 2);
          l13 = l13 + a * EPSD(jx, 1) * EPSD(ix, 3);
          l32 = l32 + a * EPSD(jx, 3) * EPSD(ix, 2);
          l33 = l33 + a * EPSD(jx, 3) * EPSD(ix, 3);
        }
      }
    dm[0 * 5 + imu] = 2.0 * l12;
    dm[1 * 5 + imu] = 2.0 * l32;
    dm[2 * 5 + imu] = sqrt(3.0) * l33;
    dm[3 * 5 + imu] = 2.0 * l13;
    dm[4 * 5 + imu] = 2.0 * l11 + l33;
  }
#undef EPSD
}

// X = L M R^T shell pair by shell pair, loop order of
// c.SYSTEM/M_rotations.f90:141-164: L and R are 1 (s), pm (p, column-major
// pm[3 m + k]) or dm (d, d_dmat; only read when a species has a d shell).  M
// and X are column-major with leading dimension nmu.
template <typename T>
__device__ inline void d_rotate(const DSpecies &sp, int in1, int in3, const T *pm, const T *M, int nmu,
                                T *X, const T *dm = nullptr) {
  int nnu = sp.norb[in3 - 1];
  for (int i = 0; i < nmu * nnu; i++) X[i] = (T)0;
  int n1 = 0;
  for (int issh = 0; issh < sp.nssh[in1 - 1]; issh++) {
    int l1 = sp.lssh[in1 - 1][issh];
    int n2 = 0;
    for (int jssh = 0; jssh < sp.nssh[in3 - 1]; jssh++) {
      int l2 = sp.lssh[in3 - 1][jssh];
      for (int m2 = 0; m2 < 2 * l2 + 1; m2++)
        for (int m1 = 0; m1 < 2 * l1 + 1; m1++)
          for (int k2 = 0; k2 < 2 * l2 + 1; k2++)
            for (int k1 = 0; k1 < 2 * l1 + 1; k1++) {
              T left = (l1 == 0) ? (T)1 : (l1 == 1 ? pm[3 * m1 + k1] : dm[5 * m1 + k1]);
              T right = (l2 == 0) ? (T)1 : (l2 == 1 ? pm[3 * m2 + k2] : dm[5 * m2 + k2]);
              T *x = &X[(n2 + k2) * nmu + (n1 + k1)];
              *x = *x + left * M[(n2 + m2) * nmu + (n1 + m1)] * right;
            }
      n2 += 2 * l2 + 1;
    }
    n1 += 2 * l1 + 1;
  }
}

// 7-point Lagrange interpolation of one 2c cell, e.FDATA/M_Fdata_2c.f90:366-393.
// Returns the status bits of the range checks; F receives nME values.
__device__ inline int d_interp2c(const DCell2c &c, const double *__restrict__ tab, double x, double *F) {
  if (x < 0.0) return ST_X_NEG;
  if (x > c.xmax * (1.0 + 1.0e-5)) return ST_X_RANGE;
  int imid = (int)(x / c.dx) + 1;
  int ileft = imid - 3, iright = imid + 3;
  if (ileft < 1) { ileft = 1; iright = 7; }
  else if (iright > c.nx) { ileft = c.nx - 6; iright = c.nx; }
  double xx[7];
#pragma unroll
  for (int k = 0; k < 7; k++) xx[k] = (double)(ileft + k - 1) * c.dx;
  for (int m = 0; m < c.nME; m++) F[m] = 0.0;
  const double *base = tab + c.off;
#pragma unroll
  for (int is = 0; is < 7; is++) {
    double prod = 1.0;
#pragma unroll
    for (int ip = 0; ip < 7; ip++) {
      if (ip != is) {
        double pden = 1.0 / (xx[is] - xx[ip]);
        prod = prod * (x - xx[ip]) * pden;
      }
    }
    const double *row = base + (int64_t)(ileft - 1 + is) * c.nME;
    for (int m = 0; m < c.nME; m++) F[m] = F[m] + row[m] * prod;
  }
  return 0;
}

// Header of 3c group gid read through the scalar cache: a waterfall over
// the distinct group indices of the active lanes (one pass when a
// wavefront's triplets share their group, as in a one-element system).  The
// empty asm keeps every field's load at the scalar index inside the loop.
typedef const __attribute__((address_space(4))) DGroup3c CGrp3c;
__device__ __forceinline__ DGroup3c group3c_scalar(const DView &v, int gid) {
#ifndef __HIP_DEVICE_COMPILE__
  return v.grp3c[gid];   // the host pass only parses device code
#else
  DGroup3c g;
  for (;;) {
    const int gc = __builtin_amdgcn_readfirstlane(gid);
    const DGroup3c t = ((CGrp3c *)v.grp3c)[gc];
    asm volatile("" ::"s"(t.nME), "s"(t.nx), "s"(t.ny), "s"(t.munu), "s"(t.cls), "s"(t.nq), "s"(t.xmax),
                 "s"(t.ymax), "s"(t.dx), "s"(t.dy), "s"(t.off), "s"(t.moff));
    if (gid == gc) {
      g = t;
      break;
    }
  }
  return g;
#endif
}
