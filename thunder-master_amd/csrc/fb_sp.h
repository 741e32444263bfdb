// Register-resident Slater-Koster rotation for species whose shells are (s) or
// (s, p) with the tables' canonical matrix-element order (make_munu /
// make_munu_3c, c.SYSTEM/M_integrals_2c.f90:297-335, M_integrals_3c.f90:299-477).
//
// rotate (c.SYSTEM/M_rotations.f90:141-164) forms X(k1,k2) as a running sum
// over (m2 outer, m1 inner) of L(k1,m1) M(m1,m2) R(k2,m2), starting from 0.
// The formulas below add the non-zero terms of M in that same order with the
// same (L*M)*R association, so in binary64 they reproduce the reference's
// rounding; the terms skipped are exact zeros.
//
// Orbital order: s, then p as (y, z, x).  P(k, m) = pm[3*m + k] rotates
// molecular p component m into crystal component k (twister,
// c.SYSTEM/M_rotations.f90:260-270).  Blocks are column-major, leading
// dimension nmu.
#pragma once

// pair class: (norb_i == 4) * 2 + (norb_j == 4); 0 s-s, 1 s-sp, 2 sp-s, 3 sp-sp
#define CLS_SS 0
#define CLS_SSP 1
#define CLS_SPS 2
#define CLS_SPSP 3

__host__ __device__ constexpr int nme2c_of(int cls) { return cls == 3 ? 6 : (cls == 0 ? 1 : 2); }
__host__ __device__ constexpr int nme3c_of(int cls) { return cls == 3 ? 10 : (cls == 0 ? 1 : 3); }

#define PP(k, m) pm[3 * (m) + (k)]

// 2c: F = {ss, s-pz, pz-s, py-py, pz-pz, px-px} (sp-sp), {ss, s-pz} (s-sp),
// {ss, pz-s} (sp-s), {ss} (s-s)
template <typename T>
__device__ __forceinline__ void rot2c_sp(int cls, const T *pm, const T *F, T *X) {
  if (cls == CLS_SPSP) {
    const T a = F[0], b = F[1], c = F[2], d = F[3], e = F[4], f = F[5];
    X[0] = a;
#pragma unroll
    for (int l = 0; l < 3; l++) X[(1 + l) * 4] = b * PP(l, 1);
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = PP(k, 1) * c;
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        T x = (PP(k, 0) * d) * PP(l, 0);
        x = x + (PP(k, 1) * e) * PP(l, 1);
        x = x + (PP(k, 2) * f) * PP(l, 2);
        X[(1 + l) * 4 + 1 + k] = x;
      }
  } else if (cls == CLS_SSP) {
    X[0] = F[0];
#pragma unroll
    for (int l = 0; l < 3; l++) X[1 + l] = F[1] * PP(l, 1);
  } else if (cls == CLS_SPS) {
    X[0] = F[0];
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = PP(k, 1) * F[1];
  } else {
    X[0] = F[0];
  }
}

// 3c: F = {ss, s-pz, pz-s, py-py, pz-pz, px-px, s-px, px-s, px-pz, pz-px} (sp-sp),
// {ss, s-pz, s-px} (s-sp), {ss, pz-s, px-s} (sp-s), {ss} (s-s); the mvalue = 1
// entries already carry sin(theta).
template <typename T>
__device__ __forceinline__ void rot3c_sp(int cls, const T *pm, const T *F, T *X) {
  if (cls == CLS_SPSP) {
    const T a = F[0], b = F[1], c = F[2], d = F[3], e = F[4], f = F[5], g = F[6], h = F[7], i = F[8], j = F[9];
    X[0] = a;
#pragma unroll
    for (int l = 0; l < 3; l++) X[(1 + l) * 4] = b * PP(l, 1) + g * PP(l, 2);
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = PP(k, 1) * c + PP(k, 2) * h;
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        T x = (PP(k, 0) * d) * PP(l, 0);
        x = x + (PP(k, 1) * e) * PP(l, 1);
        x = x + (PP(k, 2) * i) * PP(l, 1);
        x = x + (PP(k, 1) * j) * PP(l, 2);
        x = x + (PP(k, 2) * f) * PP(l, 2);
        X[(1 + l) * 4 + 1 + k] = x;
      }
  } else if (cls == CLS_SSP) {
    X[0] = F[0];
#pragma unroll
    for (int l = 0; l < 3; l++) X[1 + l] = F[1] * PP(l, 1) + F[2] * PP(l, 2);
  } else if (cls == CLS_SPS) {
    X[0] = F[0];
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = PP(k, 1) * F[1] + PP(k, 2) * F[2];
  } else {
    X[0] = F[0];
  }
}

// element (row, col) of the rotated 3c value block (rot3c_sp)
template <int CLS>
__device__ __forceinline__ double rot3c_el(int row, int col, const double *pm, const double *F) {
  if (row == 0 && col == 0) return F[0];
  if (CLS == CLS_SPSP) {
    if (row == 0) return F[1] * PP(col - 1, 1) + F[6] * PP(col - 1, 2);
    if (col == 0) return PP(row - 1, 1) * F[2] + PP(row - 1, 2) * F[7];
    const int k = row - 1, l = col - 1;
    double x = (PP(k, 0) * F[3]) * PP(l, 0);
    x = x + (PP(k, 1) * F[4]) * PP(l, 1);
    x = x + (PP(k, 2) * F[8]) * PP(l, 1);
    x = x + (PP(k, 1) * F[9]) * PP(l, 2);
    return x + (PP(k, 2) * F[5]) * PP(l, 2);
  }
  if (CLS == CLS_SSP) return F[1] * PP(col - 1, 1) + F[2] * PP(col - 1, 2);
  if (CLS == CLS_SPS) return PP(row - 1, 1) * F[1] + PP(row - 1, 2) * F[2];
  return 0.0;
}
#undef PP
