// Seven-point Lagrange interpolation helpers of the 2c assemblers
// (e.FDATA/M_Fdata_2c.f90:366-393): window, products with the host-evaluated
// denominators, and the interpolation of one cell with a compile-time element
// count.
#pragma once

#include "fb_device.h"

// the seven Lagrange products of M_Fdata_2c.f90:380-392 for window ileft
__device__ __forceinline__ void lagrange7(const DView &v, const DGrid2c &gr, double x, int ileft, double *prod) {
  const double *pd = v.pden + gr.pden_off + (int64_t)(ileft - 1) * 49;
  double xx[7];
#pragma unroll
  for (int k = 0; k < 7; k++) xx[k] = (double)(ileft + k - 1) * gr.dx;
#pragma unroll
  for (int is = 0; is < 7; is++) {
    double p = 1.0;
#pragma unroll
    for (int ip = 0; ip < 7; ip++)
      if (ip != is) p = p * (x - xx[ip]) * pd[is * 7 + ip];
    prod[is] = p;
  }
}

// The same products with the reciprocals formed in-kernel with the
// reference's expression, pden = 1/(xx_is - xx_ip) (correctly rounded
// binary64 division, as on the host).  fl(a - b) = -fl(b - a) and
// 1/(-y) = -(1/y) exactly, so the 21 divisions of is < ip give the other 21
// by negation.  A lane computes them instead of gathering the 49 host values
// of its window: the gathers of a wavefront's different windows cost more
// than the divisions.
__device__ __forceinline__ void lagrange7_div(double x, double dx, int ileft, double *prod) {
  double xx[7], r[7][7];
#pragma unroll
  for (int k = 0; k < 7; k++) xx[k] = (double)(ileft + k - 1) * dx;
#pragma unroll
  for (int is = 0; is < 7; is++)
#pragma unroll
    for (int ip = is + 1; ip < 7; ip++) {
      r[is][ip] = 1.0 / (xx[is] - xx[ip]);
      r[ip][is] = -r[is][ip];
    }
#pragma unroll
  for (int is = 0; is < 7; is++) {
    double p = 1.0;
#pragma unroll
    for (int ip = 0; ip < 7; ip++)
      if (ip != is) p = p * (x - xx[ip]) * r[is][ip];
    prod[is] = p;
  }
}

typedef const __attribute__((address_space(4))) int CInt;
typedef const __attribute__((address_space(4))) DCell2c CCell2c;
typedef const __attribute__((address_space(4))) DGrid2c CGrid2c;

// Cell of interaction (slot, isub) for the species pair sp2 = in1 | in2 << 16,
// read through the scalar cache: a waterfall over the distinct (sp2, isub)
// of the active lanes (one pass when the wavefront's pairs share their
// species, as in a one-element system).  The empty asm keeps the loads at the
// scalar index inside the loop.  Returns the cell index (negative: no table).
__device__ __forceinline__ int cell_scalar(const DView &v, int sp2, int slot, int isub, DCell2c &c) {
#ifndef __HIP_DEVICE_COMPILE__
  // the host pass only parses device code
  const int cid = v.idx2c[(((sp2 & 0xffff) - 1) * v.sp.ns + ((sp2 >> 16) - 1)) * FB_NSLOT2 * FB_MAXSUB2 +
                         slot * FB_MAXSUB2 + isub];
  if (cid >= 0) c = v.cell2c[cid];
  return cid;
#else
  int cid = -1;
  for (;;) {
    const int k = __builtin_amdgcn_readfirstlane(sp2), ks = __builtin_amdgcn_readfirstlane(isub);
    const int a = k & 0xffff, b = k >> 16;
    const int id = ((CInt *)v.idx2c)[(((a - 1) * v.sp.ns + (b - 1)) * FB_NSLOT2 + slot) * FB_MAXSUB2 + ks];
    DCell2c t{};
    if (id >= 0) t = ((CCell2c *)v.cell2c)[id];
    asm volatile("" ::"s"(id), "s"(t.xmax), "s"(t.off), "s"(t.grid));
    if (sp2 == k && isub == ks) {
      cid = id;
      c = t;
      break;
    }
  }
  return cid;
#endif
}

__device__ __forceinline__ DGrid2c grid_scalar(const DView &v, int g) {
#ifndef __HIP_DEVICE_COMPILE__
  return v.grid2c[g];
#else
  return ((CGrid2c *)v.grid2c)[g];   // g comes from a scalar cell header
#endif
}

// window of M_Fdata_2c.f90:366-375
__device__ __forceinline__ int window7(double x, double dx, int nx) {
  int imid = (int)(x / dx) + 1;
  int ileft = imid - 3;
  if (ileft < 1) ileft = 1;
  else if (imid + 3 > nx) ileft = nx - 6;
  return ileft;
}

// interpolate one canonical cell with precomputed products; F has nme entries
template <int NME>
__device__ __forceinline__ void interp_cell(const DView &v, const DCell2c &c, int ileft, const double *prod, double *F) {
  const double *base = v.tab2c + c.off + (int64_t)(ileft - 1) * NME;
#pragma unroll
  for (int m = 0; m < NME; m++) F[m] = 0.0;
#pragma unroll
  for (int is = 0; is < 7; is++)
#pragma unroll
    for (int m = 0; m < NME; m++) F[m] = F[m] + base[is * NME + m] * prod[is];
}

