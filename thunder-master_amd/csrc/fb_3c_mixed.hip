// Three-centre bcna assembly, mixed precision (M_assemble_3c_DOGS.f90:86-340),
// for s/p species with canonical tables.
//
// Same segment ownership as the binary64 kernels (a workgroup owns the target
// blocks whose triplet segment starts in its tile and sums each segment in the
// reference's order), but the work inside a chunk of triplets is spread over
// all lanes in four phases:
//   A  one lane per triplet: binary64 geometry -- bond and third-atom vectors,
//      x = |r_na - bond centre|, y = bond length, the frame epsilon(rhat,
//      sighat) -- and the reference's grid cell (ix, iy) and method choice
//      (gradient of ME 1 of theta cell 1 against 1e-6, from the binary64
//      table), so every discrete decision equals the reference's.  Stencil
//      weights, Legendre P0..P4 and the p rotation go to LDS in FP32.
//   B  one lane per (triplet, theta, matrix element): 4x4 (or 3-point)
//      stencil of the FP32 slab [iy][ix][theta][ME]; the lanes of one triplet
//      read the 5*nME contiguous values of each stencil point together.
//   C  one lane per triplet: Legendre sum, sin(theta) on mvalue = 1 elements,
//      Slater-Koster rotation (fb_sp.h) in FP32, times P_eq2, charged terms.
//   D  one lane per owned segment: binary64 accumulation of the FP32
//      contributions in segment order; one store of the final block.
// Binary64 adds run at the FP32 rate on gfx950 without packed math, so the
// accumulation is plain binary64 rather than a compensated FP32 sum.
#include "fb_device.h"
#include "fb_sp.h"

#define NB (FB_MO * FB_MO)
#define NME3 10
#define NS3 (FB_NTHETA * NME3)

struct TPrm {
  int base, ncol, rowstride;
  int nme, cls, m1, gsub;     // ME count, pair class, method-1 flag, charged (1) or not (0)
  int p;                      // target pair
  float wx[4], wy[4];         // cubic Lagrange weights (method 3)
  float w1[3];                // (1-px-py, px, py) of the 3-point rule (method 1)
  int m1k[FB_MAXSH];          // method-1 flag of each charged group (its own gradient test)
  float P[5], sint;
  float pm[9];
  float dq[FB_MAXSH];         // shell charges of the third atom
  float stinky, d13, d23, z;
  int gk[FB_MAXSH];           // FP32 slab offset difference of each charged group
};

__device__ __forceinline__ int grp_id(const DView &v, int in1, int in2, int in3, int isub) {
  return v.idx3c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * v.sp.ns + (in3 - 1)) * FB_MAXSUB3 + isub];
}

// 4-point Lagrange weights on nodes -1, 0, 1, 2 (the Horner form of
// M_Fdata_3c.f90:439-477 is this polynomial, divided by 6 per direction)
__device__ __forceinline__ void cubic_w(double p, float *w) {
  double pm1 = p - 1.0, pm2 = p - 2.0, pp1 = p + 1.0;
  w[0] = (float)(-p * pm1 * pm2 / 6.0);
  w[1] = (float)(pp1 * pm1 * pm2 / 2.0);
  w[2] = (float)(-pp1 * p * pm2 / 2.0);
  w[3] = (float)(pp1 * p * pm1 / 6.0);
}

// one (triplet, theta, ME) slot s = theta*nME + ME of the stencil: the lanes of
// a triplet read the 5*nME contiguous values of every stencil point together
__device__ __forceinline__ float interp_slot(const float *__restrict__ tab, int base, int rowstride, int ncol, int s,
                                             const TPrm &q, int m1) {
  if (m1) {
    const float *c00 = tab + base + rowstride + ncol + s;   // (ix, iy)
    float v = q.w1[0] * c00[0];
    v = __builtin_fmaf(q.w1[1], c00[ncol], v);
    return __builtin_fmaf(q.w1[2], c00[rowstride], v);
  }
  const float *c = tab + base + s;
  const float wy0 = q.wy[0], wy1 = q.wy[1], wy2 = q.wy[2], wy3 = q.wy[3];
  float acc = 0.f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const float *col = c + i * ncol;
    float g = wy0 * col[0];
    g = __builtin_fmaf(wy1, col[rowstride], g);
    g = __builtin_fmaf(wy2, col[2 * rowstride], g);
    g = __builtin_fmaf(wy3, col[3 * rowstride], g);
    acc = __builtin_fmaf(q.wx[i], g, acc);
  }
  return acc;
}

template <int CH>
__global__ __launch_bounds__(CH) void k_3c_mixed(DView v) {
  __shared__ TPrm sq[CH];
  __shared__ float sT[CH][NS3 + 1];
  __shared__ float sC[CH][NB + 1];
  const int tid = threadIdx.x;
  const int64_t ts = (int64_t)blockIdx.x * CH;
  const int64_t te = min(ts + CH, v.ntrip);
  int64_t first, last, r0, r1;
  {
    int64_t sgs = v.trip_seg[ts];
    first = (v.seg_off[sgs] == ts) ? sgs : sgs + 1;
    last = v.trip_seg[te - 1];
    if (first > last) return;
    r0 = v.seg_off[first];
    r1 = v.seg_off[last + 1];
  }
  const int64_t my = first + tid;
  const bool own = my <= last;
  int64_t sa = 0, sb = 0;
  if (own) { sa = v.seg_off[my]; sb = v.seg_off[my + 1]; }
  double acc[NB];
#pragma unroll
  for (int i = 0; i < NB; i++) acc[i] = 0.0;
  int status = 0;

  for (int64_t c0 = r0; c0 < r1; c0 += CH) {
    const int nt = (int)min((int64_t)CH, r1 - c0);
    // ---------------- phase A: geometry and stencil parameters
    if (tid < nt) {
      TPrm &q = sq[tid];
      const int64_t t = c0 + tid;
      const int4 tr = v.trip[t];
      const int ial = tr.x, gia = tr.y, gja = tr.z;
      const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
      const int s = v.atom_struct[ial];
      const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
      double r1v[3], r2v[3], rna[3], r21[3], sighat[3], rnabc[3], rhat[3], eps[9];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        r1v[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
        r2v[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
        rna[k] = v.ratom[3 * ial + k];
        r21[k] = r2v[k] - r1v[k];
      }
      const int in1 = v.imass[gia], in2 = v.imass[gja], indna = v.imass[ial];
      q.p = v.trip_pair[t];
      q.cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
      const double z = d_distance(r1v, r2v);
      if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
      else {
#pragma unroll
        for (int k = 0; k < 3; k++) sighat[k] = r21[k] / z;
      }
#pragma unroll
      for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1v[k] + r21[k] / 2.0);
      const double x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
      if (x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
      else {
#pragma unroll
        for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] / x;
      }
      const double cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
      d_epsilon(rhat, sighat, eps);
      float pmf[9];
      d_pmat<float>(eps, pmf);
#pragma unroll
      for (int k = 0; k < 9; k++) q.pm[k] = pmf[k];
      const double cost2 = cost * cost;
      q.sint = (float)sqrt(fmax(1.0e-5, 1.0 - cost2));
      q.P[0] = 1.f;
      q.P[1] = (float)cost;
      q.P[2] = (float)(0.5 * (3.0 * cost2 - 1.0));
      q.P[3] = (float)(0.5 * cost * (5.0 * cost2 - 3.0));
      q.P[4] = (float)(0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2)));
      int gid = grp_id(v, in1, in2, indna, 0);
      if (gid < 0) {
        status |= ST_NO_TABLE;
        gid = 0;
      }
      const DGroup3c g = v.grp3c[gid];
      q.nme = g.nME;
      const double dx = g.xmax / (g.nx - 1.0), dy = g.ymax / (g.ny - 1.0);
      int ix = (int)(x / dx) + 1, iy = (int)(z / dy) + 1;
      if (ix < 2) ix = 2;
      if (iy < 2) iy = 2;
      if (ix > g.nx - 2) ix = g.nx - 2;
      if (iy > g.ny - 2) iy = g.ny - 2;
      const double px = x / dx - ix + 1.0, py = z / dy - iy + 1.0;
      const double *t64 = v.tab3c + g.off;
      const double f00 = t64[((int64_t)(iy - 1) * g.nx + (ix - 1)) * g.nME];
      const double gx = (t64[((int64_t)(iy - 1) * g.nx + ix) * g.nME] - f00) / dx;
      const double gy = (t64[((int64_t)iy * g.nx + (ix - 1)) * g.nME] - f00) / dy;
      const double gg = gx * gx + gy * gy;
      q.m1 = (gg < 1.0e-6) ? 1 : 0;
      float w[4];
      q.w1[0] = (float)(1.0 - px - py);
      q.w1[1] = (float)px;
      q.w1[2] = (float)py;
      cubic_w(px, w);
#pragma unroll
      for (int k = 0; k < 4; k++) q.wx[k] = w[k];
      cubic_w(py, w);
#pragma unroll
      for (int k = 0; k < 4; k++) q.wy[k] = w[k];
      q.ncol = FB_NTHETA * g.nME;
      q.rowstride = g.nx * q.ncol;
      q.base = (int)(g.off32 + ((int64_t)(iy - 2) * g.nx + (ix - 2)) * q.ncol);
      // charged terms (M_assemble_3c_DOGS.f90:250-318)
      q.gsub = 0;
#pragma unroll
      for (int k = 0; k < FB_MAXSH; k++) { q.dq[k] = 0.f; q.gk[k] = 0; q.m1k[k] = 0; }
      for (int isorp = 1; isorp <= v.sp.nssh[indna - 1]; isorp++) {
        double dq = v.dQ[ial * FB_MAXSH + isorp - 1];
        if (dq == 0.0) continue;
        int gk = grp_id(v, in1, in2, indna, isorp);
        if (gk < 0) {
          status |= ST_NO_TABLE;
          continue;
        }
        q.dq[isorp - 1] = (float)dq;
        const DGroup3c gq = v.grp3c[gk];
        q.gk[isorp - 1] = (int)(gq.off32 - g.off32);
        // getMEs_Fdata_3c chooses the method from each table's own gradient
        const double *tq = v.tab3c + gq.off;
        const double h00 = tq[((int64_t)(iy - 1) * g.nx + (ix - 1)) * g.nME];
        const double hx = (tq[((int64_t)(iy - 1) * g.nx + ix) * g.nME] - h00) / dx;
        const double hy = (tq[((int64_t)iy * g.nx + (ix - 1)) * g.nME] - h00) / dy;
        q.m1k[isorp - 1] = (hx * hx + hy * hy < 1.0e-6) ? 1 : 0;
        q.gsub = 1;
      }
      if (q.gsub) {
        const double d13 = d_distance(rna, r1v), d23 = d_distance(rna, r2v);
        double rc1 = v.sp.rcmin[in1 - 1], rc2 = v.sp.rcmin[in2 - 1], rc3 = v.sp.rcmin[indna - 1];
        q.stinky = (float)(d_smoother(d13, rc1 + rc3, 0.8) * d_smoother(d23, rc2 + rc3, 0.8));
        q.d13 = (float)d13;
        q.d23 = (float)d23;
        q.z = (float)z;
      }
    }
    __syncthreads();
    // ---------------- phase B: interpolation, one lane per (triplet, theta, ME)
    for (int it = tid; it < nt * NS3; it += CH) {
      const int t = it / NS3, sl = it - t * NS3;
      const TPrm &q = sq[t];
      if (sl >= q.ncol) continue;
      float F = interp_slot(v.tab3c32, q.base, q.rowstride, q.ncol, sl, q, q.m1);
      if (q.gsub) {
        for (int k = 0; k < FB_MAXSH; k++) {
          const float dq = q.dq[k];
          if (dq == 0.f) continue;
          F = __builtin_fmaf(dq * q.stinky, interp_slot(v.tab3c32, q.base + q.gk[k], q.rowstride, q.ncol, sl, q, q.m1k[k]), F);
        }
      }
      sT[t][sl] = F;
    }
    __syncthreads();
    // ---------------- phase C: rotation, one lane per triplet
    if (tid < nt) {
      const TPrm &q = sq[tid];
      float F[NME3], pm[9], X[16];
      const int nm0 = (q.cls == 3) ? 6 : (q.cls == 0 ? 1 : 2);
#pragma unroll
      for (int m = 0; m < NME3; m++) {
        float f = 0.f;
        if (m < q.nme) {
#pragma unroll
          for (int th = 0; th < FB_NTHETA; th++) f = __builtin_fmaf(q.P[th], sT[tid][th * q.nme + m], f);
          if (m >= nm0) f *= q.sint;
        }
        F[m] = f;
      }
#pragma unroll
      for (int k = 0; k < 9; k++) pm[k] = q.pm[k];
#pragma unroll
      for (int i = 0; i < 16; i++) X[i] = 0.f;
      switch (q.cls) {
        case 3: rot3c_sp<float>(3, pm, F, X); break;
        case 2: rot3c_sp<float>(2, pm, F, X); break;
        case 1: rot3c_sp<float>(1, pm, F, X); break;
        default: rot3c_sp<float>(0, pm, F, X); break;
      }
      const int nb = (q.cls & 2 ? 4 : 1) * (q.cls & 1 ? 4 : 1);
      if (q.gsub) {
        float dqs = 0.f;
#pragma unroll
        for (int k = 0; k < FB_MAXSH; k++) dqs += q.dq[k];
        const double *sbp = v.mat[FB_MAT_S] + v.boff[q.p];
        const double *dbp = v.mat[FB_MAT_DIP] + v.boff[q.p];
#pragma unroll
        for (int e = 0; e < 16; e++) {
          if (e >= nb) break;
          double sterm = sbp[e] / 2.0, dterm = dbp[e] / (double)q.z;
          double emnpl = (sterm - dterm) / (double)q.d13 + (sterm + dterm) / (double)q.d23;
          X[e] = X[e] * (float)FB_P_EQ2 + (float)(dqs * (1.0 - q.stinky) * emnpl * FB_P_EQ2);
        }
#pragma unroll
        for (int e = 0; e < 16; e++) sC[tid][e] = (e < nb) ? X[e] : 0.f;
      } else {
#pragma unroll
        for (int e = 0; e < 16; e++) sC[tid][e] = X[e] * (float)FB_P_EQ2;
      }
    }
    __syncthreads();
    // ---------------- phase D: ordered binary64 accumulation per owned segment
    if (own) {
      int64_t k0 = max(sa, c0), k1 = min(sb, c0 + nt);
      for (int64_t k = k0; k < k1; k++)
#pragma unroll
        for (int i = 0; i < NB; i++) acc[i] += (double)sC[k - c0][i];
    }
    __syncthreads();
  }
  if (own) {
    const int p = v.seg_pair[my];
    const int ia = v.pair_atom[p];
    const int s = v.atom_struct[ia];
    const int ja = v.st[s].atom_off + v.neigh_j[p] - 1;
    const int nb = v.sp.norb[v.imass[ia] - 1] * v.sp.norb[v.imass[ja] - 1];
    double *dst = v.mat[FB_MAT_VNA] + v.boff[p];
#pragma unroll
    for (int i = 0; i < NB; i++)
      if (i < nb) dst[i] = acc[i];
  }
  if (status) atomicOr(v.status, status);
}

extern "C" int fbk_3c_mixed(const DView &v, int chunk, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  if (chunk == 256) {
    hipLaunchKernelGGL(k_3c_mixed<256>, dim3((unsigned)((v.ntrip + 255) / 256)), dim3(256), 0, st, v);
  } else if (chunk == 64) {
    hipLaunchKernelGGL(k_3c_mixed<64>, dim3((unsigned)((v.ntrip + 63) / 64)), dim3(64), 0, st, v);
  } else {
    hipLaunchKernelGGL(k_3c_mixed<128>, dim3((unsigned)((v.ntrip + 127) / 128)), dim3(128), 0, st, v);
  }
  return (int)hipGetLastError();
}
