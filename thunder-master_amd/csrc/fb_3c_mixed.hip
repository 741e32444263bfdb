// Three-centre bcna assembly, mixed precision (M_assemble_3c_DOGS.f90:86-340),
// for s/p species with canonical tables.
//
// A workgroup owns whole target segments (the triplets of one Vna block, in
// the reference's order), packed at list build into tiles of at most one
// chunk of triplets, so that it usually makes a single pass.  The work inside
// a chunk is spread over all lanes in four phases:
//   A  one lane per triplet: geometry.  Positions, bond and third-atom vectors,
//      x = |r_na - bond centre| and y = bond length are binary64 with the
//      reference's expressions, and so are the decisions taken on them: the
//      grid cell (ix, iy) = int(x/dx)+1, int(y/dy)+1, the method choice
//      (gradient of ME 1 of theta cell 1 of each table against 1e-6, from the
//      binary64 table) and the thresholds of the frame.  The frame
//      epsilon(rhat, sighat), cos theta and sin theta only feed FP32 values:
//      they come from the binary64 cross and dot products of the bond and
//      third-atom vectors, scaled in FP32; triplets whose threshold decisions
//      the products cannot settle take the reference's binary64 steps.
//   B  one lane per (triplet, quad of matrix elements): the 16 (or 3)
//      stencil points of the FP32 quad slab [iy][ix][theta][ME quad], five
//      float4 per point, Legendre-folded at each point, then the x/y
//      Lagrange weights.
//   C  one lane per triplet: sin(theta) on mvalue = 1 elements, Slater-Koster
//      rotation (fb_sp.h) in FP32, times P_eq2, charged terms.
//   D  one lane per (owned segment, block element): binary64 accumulation of
//      the FP32 contributions in segment order from LDS, stored to the block.
// Binary64 adds run at the FP32 rate on gfx950 without packed math, so the
// accumulation is plain binary64 rather than a compensated FP32 sum.
//
// k_3c_mlp is the learned-surrogate form: a persistent launch of 256-thread
// workgroups (four per CU) that pull one-pass tiles of up to 256 triplets from
// a work queue (one range per XCD, the others when it runs dry).  Phase B
// runs the stencil only for triplets whose calibrated network error exceeds
// the bound; phase M evaluates the network for the rest on MFMA, 16 triplets
// per wavefront pass, reading each weight fragment from global memory (L1)
// at its MFMA.
#include "fb_device.h"
#include "fb_mlp.h"
#include "fb_sp.h"
#include "fb_2c_sp.h"

#define NB (FB_MO * FB_MO)
#define NME3 10

// The rotation of rot3c_sp (fb_sp.h) in FP32 with fused multiply-adds, the
// p-p block as (P M) P^T: 42 operations instead of 126 for an sp-sp block
#define PPF(k, m) pm[3 * (m) + (k)]
template <int CLS>
__device__ __forceinline__ void rot3c_fma(const float *pm, const float *F, float *X) {
  X[0] = F[0];
  if (CLS == CLS_SPSP) {
    const float b = F[1], c = F[2], d = F[3], e = F[4], f = F[5], g = F[6], h = F[7], i = F[8], j = F[9];
#pragma unroll
    for (int l = 0; l < 3; l++) X[(1 + l) * 4] = __builtin_fmaf(g, PPF(l, 2), b * PPF(l, 1));
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = __builtin_fmaf(PPF(k, 2), h, PPF(k, 1) * c);
    float T0[3], T1[3], T2[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      T0[k] = PPF(k, 0) * d;
      T1[k] = __builtin_fmaf(PPF(k, 2), i, PPF(k, 1) * e);
      T2[k] = __builtin_fmaf(PPF(k, 2), f, PPF(k, 1) * j);
    }
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int k = 0; k < 3; k++)
        X[(1 + l) * 4 + 1 + k] = __builtin_fmaf(T2[k], PPF(l, 2), __builtin_fmaf(T1[k], PPF(l, 1), T0[k] * PPF(l, 0)));
  } else if (CLS == CLS_SSP) {
#pragma unroll
    for (int l = 0; l < 3; l++) X[1 + l] = __builtin_fmaf(F[2], PPF(l, 2), F[1] * PPF(l, 1));
  } else if (CLS == CLS_SPS) {
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = __builtin_fmaf(PPF(k, 2), F[2], PPF(k, 1) * F[1]);
  }
}
#undef PPF

struct TPrm {
  unsigned base;              // quad slab index (floats) of stencil corner (ix-1, iy-1), neutral group
  unsigned rsc;               // floats per iy row of the quad slab (bits 0-23), pair class (24-25), method-1 flag (26)
  float px, py;               // position in the cell; phase B forms the weights
  float P[4], sint;           // Legendre P1..P4 (P0 = 1)
  float pm[6];                // columns 0 and 1 of the p rotation; column 2 is their cross product
  __device__ int rowstride() const { return (int)(rsc & 0xffffffu); }
  __device__ int cls() const { return (int)(rsc >> 24) & 3; }
  __device__ bool m1() const { return (rsc >> 26) & 1u; }
  __device__ int nme() const { return nme3c_of(cls()); }   // canonical s/p tables
  __device__ int nq() const { return cls() == CLS_SPSP ? 3 : 1; }
};
// 15 words: an odd row stride, so lane-per-triplet accesses hit distinct banks
static_assert(sizeof(TPrm) == 15 * 4, "TPrm row stride must stay odd in words");

// the group table through the constant address space: uniform indices load
// with scalar loads
typedef const __attribute__((address_space(4))) DGroup3c CGroup3c;

// The constants of group gid (per lane), by scalar loads: one pass per
// distinct group of the wavefront (the group table is read-only during a
// launch).  The empty asm uses the loaded fields inside the loop, where the
// index is a scalar; without it the loads sink past the loop and address
// with the per-lane index.
__device__ __forceinline__ DGroup3c group_scalar(const DView &v, int gid) {
#ifndef __HIP_DEVICE_COMPILE__
  return v.grp3c[gid];   // the host pass only parses device code
#else
  const CGroup3c *cg = (const CGroup3c *)v.grp3c;
  DGroup3c g;
  for (;;) {
    const int gc = __builtin_amdgcn_readfirstlane(gid);
    const DGroup3c t = cg[gc];
    asm volatile("" ::"s"(t.nME), "s"(t.nx), "s"(t.ny), "s"(t.cls), "s"(t.nq), "s"(t.xmax), "s"(t.ymax), "s"(t.dx),
                 "s"(t.dy), "s"(t.rdx), "s"(t.rdy), "s"(t.off32q), "s"(t.moff));
    if (gid == gc) {
      g = t;
      break;
    }
  }
  return g;
#endif
}

struct TChg {                 // charged third atom (DOGS terms)
  float dq[FB_MAXSH];
  int gk[FB_MAXSH];           // slab offset of each charged group relative to the neutral one
  int m1k[FB_MAXSH];
  float stinky, d13, d23, z, dqs;
};

__device__ __forceinline__ int grp_id(const DView &v, int in1, int in2, int in3, int isub) {
  return v.idx3c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * v.sp.ns + (in3 - 1)) * FB_MAXSUB3 + isub];
}

// 4-point Lagrange weights on nodes -1, 0, 1, 2 (the Horner form of
// M_Fdata_3c.f90:439-477 is this polynomial, divided by 6 per direction)
__device__ __forceinline__ void cubic_w(float p, float *w) {
  const float pm1 = p - 1.f, pm2 = p - 2.f, pp1 = p + 1.f, s6 = 1.f / 6.f;
  w[0] = -p * pm1 * pm2 * s6;
  w[1] = pp1 * pm1 * pm2 * 0.5f;
  w[2] = -pp1 * p * pm2 * 0.5f;
  w[3] = pp1 * p * pm1 * s6;
}

// Buffer resource over the FP32 slab: loads take one 32-bit byte offset in a
// VGPR plus an immediate, and addresses outside the slab read as zero.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slab_rsrc(const float *tab, int64_t nfloats) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)tab, (short)0, (int)(nfloats * 4), 0x00020000);
}

__device__ __forceinline__ float4 ld4(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
}

// Legendre fold of one stencil point for one ME quad: sum_theta P_theta
// T[theta][quad]; off in bytes, TS bytes per theta row of a point
template <unsigned TS>
__device__ __forceinline__ float4 fold4(__amdgpu_buffer_rsrc_t r, unsigned off, const float *P) {
  float4 v = ld4(r, off);
#pragma unroll
  for (int t = 1; t < 5; t++) {
    const float4 a = ld4(r, off + t * TS);
    v.x = __builtin_fmaf(P[t], a.x, v.x);
    v.y = __builtin_fmaf(P[t], a.y, v.y);
    v.z = __builtin_fmaf(P[t], a.z, v.z);
    v.w = __builtin_fmaf(P[t], a.w, v.w);
  }
  return v;
}

__device__ __forceinline__ float4 axpy4(float w, float4 a, float4 y) {
  return make_float4(__builtin_fmaf(w, a.x, y.x), __builtin_fmaf(w, a.y, y.y), __builtin_fmaf(w, a.z, y.z),
                     __builtin_fmaf(w, a.w, y.w));
}

// one (triplet, ME quad): 16-point tensor stencil (method 3) or 3-point rule
// (M1, method 1); o: byte offset of the quad at stencil corner (ix-1, iy-1),
// rb: bytes per iy row; NQ quads per point.  Offsets within a row are
// immediates of the loads.
template <int NQ, bool M1>
__device__ __forceinline__ float4 stencil4(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned rb, float px, float py,
                                           const float *P) {
  constexpr unsigned TS = 16 * NQ, PT = 5 * TS;   // bytes per theta row, per ix step
  if (M1) {
    const unsigned c = o + rb + PT;   // (ix, iy)
    const float4 f0 = fold4<TS>(r, c, P), f1 = fold4<TS>(r, c + PT, P), f2 = fold4<TS>(r, c + rb, P);
    const float w0 = (1.f - px) - py;
    return axpy4(py, f2, axpy4(px, f1, make_float4(w0 * f0.x, w0 * f0.y, w0 * f0.z, w0 * f0.w)));
  }
  float wx[4], wy[4];
  cubic_w(px, wx);
  cubic_w(py, wy);
  float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int j = 0; j < 4; j++) {
    const unsigned row = o + (unsigned)j * rb;
    const float4 g0 = fold4<TS>(r, row, P), g1 = fold4<TS>(r, row + PT, P);
    float4 g = make_float4(wx[0] * g0.x, wx[0] * g0.y, wx[0] * g0.z, wx[0] * g0.w);
    g = axpy4(wx[1], g1, g);
    const float4 g2 = fold4<TS>(r, row + 2 * PT, P), g3 = fold4<TS>(r, row + 3 * PT, P);
    g = axpy4(wx[2], g2, g);
    g = axpy4(wx[3], g3, g);
    acc = axpy4(wy[j], g, acc);
  }
  return acc;
}

template <int NQ>
__device__ __forceinline__ float4 stencil4(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned rb, float px, float py,
                                           const float *P, int m1) {
  return m1 ? stencil4<NQ, true>(r, o, rb, px, py, P) : stencil4<NQ, false>(r, o, rb, px, py, P);
}

// One point (ix-2+i, iy-2+j) of the 16-point stencil times its weights,
// wx_i wy_j fold(i, j); o: byte offset of the quad at the stencil corner
template <int NQ>
__device__ __forceinline__ float4 stencil_point(__amdgpu_buffer_rsrc_t r, unsigned o, unsigned rb, float px, float py,
                                                int i, int j, const float *P) {
  constexpr unsigned TS = 16 * NQ, PT = 5 * TS;
  float wx[4], wy[4];
  cubic_w(px, wx);
  cubic_w(py, wy);
  const float w = (i == 0 ? wx[0] : (i == 1 ? wx[1] : (i == 2 ? wx[2] : wx[3]))) *
                  (j == 0 ? wy[0] : (j == 1 ? wy[1] : (j == 2 ? wy[2] : wy[3])));
  const float4 f = fold4<TS>(r, o + (unsigned)j * rb + (unsigned)i * PT, P);
  return make_float4(w * f.x, w * f.y, w * f.z, w * f.w);
}

// sum over the 16 lanes 16m..16m+15 of the wavefront, every lane receiving it
__device__ __forceinline__ float sum16(float v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 4);
  return v + __shfl_xor(v, 8);
}
__device__ __forceinline__ float4 sum16(float4 v) { return make_float4(sum16(v.x), sum16(v.y), sum16(v.z), sum16(v.w)); }

// Barrier of one tile pass: a 64-triplet tile is one wavefront, which needs
// only ordering of its LDS accesses (so that several such tiles can share a
// workgroup, one per wavefront, each with its own pass count)
template <int CH>
__device__ __forceinline__ void tile_sync() {
  if (CH == 64) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  } else {
    __syncthreads();
  }
}

// lane of a tile: a 64-triplet tile is one wavefront of its workgroup
template <int CH>
__device__ __forceinline__ int tile_tid() {
  return CH == 64 ? (int)(threadIdx.x & 63u) : (int)threadIdx.x;
}

// LDS of one tile pass
template <int CH, bool CHARGED, bool MLP>
struct S3c {
  TPrm sq[CH];
  float sF[CH][NME3];                 // molecular-frame MEs (phase C's contributions reuse sq and sF)
  TChg sx[CHARGED ? CH : 1];
  int sSeg[CH + 1];                   // triplet offsets (below 2^31: the list build checks)
  int64_t sBo[CH];
  unsigned char sNb[CH];
  float sU[MLP ? CH : 1][4];          // network inputs and envelope
  short sG[MLP ? CH : 1];             // network group of the triplet, -1: table stencil
  short sNet[MLP ? CH : 1];           // chunk triplets on the network, in order
  short sFb[CH];                      // chunk triplets on the stencil, in four classes (see tile3c)
  int sWn[CH / 64][5];                // per-wavefront counts of the network list and the four classes
};

// Phase B over one class of the stencil list, entries [kb, kb + n) of sFb,
// NQ lanes per triplet (one per ME quad).  M1, the 3-point rule: one lane per
// quad.  The 16-point stencil: 16 lanes per quad, one per stencil point,
// whose weighted folds are summed across the lanes (a point's five loads
// are then all the latency a lane waits for).  Charged subtypes choose
// their method per table.
template <int NQ, bool M1, int CH, bool CHARGED, bool MLP>
__device__ __forceinline__ void interp_class(S3c<CH, CHARGED, MLP> &S, __amdgpu_buffer_rsrc_t tab, int kb, int n) {
  constexpr int R = M1 ? 1 : 16;   // lanes per (triplet, quad)
  for (int it = tile_tid<CH>(); it < R * NQ * n; it += CH) {
    const int kq = it / R, p = it - R * kq;
    const int k = kq / NQ, qd = kq - NQ * k;
    const int t = S.sFb[kb + k];
    const TPrm &q = S.sq[t];
    const float P[5] = {1.f, q.P[0], q.P[1], q.P[2], q.P[3]};
    const float px = q.px, py = q.py;
    const unsigned o = (q.base + 4u * (unsigned)qd) * 4u, rb = (unsigned)q.rowstride() * 4u;
    float4 F = M1 ? stencil4<NQ, true>(tab, o, rb, px, py, P)
                  : sum16(stencil_point<NQ>(tab, o, rb, px, py, p & 3, p >> 2, P));
    if (CHARGED) {
      const TChg &c = S.sx[t];
      for (int kk = 0; kk < FB_MAXSH; kk++) {
        const float dq = c.dq[kk];
        if (dq == 0.f) continue;
        const unsigned ok = o + (unsigned)c.gk[kk] * 4u;
        float4 Fk;
        if (M1) Fk = stencil4<NQ>(tab, ok, rb, px, py, P, c.m1k[kk]);
        // all 16 lanes of the quad take the same branch: the subtype's method is per triplet
        else if (c.m1k[kk]) Fk = stencil4<NQ, true>(tab, ok, rb, px, py, P);
        else Fk = sum16(stencil_point<NQ>(tab, ok, rb, px, py, p & 3, p >> 2, P));
        F = axpy4(dq * c.stinky, Fk, F);
      }
    }
    if (p == 0) {
      // quad 2 of an sp-sp triplet holds MEs 8, 9 and two padding zeros
      float2 *const d = reinterpret_cast<float2 *>(&S.sF[t][4 * qd]);
      d[0] = make_float2(F.x, F.y);
      if (qd < 2) d[1] = make_float2(F.z, F.w);
    }
  }
}

// One tile of whole target segments (first, last, r0, r1), in chunks of at
// most CH triplets.  MLP: the learned surrogate replaces the stencil of phase
// B for every triplet whose calibrated bound passes (neutral builds only).
template <int CH, bool CHARGED, bool MLP>
__device__ __forceinline__ void tile3c(const DView &v, const int4 td, S3c<CH, CHARGED, MLP> &S, unsigned *sCnt,
                                       int &status) {
  // contributions reuse the parameter storage, rows padded to 17 words
  float(*sC)[NB + 1] = reinterpret_cast<float(*)[NB + 1]>(&S.sq[0]);
  typedef S3c<CH, CHARGED, MLP> S_t;
  static_assert(offsetof(S_t, sF) == sizeof(S.sq) && sizeof(S.sq) + sizeof(S.sF) >= CH * (NB + 1) * sizeof(float),
                "the contribution rows must fit in the parameter and ME rows");
  const int tid = tile_tid<CH>();
  const int64_t first = td.x, last = td.y, r0 = td.z, r1 = td.w;
  if (first > last) return;
  // Accumulator slot sl = (segment, block element); lane tid owns the slots
  // tid + j*CH, j < NB.  The workgroup owns the target blocks, so a segment
  // that spans chunks keeps its binary64 partial sum in the block itself.
  const int nown = (int)(last - first + 1);
  for (int k = tid; k <= nown; k += CH) S.sSeg[k] = (int)v.seg_off[first + k];
  for (int k = tid; k < nown; k += CH) {
    const int p = v.seg_pair[first + k];
    const int64_t b0 = v.boff[p];
    S.sBo[k] = b0;
    S.sNb[k] = (unsigned char)(v.boff[p + 1] - b0);
  }
  double *const vna = v.mat[FB_MAT_VNA];
  const __amdgpu_buffer_rsrc_t tab = slab_rsrc(v.tab3c32q, v.ntab3c32q);

  for (int64_t c0 = r0; c0 < r1; c0 += CH) {
    const int nt = (int)min((int64_t)CH, r1 - c0);
    // ---------------- phase A: geometry and stencil parameters
    if (tid < nt) {
      TPrm &q = S.sq[tid];
      const int64_t t = c0 + tid;
      const int4 tr = v.trip[t];
      const int aux = v.trip_aux[t];
      const int ial = tr.x, gia = tr.y, gja = tr.z;
      const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
      const int s = aux >> 16;
      const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
      double r1v[3], r2v[3], rna[3], r21[3], sighat[3], rnabc[3], rhat[3];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        r1v[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
        r2v[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
        rna[k] = v.ratom[3 * ial + k];
        r21[k] = r2v[k] - r1v[k];
      }
      int gid = aux & 0xffff;
      if (gid == 0xffff) gid = -1;
      if (gid < 0) {
        status |= ST_NO_TABLE;
        gid = 0;
      }
      const DGroup3c g = group_scalar(v, gid);
      const double z = d_distance(r1v, r2v);
#pragma unroll
      for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1v[k] + r21[k] * 0.5);
      const double x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
      // |sighat x rhat| = |c| / (z x) with c = r21 x rnabc, and cos theta =
      // (r21 . rnabc) / (z x).  epsilon's collinearity branch (|sighat x
      // rhat| < 1e-4) is c.c < 1e-8 (z x)^2; within a relative 1e-6 of that
      // bound, and when z or x is below the 1e-5 of the unit vectors, the
      // reference's own binary64 steps decide.
      double cr[3];
      d_cross(r21, rnabc, cr);
      const double c2 = cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2];
      const double zx = z * x, lim = 1.0e-8 * (zx * zx);
      const bool fast = z >= 1.0e-5 && x >= 1.0e-5 && c2 > lim && c2 - lim > 1.0e-6 * lim;
      float eps[9], cosf_, sint;
      if (fast) {
        const float izx = __builtin_amdgcn_rcpf((float)zx), iz = __builtin_amdgcn_rcpf((float)z);
        const float ic = __builtin_amdgcn_rsqf((float)c2);
        const float dt = (float)(r21[0] * rnabc[0] + r21[1] * rnabc[1] + r21[2] * rnabc[2]);
        float zp[3], yp[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          zp[k] = (float)r21[k] * iz;
          yp[k] = (float)cr[k] * ic;
        }
        eps[0] = yp[1] * zp[2] - yp[2] * zp[1];
        eps[1] = yp[2] * zp[0] - yp[0] * zp[2];
        eps[2] = yp[0] * zp[1] - yp[1] * zp[0];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          eps[3 + k] = yp[k];
          eps[6 + k] = zp[k];
        }
        cosf_ = dt * izx;
        // sqrt(max(1e-5, 1 - cos^2)) of M_Fdata_3c.f90:349 as max(sqrt(1e-5), |sin|)
        sint = fmaxf(__builtin_amdgcn_sqrtf((float)c2) * izx, 3.16227766e-3f);
      } else {
        if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
        else {
          const double iz_ = 1.0 / z;
#pragma unroll
          for (int k = 0; k < 3; k++) sighat[k] = r21[k] * iz_;
        }
        if (x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
        else {
          const double ix_ = 1.0 / x;
#pragma unroll
          for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] * ix_;
        }
        const double cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
        // epsilon(rhat, sighat), b.FUNCTIONS/epsilon.f90:120-177, with the
        // reference's thresholds (sighat is a unit vector, so the r2 = 0 branch
        // cannot occur); sighat and rhat are unit vectors already (rhat = 0
        // when x < 1e-5), so the normalisations of epsilon.f90:136-145 change
        // them by rounding only
        double zp[3] = {sighat[0], sighat[1], sighat[2]}, yp[3] = {0.0, 0.0, 0.0}, xp[3];
        double ypm = 0.0;
        if (x >= 1.0e-5) {
          d_cross(zp, rhat, yp);
          ypm = d_mag(yp);
        }
        if (ypm < 1.0e-4) {
          if (fabs(zp[0]) > 1.0e-4) { yp[0] = -(zp[1] + zp[2]) / zp[0]; yp[1] = 1.0; yp[2] = 1.0; }
          else if (fabs(zp[1]) > 1.0e-4) { yp[0] = 1.0; yp[1] = -(zp[0] + zp[2]) / zp[1]; yp[2] = 1.0; }
          else { yp[0] = 1.0; yp[1] = 1.0; yp[2] = -(zp[0] + zp[1]) / zp[2]; }
          ypm = d_mag(yp);
        }
        const double iy_ = 1.0 / ypm;
#pragma unroll
        for (int k = 0; k < 3; k++) yp[k] *= iy_;
        d_cross(yp, zp, xp);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          eps[k] = (float)xp[k];
          eps[3 + k] = (float)yp[k];
          eps[6 + k] = (float)zp[k];
        }
        cosf_ = (float)cost;
        sint = (float)sqrt(fmax(1.0e-5, 1.0 - cost * cost));
      }
      {
        float pmf[9];
        d_pmat<float>(eps, pmf);
#pragma unroll
        for (int k = 0; k < 6; k++) q.pm[k] = pmf[k];
      }
      const float cost = cosf_, cost2 = cost * cost;
      q.sint = sint;
      q.P[0] = cost;
      q.P[1] = 0.5f * (3.f * cost2 - 1.f);
      q.P[2] = 0.5f * cost * (5.f * cost2 - 3.f);
      q.P[3] = 0.125f * (3.f - cost2 * (30.f - 35.f * cost2));
      // grid cell and method, M_Fdata_3c.f90:367-385: ix = int(x/dx) + 1.
      // x (1/dx) lies within 4e-16 relative of x/dx, so its integer part is
      // that of the quotient unless it falls within 1e-12 of an integer; then
      // the division is taken.
      double qx = x * g.rdx, qy = z * g.rdy;
      {
        const double fx = qx - floor(qx), fy = qy - floor(qy);
        const bool nx_ = fx < 1.0e-12 || fx > 1.0 - 1.0e-12, ny_ = fy < 1.0e-12 || fy > 1.0 - 1.0e-12;
        // a branch of the wavefront (not a select): the quotients are rare
        if (__builtin_amdgcn_ballot_w64(nx_ || ny_) != 0) {
          if (nx_) qx = x / g.dx;
          if (ny_) qy = z / g.dy;
        }
      }
      int ix = (int)qx + 1, iy = (int)qy + 1;
      if (ix < 2) ix = 2;
      if (iy < 2) iy = 2;
      if (ix > g.nx - 2) ix = g.nx - 2;
      if (iy > g.ny - 2) iy = g.ny - 2;
      const double px = qx - ix + 1.0, py = qy - iy + 1.0;
      // method of the cell, precomputed from the binary64 table with the
      // reference's gradient test (M_Fdata_3c.f90:383-385)
      const int64_t icell = (int64_t)(iy - 1) * g.nx + (ix - 1);
      q.rsc = (unsigned)(g.nx * 20 * g.nq) | ((unsigned)g.cls << 24) | ((unsigned)v.meth3c[g.moff + icell] << 26);
      q.px = (float)px;
      q.py = (float)py;
      q.base = (unsigned)(g.off32q + ((int64_t)(iy - 2) * g.nx + (ix - 2)) * 20 * g.nq);
      if (MLP) {
        // per-triplet bound: P_eq2 x the calibrated Frobenius error of the
        // sub-cell's molecular-frame block (bounds every rotated element),
        // with a factor 2 of margin over the sampled maximum
        const DMlp *w = v.mlp + gid;
        bool ok = w->ready && x <= g.xmax && z <= g.ymax;
        if (ok) {
          const float ce = v.mlp_err[w->err_off + mlp_subcell(x, z, g.xmax, g.ymax, g.nx, g.ny)];
          ok = 2.0f * (float)FB_P_EQ2 * ce <= v.mlp_tau;
        }
        S.sU[tid][0] = (float)(2.0 * x / g.xmax - 1.0);
        S.sU[tid][1] = (float)(2.0 * z / g.ymax - 1.0);
        S.sU[tid][2] = cost;
        S.sU[tid][3] = ok ? mlp_envelope(w, (float)x, (float)z) : 0.f;
        S.sG[tid] = (short)(ok ? gid : -1);
        atomicAdd(&sCnt[0], ok ? 1u : 0u);
        atomicAdd(&sCnt[1], 1u);
      }
      if (CHARGED) {
        const int in1 = v.imass[gia], in2 = v.imass[gja], indna = v.imass[ial];
        TChg &c = S.sx[tid];
        c.dqs = 0.f;
#pragma unroll
        for (int k = 0; k < FB_MAXSH; k++) { c.dq[k] = 0.f; c.gk[k] = 0; c.m1k[k] = 0; }
        for (int isorp = 1; isorp <= v.sp.nssh[indna - 1]; isorp++) {
          double dq = v.dQ[ial * FB_MAXSH + isorp - 1];
          if (dq == 0.0) continue;
          int gk = grp_id(v, in1, in2, indna, isorp);
          if (gk < 0) {
            status |= ST_NO_TABLE;
            continue;
          }
          const DGroup3c gq = v.grp3c[gk];
          c.dq[isorp - 1] = (float)dq;
          c.dqs += (float)dq;
          c.gk[isorp - 1] = (int)(gq.off32q - g.off32q);
          // getMEs_Fdata_3c chooses the method from each table's own gradient
          c.m1k[isorp - 1] = v.meth3c[gq.moff + icell];
        }
        const double d13 = d_distance(rna, r1v), d23 = d_distance(rna, r2v);
        double rc1 = v.sp.rcmin[in1 - 1], rc2 = v.sp.rcmin[in2 - 1], rc3 = v.sp.rcmin[indna - 1];
        c.stinky = (float)(d_smoother(d13, rc1 + rc3, 0.8) * d_smoother(d23, rc2 + rc3, 0.8));
        c.d13 = (float)d13;
        c.d23 = (float)d23;
        c.z = (float)z;
      }
    } else if (MLP && tid < CH) {
      S.sG[tid] = -1;   // lanes past the chunk end take no network pass
    }
    // The chunk's triplets split into a network list and a stencil list, each
    // dense and in triplet order, so that each phase costs in proportion to
    // its share.  The stencil list holds four classes one after another:
    // 3-point cells with 3 ME quads, then with 1, then 16-point cells with 3
    // and with 1.  Phase B then gives each triplet as many lanes as it has
    // quads, and a wavefront rarely mixes a 16-point cell (five times the
    // work) or the other quad count into its lanes.
    int nnet = 0, nc[4] = {0, 0, 0, 0};
    {
      const int wv = tid >> 6, l = tid & 63;
      const bool on = MLP && tid < nt && S.sG[tid] >= 0, st = tid < nt && !on;
      const int cl = st ? (S.sq[tid].m1() ? 0 : 2) + (S.sq[tid].nq() == 3 ? 0 : 1) : -1;
      const unsigned long long bn = __ballot(on);
      unsigned long long bc[4];
#pragma unroll
      for (int c = 0; c < 4; c++) bc[c] = __ballot(cl == c);
      int pn = 0, pc[4] = {0, 0, 0, 0};
      if (CH == 64) {
        nnet = __popcll(bn);
#pragma unroll
        for (int c = 0; c < 4; c++) nc[c] = __popcll(bc[c]);
      } else {
        if (l == 0) {
          S.sWn[wv][0] = __popcll(bn);
#pragma unroll
          for (int c = 0; c < 4; c++) S.sWn[wv][1 + c] = __popcll(bc[c]);
        }
        tile_sync<CH>();
#pragma unroll
        for (int w = 0; w < CH / 64; w++) {
          if (w < wv) {
            pn += S.sWn[w][0];
#pragma unroll
            for (int c = 0; c < 4; c++) pc[c] += S.sWn[w][1 + c];
          }
          nnet += S.sWn[w][0];
#pragma unroll
          for (int c = 0; c < 4; c++) nc[c] += S.sWn[w][1 + c];
        }
      }
      const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
      if (on) S.sNet[pn + __popcll(bn & below)] = (short)tid;
      if (cl >= 0) {
        int base = 0;
#pragma unroll
        for (int c = 0; c < 4; c++)
          if (c < cl) base += nc[c];
        S.sFb[base + pc[cl] + __popcll(bc[cl] & below)] = (short)tid;
      }
    }
    tile_sync<CH>();
    // ---------------- phase B: interpolation, one lane per (triplet, ME
    // quad), the four classes of the stencil list one after another
    interp_class<3, true>(S, tab, 0, nc[0]);
    interp_class<1, true>(S, tab, nc[0], nc[1]);
    interp_class<3, false>(S, tab, nc[0] + nc[1], nc[2]);
    interp_class<1, false>(S, tab, nc[0] + nc[1] + nc[2], nc[3]);
    if (MLP) {
      // keeps the weight loads of phase M from being hoisted into the stencil
      // loop (their registers would overlap)
      tile_sync<CH>();
      // ---------------- phase M: network evaluation, 16 triplets per
      // wavefront pass; lane group l/16 folds the matrix elements l/16 + 4k of
      // its triplet (the theta cells times P_theta(cos) and the envelope)
      const int wv = tid >> 6, l = tid & 63, qg = l >> 4;
      for (int blk = wv; blk * 16 < nnet; blk += CH / 64) {
        const int k = blk * 16 + (l & 15);
        const int t = (k < nnet) ? (int)S.sNet[k] : 0;
        const int gsel = (k < nnet) ? S.sG[t] : -1;
        unsigned long long pend = __ballot(gsel >= 0);
        while (pend) {
          const int lead = __ffsll((long long)pend) - 1;
          const int gnow = __shfl(gsel, lead);
          const MlpW *w = &v.mlp[gnow].w;
          const bool mine = (gsel == gnow);
          float out[16];
          mlp_eval16(w, mine ? S.sU[t][0] : 0.f, mine ? S.sU[t][1] : 0.f, mine ? S.sU[t][2] : 0.f, out);
          if (mine) {
            const TPrm &q = S.sq[t];
            const float env = S.sU[t][3];
            const float P[5] = {env, q.P[0] * env, q.P[1] * env, q.P[2] * env, q.P[3] * env};
#pragma unroll
            for (int ml = 0; ml < 3; ml++) {
              const int m = qg + 4 * ml;
              if (m < q.nme()) {
                float F = P[0] * out[ml * 5];
#pragma unroll
                for (int th = 1; th < 5; th++) F = __builtin_fmaf(P[th], out[ml * 5 + th], F);
                S.sF[t][m] = F;
              }
            }
          }
          pend &= ~__ballot(mine);
        }
      }
    }
    tile_sync<CH>();
    // ---------------- phase C: rotation, one lane per triplet; P_eq2 (and
    // sin theta on the mvalue = 1 elements) scale the MEs before it
    float X[16];
    int cls = 0, p = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) X[i] = 0.f;
    if (tid < nt) {
      const TPrm &q = S.sq[tid];
      float F[NME3], pm[9];
      cls = q.cls();
      if (CHARGED) p = v.trip_pair[c0 + tid];
      const int nm0 = (cls == 3) ? 6 : (cls == 0 ? 1 : 2);
      const float s0 = (float)FB_P_EQ2, s1 = (float)FB_P_EQ2 * q.sint;
      const int nme = q.nme();
#pragma unroll
      for (int m = 0; m < NME3; m++) F[m] = (m < nme) ? S.sF[tid][m] * (m >= nm0 ? s1 : s0) : 0.f;
#pragma unroll
      for (int k = 0; k < 6; k++) pm[k] = q.pm[k];
      // column 2 = column 0 x column 1 (a rotation; (y, z, x) is a cyclic order)
      pm[6] = pm[1] * pm[5] - pm[2] * pm[4];
      pm[7] = pm[2] * pm[3] - pm[0] * pm[5];
      pm[8] = pm[0] * pm[4] - pm[1] * pm[3];
      switch (cls) {
        case 3: rot3c_fma<3>(pm, F, X); break;
        case 2: rot3c_fma<2>(pm, F, X); break;
        case 1: rot3c_fma<1>(pm, F, X); break;
        default: X[0] = F[0]; break;
      }
    }
    tile_sync<CH>();   // every lane has read its parameters before sC overwrites them
    if (tid < nt) {
      const int nb = (cls & 2 ? 4 : 1) * (cls & 1 ? 4 : 1);
      if (CHARGED) {
        const TChg &c = S.sx[tid];
        const double *sbp = v.mat[FB_MAT_S] + v.boff[p];
        const double *dbp = v.mat[FB_MAT_DIP] + v.boff[p];
#pragma unroll
        for (int e = 0; e < 16; e++) {
          float val = 0.f;
          if (e < nb) {
            double sterm = sbp[e] * 0.5, dterm = dbp[e] / (double)c.z;
            double emnpl = (sterm - dterm) / (double)c.d13 + (sterm + dterm) / (double)c.d23;
            val = X[e] + (float)(c.dqs * (1.0 - c.stinky) * emnpl * FB_P_EQ2);
          }
          sC[tid][e] = val;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 16; e++) sC[tid][e] = X[e];
      }
    }
    tile_sync<CH>();
    // ---------------- phase D: ordered binary64 accumulation, one lane per
    // (owned segment, block element), segment order preserved: the sum starts
    // from zero in the segment's first chunk and continues from the block
    // (written by this lane in the previous chunk) in a later one
    for (int sl = tid; sl < nown * NB; sl += CH) {
      const int sg = sl / NB, e = sl - sg * NB;
      // the segment's triplets relative to the chunk start
      const int s0 = (int)(S.sSeg[sg] - c0), s1 = (int)(S.sSeg[sg + 1] - c0);
      const int k0 = max(s0, 0), k1 = min(s1, nt);
      if (k0 < k1 && e < S.sNb[sg]) {
        double *const o = vna + S.sBo[sg] + e;
        double a = (s0 < 0) ? *o : 0.0;
#pragma unroll 4
        for (int k = k0; k < k1; k++) a += (double)sC[k][e];
        *o = a;
      }
    }
    tile_sync<CH>();
  }
}

// table stencil: one tile per workgroup
template <int CH, bool CHARGED>
__global__ __launch_bounds__(CH, 4) void k_3c_mixed(DView v) {
  __shared__ S3c<CH, CHARGED, false> S;
  const int kt = CH == 64 ? 0 : (CH == 128 ? 1 : 2);
  int status = 0;
  tile3c<CH, CHARGED, false>(v, v.tiles[kt][xcd_tile(blockIdx.x, gridDim.x)], S, nullptr, status);
  if (status) atomicOr(v.status, status);
}

// Next tile of a persistent workgroup from per-XCD ranges of the 256-triplet
// tiles: queue[x * FB_QSTRIDE] counts the tiles taken from the range of XCD x (numbered
// contiguously per XCD, as xcd_tile numbers them); the workgroup's own XCD
// (b mod 8) first, then the others; -1 when all are empty.  One lane per
// taker (workgroup or wavefront).
__device__ __forceinline__ int next_tile(int *queue, unsigned ntiles) {
  const unsigned per = ntiles >> 3, rem = ntiles & 7u, x0 = blockIdx.x & 7u;
  for (unsigned k = 0; k < 8; k++) {
    const unsigned x = (x0 + k) & 7u;
    const unsigned lo = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    const unsigned cnt = per + (x < rem ? 1u : 0u);
    if (cnt == 0) continue;
    const unsigned i = (unsigned)atomicAdd(&queue[x * FB_QSTRIDE], 1);
    if (i < cnt) return (int)(lo + i);
  }
  return -1;
}

// learned surrogate with the stencil as fallback: persistent workgroups over
// the per-XCD queue of the 256-triplet tiles (next_tile); a workgroup exits
// when all ranges are empty.
__global__ __launch_bounds__(256, 4) void k_3c_mlp(DView v, int *queue) {
  __shared__ S3c<256, false, true> S;
  __shared__ int s_tile;
  __shared__ unsigned sCnt[2];
  const int tid = threadIdx.x;
  if (tid < 2) sCnt[tid] = 0u;
  const unsigned nt = (unsigned)v.ntiles[2];
  int status = 0;
  for (;;) {
    __syncthreads();
    if (tid == 0) s_tile = next_tile(queue, nt);
    __syncthreads();
    const int t = s_tile;
    if (t < 0) break;
    tile3c<256, false, true>(v, v.tiles[2][t], S, sCnt, status);
  }
  if (tid == 0) {
    atomicAdd(&v.mlp_count[0], (unsigned long long)sCnt[0]);
    atomicAdd(&v.mlp_count[1], (unsigned long long)sCnt[1]);
  }
  if (status) atomicOr(v.status, status);
}

// The whole neutral H+S build of s/p species in one launch: blocks
// [0, n2p) are the 2c workgroup items of v.wg2c (S, T, dipole_z, the Vna
// ontop and atom terms, fb_2c_sp.h; the Vna blocks of 3c targets are left to
// the 3c tiles), n2p = v.nwg2c rounded up to a multiple of 8 so that the XCD
// of a block stays blockIdx mod 8; each block after them runs four
// 64-triplet 3c tiles, one per wavefront, numbered contiguously per XCD (the
// tiles that win as separate launches, without a workgroup barrier between
// the waves of a block).  The 2c blocks are dispatched
// first, so their long per-lane chains (one pair per lane, six interactions)
// start at once and the 3c tiles fill the issue slots they leave idle.
// Neutral builds only: charged 3c terms read the S and dipole_z blocks of the
// 2c items.
union SFused {
  S2c<0> s2;
  S3c<64, false, false> s3[4];
};

// W: waves per SIMD the register allocation targets (3: the 2c body's 159
// VGPRs fit; 4: the 3c tiles' occupancy, the 2c body spills)
template <int W>
__global__ __launch_bounds__(256, W) void k_hs_fused(DView v, int n2p) {
  __shared__ SFused S;
  const int b = blockIdx.x;
  if (b < n2p) {
    if (b < v.nwg2c) wg2c_sp<0>(v, 15 | FB_2C_SKIP_TGT, b, S.s2);
    return;
  }
  const int b3 = b - n2p;
  // four 64-triplet tiles per workgroup, one per wavefront
  const int wv = threadIdx.x >> 6;
  const int64_t tile = 4 * xcd_tile(b3, gridDim.x - n2p) + wv;
  if (tile >= v.ntiles[0]) return;
  int status = 0;
  tile3c<64, false, false>(v, v.tiles[0][tile], S.s3[wv], nullptr, status);
  if (status) atomicOr(v.status, status);
}

// Persistent form of k_hs_fused: the same 2c blocks first, then a fixed
// number of blocks whose wavefronts each take 64-triplet tiles from the
// per-XCD queue (next_tile over v.tiles[0]) until every range is empty.
// Every wavefront's loop ends: each pass either takes an unused tile or
// finds all eight ranges exhausted.
template <int W>
__global__ __launch_bounds__(256, W) void k_hs_persist(DView v, int n2p, int *queue) {
  __shared__ SFused S;
  const int b = blockIdx.x;
  if (b < n2p) {
    if (b < v.nwg2c) wg2c_sp<0>(v, 15 | FB_2C_SKIP_TGT, b, S.s2);
    return;
  }
  const int wv = threadIdx.x >> 6;
  const unsigned nt = (unsigned)v.ntiles[0];
  int status = 0;
  for (;;) {
    int t = 0;
    if ((threadIdx.x & 63) == 0) t = next_tile(queue, nt);
    t = __shfl(t, 0);
    if (t < 0) break;
    tile3c<64, false, false>(v, v.tiles[0][t], S.s3[wv], nullptr, status);
    tile_sync<64>();   // the next tile reuses the wavefront's LDS rows
  }
  if (status) atomicOr(v.status, status);
}

template <bool CHARGED>
static int launch_mixed(const DView &v, int chunk, hipStream_t st) {
  const int kt = chunk == 64 ? 0 : (chunk == 256 ? 2 : 1);
  if (v.ntiles[kt] == 0) return 0;
  const dim3 grid((unsigned)v.ntiles[kt]);
  if (kt == 2) {
    hipLaunchKernelGGL((k_3c_mixed<256, CHARGED>), grid, dim3(256), 0, st, v);
  } else if (kt == 0) {
    hipLaunchKernelGGL((k_3c_mixed<64, CHARGED>), grid, dim3(64), 0, st, v);
  } else {
    hipLaunchKernelGGL((k_3c_mixed<128, CHARGED>), grid, dim3(128), 0, st, v);
  }
  return (int)hipGetLastError();
}

extern "C" int fbk_fill_i32(int *a, int64_t n, int val, hipStream_t st);

// queue: FB_QUEUE_INTS ints of device memory, cleared here by a fill kernel; nwg:
// persistent workgroups (four per CU)
extern "C" int fbk_3c_mlp(const DView &v, int *queue, int nwg, hipStream_t st) {
  if (v.ntrip == 0 || v.ntiles[2] == 0) return 0;
  int rc = fbk_fill_i32(queue, FB_QUEUE_INTS, 0, st);
  if (rc) return rc;
  const dim3 grid((unsigned)std::max(1, std::min(nwg, v.ntiles[2])));
  hipLaunchKernelGGL(k_3c_mlp, grid, dim3(256), 0, st, v, queue);
  return (int)hipGetLastError();
}

extern "C" int fbk_3c_mixed(const DView &v, int chunk, int charged, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  return charged ? launch_mixed<true>(v, chunk, st) : launch_mixed<false>(v, chunk, st);
}

// waves: 3 or 4 per SIMD (the register allocation of k_hs_fused); queue:
// null for one block per four tiles, else FB_QUEUE_INTS ints of device memory (cleared
// here by a fill kernel) for the persistent form with waves x ncu 3c blocks
extern "C" int fbk_hs_fused(const DView &v, int waves, int *queue, int ncu, hipStream_t st) {
  if (v.nwg2c == 0 && v.ntiles[0] == 0) return 0;
  const int n2p = (v.nwg2c + 7) / 8 * 8;
  if (queue) {
    int rc = fbk_fill_i32(queue, FB_QUEUE_INTS, 0, st);
    if (rc) return rc;
    const int nq = std::max(1, std::min(waves * ncu, (v.ntiles[0] + 3) / 4));
    const dim3 grid((unsigned)(n2p + nq));
    if (waves == 3)
      hipLaunchKernelGGL(k_hs_persist<3>, grid, dim3(256), 0, st, v, n2p, queue);
    else
      hipLaunchKernelGGL(k_hs_persist<4>, grid, dim3(256), 0, st, v, n2p, queue);
    return (int)hipGetLastError();
  }
  const dim3 grid((unsigned)(n2p + (v.ntiles[0] + 3) / 4));
  if (waves == 3)
    hipLaunchKernelGGL(k_hs_fused<3>, grid, dim3(256), 0, st, v, n2p);
  else
    hipLaunchKernelGGL(k_hs_fused<4>, grid, dim3(256), 0, st, v, n2p);
  return (int)hipGetLastError();
}
