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
//      grid cell (ix, iy) = int(x/dx)+1, int(y/dy)+1 and the method choice
//      (gradient of ME 1 of theta cell 1 of each table against 1e-6, from the
//      binary64 table).  Unit vectors and the frame epsilon(rhat, sighat) are
//      binary64 with reciprocal multiplies; weights, Legendre P0..P4 and the
//      p rotation go to LDS in FP32.
//   B  one lane per (triplet, matrix element): the 16 (or 3) stencil points of
//      the FP32 slab [iy][ix][ME][theta], one float4 + one float per point,
//      Legendre-folded at each point, then the x/y Lagrange weights.
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

struct TPrm {
  unsigned base;              // FP32 slab index (floats) of stencil corner (ix-1, iy-1), neutral group
  unsigned rowstride;         // floats per iy row of the slab
  unsigned char cls, nme, m1, pad_;   // pair class, ME count, method-1 flag
  float px, py;               // position in the cell; phase B forms the weights
  float P[4], sint;           // Legendre P1..P4 (P0 = 1)
  float pm[9];
};
// 19 words: an odd row stride, so lane-per-triplet accesses hit distinct banks
static_assert(sizeof(TPrm) == 19 * 4, "TPrm row stride must stay odd in words");

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

// Legendre fold of one (point, ME): sum_theta P_theta T_theta; off in bytes
__device__ __forceinline__ float fold(__amdgpu_buffer_rsrc_t r, unsigned off, const float *P) {
  const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
  const float b = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(off + 16u), 0, 0));
  float v = P[0] * a.x;
  v = __builtin_fmaf(P[1], a.y, v);
  v = __builtin_fmaf(P[2], a.z, v);
  v = __builtin_fmaf(P[3], a.w, v);
  return __builtin_fmaf(P[4], b, v);
}

// one (triplet, ME): 16-point tensor stencil (method 3) or 3-point rule (method 1);
// base and rowstride in floats
template <int NME>
__device__ __forceinline__ float stencil(__amdgpu_buffer_rsrc_t r, unsigned base, unsigned rowstride, int m,
                                         const float *wx, const float *wy, const float *w1, const float *P, int m1) {
  constexpr unsigned PT = NME * FB_TH32 * 4;   // bytes per ix step
  const unsigned o = (base + (unsigned)m * FB_TH32) * 4u, rb = rowstride * 4u;
  if (m1) {
    const unsigned c = o + rb + PT;   // (ix, iy)
    float v = w1[0] * fold(r, c, P);
    v = __builtin_fmaf(w1[1], fold(r, c + PT, P), v);
    return __builtin_fmaf(w1[2], fold(r, c + rb, P), v);
  }
  float acc = 0.f;
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const unsigned row = o + (unsigned)j * rb;
    float g = wx[0] * fold(r, row, P);
    g = __builtin_fmaf(wx[1], fold(r, row + PT, P), g);
    g = __builtin_fmaf(wx[2], fold(r, row + 2 * PT, P), g);
    g = __builtin_fmaf(wx[3], fold(r, row + 3 * PT, P), g);
    acc = __builtin_fmaf(wy[j], g, acc);
  }
  return acc;
}

// LDS of one tile pass
template <int CH, bool CHARGED, bool MLP>
struct S3c {
  TPrm sq[CH];
  float sF[CH][NME3];
  TChg sx[CHARGED ? CH : 1];
  int sSeg[CH + 1];                   // triplet offsets (below 2^31: the list build checks)
  int64_t sBo[CH];
  unsigned char sNb[CH];
  float sU[MLP ? CH : 1][4];          // network inputs and envelope
  int sG[MLP ? CH : 1];               // network group of the triplet, -1: table stencil
  short sNet[MLP ? CH : 1], sFb[MLP ? CH : 1];   // chunk triplets on the network / on the stencil, in order
  int sWn[MLP ? CH / 64 : 1][2];      // per-wavefront counts of both lists
};

// One tile of whole target segments (first, last, r0, r1), in chunks of at
// most CH triplets.  MLP: the learned surrogate replaces the stencil of phase
// B for every triplet whose calibrated bound passes (neutral builds only).
template <int CH, bool CHARGED, bool MLP>
__device__ __forceinline__ void tile3c(const DView &v, const int4 td, S3c<CH, CHARGED, MLP> &S, unsigned *sCnt,
                                       int &status) {
  // contributions reuse the parameter storage, rows padded to 17 words
  float(*sC)[NB + 1] = reinterpret_cast<float(*)[NB + 1]>(&S.sq[0]);
  static_assert(sizeof(TPrm) >= (NB + 1) * sizeof(float), "contribution row must fit in a parameter row");
  const int tid = threadIdx.x;
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
  const __amdgpu_buffer_rsrc_t tab = slab_rsrc(v.tab3c32, v.ntab3c32);

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
      const int in1 = v.imass[gia], in2 = v.imass[gja], indna = v.imass[ial];
      q.cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
      const double z = d_distance(r1v, r2v);
      if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
      else {
        const double iz = 1.0 / z;
#pragma unroll
        for (int k = 0; k < 3; k++) sighat[k] = r21[k] * iz;
      }
#pragma unroll
      for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1v[k] + r21[k] * 0.5);
      const double x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
      if (x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
      else {
        const double ix_ = 1.0 / x;
#pragma unroll
        for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] * ix_;
      }
      const double cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
      {
        // epsilon(rhat, sighat), b.FUNCTIONS/epsilon.f90:120-177, with the
        // reference's thresholds (sighat is a unit vector, so the r2 = 0 branch
        // cannot occur)
        // sighat and rhat are unit vectors already (rhat = 0 when x < 1e-5),
        // so the normalisations of epsilon.f90:136-145 change them by rounding only
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
        const double eps[9] = {xp[0], xp[1], xp[2], yp[0], yp[1], yp[2], zp[0], zp[1], zp[2]};
        float pmf[9];
        d_pmat<float>(eps, pmf);
#pragma unroll
        for (int k = 0; k < 9; k++) q.pm[k] = pmf[k];
      }
      const double cost2 = cost * cost;
      q.sint = (float)sqrt(fmax(1.0e-5, 1.0 - cost2));
      q.P[0] = (float)cost;
      q.P[1] = (float)(0.5 * (3.0 * cost2 - 1.0));
      q.P[2] = (float)(0.5 * cost * (5.0 * cost2 - 3.0));
      q.P[3] = (float)(0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2)));
      int gid = aux & 0xffff;
      if (gid == 0xffff) gid = -1;
      if (gid < 0) {
        status |= ST_NO_TABLE;
        gid = 0;
      }
      const DGroup3c g = v.grp3c[gid];
      q.nme = g.nME;
      // grid cell and method, binary64 as M_Fdata_3c.f90:367-385
      const double dx = g.xmax / (g.nx - 1.0), dy = g.ymax / (g.ny - 1.0);
      const double qx = x / dx, qy = z / dy;
      int ix = (int)qx + 1, iy = (int)qy + 1;
      if (ix < 2) ix = 2;
      if (iy < 2) iy = 2;
      if (ix > g.nx - 2) ix = g.nx - 2;
      if (iy > g.ny - 2) iy = g.ny - 2;
      const double px = qx - ix + 1.0, py = qy - iy + 1.0;
      // method of the cell, precomputed from the binary64 table with the
      // reference's gradient test (M_Fdata_3c.f90:383-385)
      const int64_t icell = (int64_t)(iy - 1) * g.nx + (ix - 1);
      q.m1 = v.meth3c[g.moff + icell];
      q.px = (float)px;
      q.py = (float)py;
      q.rowstride = (unsigned)(g.nx * g.nME * FB_TH32);
      q.base = (unsigned)(g.off32 + ((int64_t)(iy - 2) * g.nx + (ix - 2)) * g.nME * FB_TH32);
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
        S.sU[tid][2] = (float)cost;
        S.sU[tid][3] = ok ? mlp_envelope(w, (float)x, (float)z) : 0.f;
        S.sG[tid] = ok ? gid : -1;
        atomicAdd(&sCnt[0], ok ? 1u : 0u);
        atomicAdd(&sCnt[1], 1u);
      }
      if (CHARGED) {
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
          c.gk[isorp - 1] = (int)(gq.off32 - g.off32);
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
    // the chunk's triplets split into a network list and a stencil list, each
    // dense and in triplet order, so that each phase costs in proportion to
    // its share
    int nnet = 0, nfb = nt;
    if (MLP) {
      const int wv = tid >> 6, l = tid & 63;
      const bool on = tid < nt && S.sG[tid] >= 0, fb = tid < nt && !on;
      const unsigned long long bn = __ballot(on), bf = __ballot(fb);
      if (l == 0) {
        S.sWn[wv][0] = __popcll(bn);
        S.sWn[wv][1] = __popcll(bf);
      }
      __syncthreads();
      int pn = 0, pf = 0;
      nnet = 0;
      nfb = 0;
#pragma unroll
      for (int w = 0; w < CH / 64; w++) {
        if (w < wv) {
          pn += S.sWn[w][0];
          pf += S.sWn[w][1];
        }
        nnet += S.sWn[w][0];
        nfb += S.sWn[w][1];
      }
      const unsigned long long below = (l == 0) ? 0ull : (~0ull >> (64 - l));
      if (on) S.sNet[pn + __popcll(bn & below)] = (short)tid;
      if (fb) S.sFb[pf + __popcll(bf & below)] = (short)tid;
    }
    __syncthreads();
    // ---------------- phase B: interpolation, one lane per (triplet, ME)
    for (int it = tid; it < nfb * NME3; it += CH) {
      const int k = it / NME3, m = it - k * NME3;
      const int t = MLP ? (int)S.sFb[k] : k;
      const TPrm &q = S.sq[t];
      if (m >= q.nme) continue;
      float wx[4], wy[4], w1[3], P[5];
      const float px = q.px, py = q.py;
      cubic_w(px, wx);
      cubic_w(py, wy);
      w1[0] = (1.f - px) - py;
      w1[1] = px;
      w1[2] = py;
#pragma unroll
      for (int k = 0; k < 4; k++) P[k + 1] = q.P[k];
      P[0] = 1.f;
      const unsigned base = q.base, rs = q.rowstride;
      float F;
      switch (q.nme) {
        case 10: F = stencil<10>(tab, base, rs, m, wx, wy, w1, P, q.m1); break;
        case 3: F = stencil<3>(tab, base, rs, m, wx, wy, w1, P, q.m1); break;
        default: F = stencil<1>(tab, base, rs, m, wx, wy, w1, P, q.m1); break;
      }
      if (CHARGED) {
        const TChg &c = S.sx[t];
        for (int k = 0; k < FB_MAXSH; k++) {
          const float dq = c.dq[k];
          if (dq == 0.f) continue;
          const unsigned bk = base + (unsigned)c.gk[k];
          float Fk;
          switch (q.nme) {
            case 10: Fk = stencil<10>(tab, bk, rs, m, wx, wy, w1, P, c.m1k[k]); break;
            case 3: Fk = stencil<3>(tab, bk, rs, m, wx, wy, w1, P, c.m1k[k]); break;
            default: Fk = stencil<1>(tab, bk, rs, m, wx, wy, w1, P, c.m1k[k]); break;
          }
          F = __builtin_fmaf(dq * c.stinky, Fk, F);
        }
      }
      S.sF[t][m] = F;
    }
    if (MLP) {
      // keeps the weight loads of phase M from being hoisted into the stencil
      // loop (their registers would overlap)
      __syncthreads();
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
              if (m < q.nme) {
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
    __syncthreads();
    // ---------------- phase C: rotation, one lane per triplet
    float X[16];
    int cls = 0, p = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) X[i] = 0.f;
    if (tid < nt) {
      const TPrm &q = S.sq[tid];
      float F[NME3], pm[9];
      cls = q.cls;
      if (CHARGED) p = v.trip_pair[c0 + tid];
      const int nm0 = (cls == 3) ? 6 : (cls == 0 ? 1 : 2);
      const float sint = q.sint;
#pragma unroll
      for (int m = 0; m < NME3; m++) F[m] = (m < q.nme) ? S.sF[tid][m] * (m >= nm0 ? sint : 1.f) : 0.f;
#pragma unroll
      for (int k = 0; k < 9; k++) pm[k] = q.pm[k];
      switch (cls) {
        case 3: rot3c_sp<float>(3, pm, F, X); break;
        case 2: rot3c_sp<float>(2, pm, F, X); break;
        case 1: rot3c_sp<float>(1, pm, F, X); break;
        default: rot3c_sp<float>(0, pm, F, X); break;
      }
    }
    __syncthreads();   // every lane has read its parameters before sC overwrites them
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
            val = X[e] * (float)FB_P_EQ2 + (float)(c.dqs * (1.0 - c.stinky) * emnpl * FB_P_EQ2);
          }
          sC[tid][e] = val;
        }
      } else {
#pragma unroll
        for (int e = 0; e < 16; e++) sC[tid][e] = X[e] * (float)FB_P_EQ2;
      }
    }
    __syncthreads();
    // ---------------- phase D: ordered binary64 accumulation, one lane per
    // (owned segment, block element), segment order preserved: the sum starts
    // from zero in the segment's first chunk and continues from the block
    // (written by this lane in the previous chunk) in a later one
#pragma unroll
    for (int j = 0; j < NB; j++) {
      const int sl = tid + j * CH;
      if (sl < nown * NB) {
        const int sg = sl / NB, e = sl - sg * NB;
        const int64_t s0 = S.sSeg[sg];
        const int64_t k0 = max(s0, c0), k1 = min((int64_t)S.sSeg[sg + 1], c0 + nt);
        if (k0 < k1 && e < S.sNb[sg]) {
          double *const o = vna + S.sBo[sg] + e;
          double a = (s0 < c0) ? *o : 0.0;
          for (int64_t k = k0; k < k1; k++) a += (double)sC[k - c0][e];
          *o = a;
        }
      }
    }
    __syncthreads();
  }
}

// table stencil: one tile per workgroup
template <int CH, bool CHARGED>
__global__ __launch_bounds__(CH) void k_3c_mixed(DView v) {
  __shared__ S3c<CH, CHARGED, false> S;
  const int kt = CH == 64 ? 0 : (CH == 128 ? 1 : 2);
  int status = 0;
  tile3c<CH, CHARGED, false>(v, v.tiles[kt][xcd_tile(blockIdx.x, gridDim.x)], S, nullptr, status);
  if (status) atomicOr(v.status, status);
}

// Next tile of a persistent workgroup from per-XCD ranges of the 256-triplet
// tiles: queue[x] counts the tiles taken from the range of XCD x (numbered
// contiguously per XCD, as xcd_tile numbers them); the workgroup's own XCD
// (b mod 8) first, then the others; -1 when all are empty.  Thread 0 only.
__device__ __forceinline__ int next_tile(int *queue, unsigned ntiles) {
  const unsigned per = ntiles >> 3, rem = ntiles & 7u, x0 = blockIdx.x & 7u;
  for (unsigned k = 0; k < 8; k++) {
    const unsigned x = (x0 + k) & 7u;
    const unsigned lo = x < rem ? x * (per + 1) : rem * (per + 1) + (x - rem) * per;
    const unsigned cnt = per + (x < rem ? 1u : 0u);
    if (cnt == 0) continue;
    const unsigned i = (unsigned)atomicAdd(&queue[x], 1);
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
// of a block stays blockIdx mod 8; the blocks after them are the 256-triplet
// 3c tiles, numbered contiguously per XCD.  The 2c blocks are dispatched
// first, so their long per-lane chains (one pair per lane, six interactions)
// start at once and the 3c tiles fill the issue slots they leave idle.
// Neutral builds only: charged 3c terms read the S and dipole_z blocks of the
// 2c items.
union SFused {
  S2c<0> s2;
  S3c<256, false, false> s3;
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
  int status = 0;
  tile3c<256, false, false>(v, v.tiles[2][xcd_tile(b - n2p, gridDim.x - n2p)], S.s3, nullptr, status);
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

// queue: 8 ints of device memory, cleared here by a fill kernel; nwg:
// persistent workgroups (four per CU)
extern "C" int fbk_3c_mlp(const DView &v, int *queue, int nwg, hipStream_t st) {
  if (v.ntrip == 0 || v.ntiles[2] == 0) return 0;
  int rc = fbk_fill_i32(queue, 8, 0, st);
  if (rc) return rc;
  const dim3 grid((unsigned)std::max(1, std::min(nwg, v.ntiles[2])));
  hipLaunchKernelGGL(k_3c_mlp, grid, dim3(256), 0, st, v, queue);
  return (int)hipGetLastError();
}

extern "C" int fbk_3c_mixed(const DView &v, int chunk, int charged, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  return charged ? launch_mixed<true>(v, chunk, st) : launch_mixed<false>(v, chunk, st);
}

// waves: 3 or 4 per SIMD (the register allocation of k_hs_fused)
extern "C" int fbk_hs_fused(const DView &v, int waves, hipStream_t st) {
  if (v.nwg2c == 0 && v.ntiles[2] == 0) return 0;
  const int n2p = (v.nwg2c + 7) / 8 * 8;
  const dim3 grid((unsigned)(n2p + v.ntiles[2]));
  if (waves == 3)
    hipLaunchKernelGGL(k_hs_fused<3>, grid, dim3(256), 0, st, v, n2p);
  else
    hipLaunchKernelGGL(k_hs_fused<4>, grid, dim3(256), 0, st, v, n2p);
  return (int)hipGetLastError();
}
