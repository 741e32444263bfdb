// Three-centre neutral-atom (bcna) assembly, j.ASSEMBLERS/M_assemble_3c_DOGS.f90:86-340.
//
// Triplets arrive sorted by their target block (iatom, mneigh), stably, so a
// segment of consecutive triplets holds every contribution to one block in the
// reference's accumulation order (ialpha ascending, then the common-pair
// order).  A workgroup owns the segments that start inside its tile of
// consecutive triplets: it computes the contributions of their triplets in
// chunks, and one thread per segment adds them in order and stores the block
// once.  No atomics; results are deterministic.  A triplet contributes one
// term (neutral) or, with shell charges on its third atom, one more term per
// charged shell; the reference adds each term to the block on its own
// (M_assemble_3c_DOGS.f90:298-318), so the terms are kept apart and the
// segment owner adds them one by one in that order.
//
// k_3c_fp64 is the plain binary64 kernel: one thread per triplet evaluates the
// reference's interpolation (e.FDATA/M_Fdata_3c.f90:293-517) literally.
#include "fb_device.h"
#include "fb_3c_dev.h"

#define T3 128
#define NB (FB_MO * FB_MO)
#define NME3 10   // max non-zero 3c matrix elements for s/p species

__device__ __forceinline__ int group3c_id(const DView &v, int in1, int in2, int in3, int isub) {
  return v.idx3c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * v.sp.ns + (in3 - 1)) * FB_MAXSUB3 + isub];
}

// triplet geometry, M_assemble_3c_DOGS.f90:175-271
struct Geo3 {
  double x, z, cost, d13, d23, stinky;
  double eps[9];
  int in1, in2, indna, nmu, nnu;
};

__device__ inline void d_geo3(const DView &v, int4 tr, Geo3 &G) {
  const int ial = tr.x, gia = tr.y, gja = tr.z;
  const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
  const int s = v.atom_struct[ial];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  double r1[3], r2[3], rna[3], r21[3], sighat[3], rnabc[3], rhat[3];
  for (int k = 0; k < 3; k++) {
    r1[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
    r2[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
    rna[k] = v.ratom[3 * ial + k];
    r21[k] = r2[k] - r1[k];
  }
  G.in1 = v.imass[gia];
  G.in2 = v.imass[gja];
  G.indna = v.imass[ial];
  G.nmu = v.sp.norb[G.in1 - 1];
  G.nnu = v.sp.norb[G.in2 - 1];
  G.z = d_distance(r1, r2);
  if (G.z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) sighat[k] = r21[k] / G.z;
  for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1[k] + r21[k] / 2.0);
  G.x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
  G.d13 = d_distance(rna, r1);
  G.d23 = d_distance(rna, r2);
  if (G.x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
  else for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] / G.x;
  G.cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
  d_epsilon(rhat, sighat, G.eps);
  double rc1 = v.sp.rcmin[G.in1 - 1], rc2 = v.sp.rcmin[G.in2 - 1], rc3 = v.sp.rcmin[G.indna - 1];
  G.stinky = d_smoother(G.d13, rc1 + rc3, 0.8);
  G.stinky = G.stinky * d_smoother(G.d23, rc2 + rc3, 0.8);
}

// terms per triplet: the neutral one plus one per charged shell of the third atom
#define KT_GEN (1 + FB_MAXSH)
#define T3G 64   // triplets per tile of the generic kernel (its LDS holds KT_GEN terms each)

__global__ __launch_bounds__(T3G) void k_3c_fp64(DView v) {
  __shared__ double s_c[T3G][KT_GEN][NB + 1];
  __shared__ int s_n[T3G];
  const int64_t ts = xcd_tile(blockIdx.x, gridDim.x) * T3G;
  const int64_t te = min(ts + T3G, v.ntrip);
  int64_t first, last, r0, r1;
  if (!tile_segments(v, ts, te, first, last, r0, r1)) return;
  const int64_t my = first + threadIdx.x;
  const bool own = my <= last;
  int64_t a = 0, b = 0;
  if (own) { a = v.seg_off[my]; b = v.seg_off[my + 1]; }
  double acc[NB];
  for (int i = 0; i < NB; i++) acc[i] = 0.0;
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += T3G) {
    const int64_t t = c0 + threadIdx.x;
    int nt = 0;
    if (t < r1) {
      const int4 tr = v.trip[t];
      Geo3 G;
      d_geo3(v, tr, G);
      double pm[9], hm[NB], bx[NB];
      d_pmat<double>(G.eps, pm);
      const int nb = G.nmu * G.nnu;
      int gid = group3c_id(v, G.in1, G.in2, G.indna, 0);
      status |= d_getMEs_3c<NME3>(v, gid, G.x, G.z, G.cost, G.nmu, G.nnu, hm);
      d_rotate<double>(v.sp, G.in1, G.in2, pm, hm, G.nmu, bx);
      for (int i = 0; i < nb; i++) s_c[threadIdx.x][0][i] = bx[i] * FB_P_EQ2;
      nt = 1;
      bool charged = false;
      for (int isorp = 1; isorp <= v.sp.nssh[G.indna - 1]; isorp++)
        charged |= (v.dQ[tr.x * FB_MAXSH + isorp - 1] != 0.0);
      if (charged) {
        const int p = v.trip_pair[t];
        const double *sb = v.mat[FB_MAT_S] + v.boff[p];
        const double *db = v.mat[FB_MAT_DIP] + v.boff[p];
        for (int isorp = 1; isorp <= v.sp.nssh[G.indna - 1]; isorp++) {
          double dQ = v.dQ[tr.x * FB_MAXSH + isorp - 1];
          if (dQ == 0.0) continue;   // the reference adds an exact zero here
          gid = group3c_id(v, G.in1, G.in2, G.indna, isorp);
          status |= d_getMEs_3c<NME3>(v, gid, G.x, G.z, G.cost, G.nmu, G.nnu, hm);
          d_rotate<double>(v.sp, G.in1, G.in2, pm, hm, G.nmu, bx);
          for (int i = 0; i < nb; i++) {
            double sterm = sb[i] / 2.0, dterm = db[i] / G.z;
            double emnpl = (sterm - dterm) / G.d13 + (sterm + dterm) / G.d23;
            s_c[threadIdx.x][nt][i] = dQ * (G.stinky * bx[i] + (1.0 - G.stinky) * emnpl) * FB_P_EQ2;
          }
          nt++;
        }
      }
    }
    s_n[threadIdx.x] = nt;
    __syncthreads();
    if (own) {
      int64_t k0 = max(a, c0), k1 = min(b, c0 + T3G);
      for (int64_t k = k0; k < k1; k++)
        for (int j = 0; j < s_n[k - c0]; j++)
          for (int i = 0; i < NB; i++) acc[i] = acc[i] + s_c[k - c0][j][i];
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
    for (int i = 0; i < nb; i++) dst[i] = acc[i];
  }
  if (status) atomicOr(v.status, status);
}

extern "C" int fbk_3c_fp64(const DView &v, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  hipLaunchKernelGGL(k_3c_fp64, dim3((unsigned)((v.ntrip + T3G - 1) / T3G)), dim3(T3G), 0, st, v);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// s/p species with canonical tables: the same algorithm, with compile-time
// matrix-element counts and the register-resident rotation of fb_sp.h (same
// operations and order as the generic path, without private-memory arrays).
#include "fb_sp.h"

// terms per triplet of s/p species: neutral, plus at most two charged shells
#define KT_SP 3

// the triplet's terms into c[k][*], k < returned count (neutral first, then
// the charged shells in order)
template <int CLS>
__device__ __forceinline__ int triplet_fp64_sp(const DView &v, int64_t t, const int4 tr, const Geo3 &G,
                                               double (*c)[NB + 1], int *nterms, int *status_out) {
  constexpr int NME = nme3c_of(CLS);
  constexpr int NBL = ((CLS & 2) ? 4 : 1) * ((CLS & 1) ? 4 : 1);
  double pm[9], F[NME], bx[16];
  d_pmat<double>(G.eps, pm);
  int status = getMEs_3c_sp<NME>(v, group3c_id(v, G.in1, G.in2, G.indna, 0), G.x, G.z, G.cost, F);
  rot3c_sp<double>(CLS, pm, F, bx);
#pragma unroll
  for (int i = 0; i < NBL; i++) c[0][i] = bx[i] * FB_P_EQ2;
  int nt = 1;
  bool charged = false;
  for (int isorp = 1; isorp <= v.sp.nssh[G.indna - 1]; isorp++) charged |= (v.dQ[tr.x * FB_MAXSH + isorp - 1] != 0.0);
  if (charged) {
    const int p = v.trip_pair[t];
    const double *sb = v.mat[FB_MAT_S] + v.boff[p];
    const double *db = v.mat[FB_MAT_DIP] + v.boff[p];
    for (int isorp = 1; isorp <= v.sp.nssh[G.indna - 1] && nt < KT_SP; isorp++) {
      double dQ = v.dQ[tr.x * FB_MAXSH + isorp - 1];
      if (dQ == 0.0) continue;   // the reference adds an exact zero here
      status |= getMEs_3c_sp<NME>(v, group3c_id(v, G.in1, G.in2, G.indna, isorp), G.x, G.z, G.cost, F);
      rot3c_sp<double>(CLS, pm, F, bx);
#pragma unroll
      for (int i = 0; i < NBL; i++) {
        double sterm = sb[i] / 2.0, dterm = db[i] / G.z;
        double emnpl = (sterm - dterm) / G.d13 + (sterm + dterm) / G.d23;
        c[nt][i] = dQ * (G.stinky * bx[i] + (1.0 - G.stinky) * emnpl) * FB_P_EQ2;
      }
      nt++;
    }
  }
  *nterms = nt;
  *status_out |= status;
  return 0;
}

__global__ __launch_bounds__(T3) void k_3c_fp64_sp(DView v) {
  __shared__ double s_c[T3][KT_SP][NB + 1];
  __shared__ unsigned char s_n[T3];
  const int64_t ts = xcd_tile(blockIdx.x, gridDim.x) * T3;
  const int64_t te = min(ts + T3, v.ntrip);
  int64_t first, last, r0, r1;
  if (!tile_segments(v, ts, te, first, last, r0, r1)) return;
  const int64_t my = first + threadIdx.x;
  const bool own = my <= last;
  int64_t a = 0, b = 0;
  if (own) { a = v.seg_off[my]; b = v.seg_off[my + 1]; }
  double acc[NB];
#pragma unroll
  for (int i = 0; i < NB; i++) acc[i] = 0.0;
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += T3) {
    const int64_t t = c0 + threadIdx.x;
    int nt = 0;
    if (t < r1) {
      const int4 tr = v.trip[t];
      Geo3 G;
      d_geo3(v, tr, G);
      const int cls = (G.nmu == 4 ? 2 : 0) + (G.nnu == 4 ? 1 : 0);
      switch (cls) {
        case 3: triplet_fp64_sp<3>(v, t, tr, G, s_c[threadIdx.x], &nt, &status); break;
        case 2: triplet_fp64_sp<2>(v, t, tr, G, s_c[threadIdx.x], &nt, &status); break;
        case 1: triplet_fp64_sp<1>(v, t, tr, G, s_c[threadIdx.x], &nt, &status); break;
        default: triplet_fp64_sp<0>(v, t, tr, G, s_c[threadIdx.x], &nt, &status); break;
      }
    }
    s_n[threadIdx.x] = (unsigned char)nt;
    __syncthreads();
    if (own) {
      int64_t k0 = max(a, c0), k1 = min(b, c0 + T3);
      for (int64_t k = k0; k < k1; k++) {
        const int n = s_n[k - c0];
        for (int j = 0; j < n; j++)
#pragma unroll
          for (int i = 0; i < NB; i++) acc[i] = acc[i] + s_c[k - c0][j][i];
      }
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

extern "C" int fbk_3c_fp64_sp(const DView &v, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  hipLaunchKernelGGL(k_3c_fp64_sp, dim3((unsigned)((v.ntrip + T3 - 1) / T3)), dim3(T3), 0, st, v);
  return (int)hipGetLastError();
}
