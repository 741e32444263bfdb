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

// aux: the triplet's trip_aux word (neutral group | structure << 16)
__device__ inline void d_geo3(const DView &v, int4 tr, int aux, Geo3 &G) {
  const int ial = tr.x, gia = tr.y, gja = tr.z;
  const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
  const int s = aux >> 16;
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

// the generic kernel's sizes: s/p species (MO 4), or with d shells (MO 9)
template <int MO> struct Gen3c;
template <> struct Gen3c<FB_MO> {
  static constexpr int T = 64;     // triplets per tile (the LDS holds KT_GEN terms of each)
  static constexpr int NME = 10;   // non-zero bcna matrix elements
};
template <> struct Gen3c<FB_MOD> {
  static constexpr int T = 16;
  static constexpr int NME = FB_NME3_WIDE;
};

template <int MO>
__global__ __launch_bounds__(Gen3c<MO>::T) void k_3c_fp64(DView v) {
  constexpr int T3G = Gen3c<MO>::T, NBG = MO * MO;
  __shared__ double s_c[T3G][KT_GEN][NBG + 1];
  __shared__ int s_n[T3G];
  const int64_t ts = xcd_tile(blockIdx.x, gridDim.x) * T3G;
  const int64_t te = min(ts + T3G, v.ntrip);
  int64_t first, last, r0, r1;
  if (!tile_segments(v, ts, te, first, last, r0, r1)) return;
  const int64_t my = first + threadIdx.x;
  const bool own = my <= last;
  int64_t a = 0, b = 0;
  if (own) { a = v.seg_off[my]; b = v.seg_off[my + 1]; }
  double acc[NBG];
  for (int i = 0; i < NBG; i++) acc[i] = 0.0;
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += T3G) {
    const int64_t t = c0 + threadIdx.x;
    int nt = 0;
    if (t < r1) {
      const int4 tr = v.trip[t];
      Geo3 G;
      d_geo3(v, tr, v.trip_aux[t], G);
      double pm[9], dm[MO == FB_MOD ? 25 : 1], hm[NBG], bx[NBG];
      d_pmat<double>(G.eps, pm);
      if (MO == FB_MOD) d_dmat(G.eps, dm);
      const int nb = G.nmu * G.nnu;
      int gid = group3c_id(v, G.in1, G.in2, G.indna, 0);
      status |= d_getMEs_3c<Gen3c<MO>::NME>(v, gid, G.x, G.z, G.cost, G.nmu, G.nnu, hm);
      d_rotate<double>(v.sp, G.in1, G.in2, pm, hm, G.nmu, bx, dm);
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
          status |= d_getMEs_3c<Gen3c<MO>::NME>(v, gid, G.x, G.z, G.cost, G.nmu, G.nnu, hm);
          d_rotate<double>(v.sp, G.in1, G.in2, pm, hm, G.nmu, bx, dm);
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
          for (int i = 0; i < NBG; i++) acc[i] = acc[i] + s_c[k - c0][j][i];
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
  if (fb_wide_species(v.sp)) {
    constexpr int T = Gen3c<FB_MOD>::T;
    hipLaunchKernelGGL(k_3c_fp64<FB_MOD>, dim3((unsigned)((v.ntrip + T - 1) / T)), dim3(T), 0, st, v);
  } else {
    constexpr int T = Gen3c<FB_MO>::T;
    hipLaunchKernelGGL(k_3c_fp64<FB_MO>, dim3((unsigned)((v.ntrip + T - 1) / T)), dim3(T), 0, st, v);
  }
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// s/p species with canonical tables: the same algorithm, with compile-time
// matrix-element counts and the register-resident rotation of fb_sp.h (same
// operations and order as the generic path, without private-memory arrays).
#include "fb_sp.h"

// terms per triplet of s/p species: neutral, plus at most two charged shells
#define KT_SP 3

// ---------------------------------------------------------------------------
// The binary64 s/p kernel in four phases over tiles of whole target segments
// (the 64-triplet tiles of the mixed kernel, one wavefront each):
//   A  one lane per triplet: geometry (d_geo3), the cell and px, py as
//      getMEs_3c_sp forms them, the method of each term's table (the
//      reference's gradient test, precomputed per cell from the same binary64
//      table), Legendre values and the p rotation;
//   B  one lane per (triplet, term, matrix element): the theta sum of the
//      interpolants with the reference's expressions and order;
//   C  one lane per (triplet, term): rotation and the term's factors;
//   D  one lane per (segment, block element): the terms added in triplet
//      order, each term on its own.
// Every binary64 operation and its order is that of the one-lane-per-triplet
// form (getMEs_3c_sp, rot3c_sp, the terms added one by one), so the blocks are
// the reference's bits; the phases only change which lane does what.
#define CH64 64

struct TPrm64 {
  double px, py, sint, p[4];   // p1..p4 (p0 = 1)
  double pm[6];                // columns 0 and 1 of the p rotation; column 2 is their cross product
  int64_t base;                // FP64 slab index of (theta 0, ix, iy, ME 0), neutral group
  int rowstride, tstride;      // slab values per iy row and per theta cell
  unsigned char cls, nterms, m1[KT_SP];
  int gk[KT_SP];               // slab offset of each term's group relative to the neutral one
};

struct TChg64 {                // charged terms: M_assemble_3c_DOGS.f90:298-318
  double dq[KT_SP], z, d13, d23, stinky;
  int pair;
};

template <bool CHARGED>
struct S3c64 {
  TPrm64 sq[CH64];
  double sF[CH64][CHARGED ? KT_SP : 1][NME3];
  TChg64 sx[CHARGED ? CH64 : 1];   // after sF: the contribution rows cover sq, sF and sx
  int sSeg[CH64 + 1];
  int64_t sBo[CH64];
  unsigned char sNb[CH64];
  unsigned char sOrd[CH64];    // the chunk's triplets, 3-point cells first
};

// interim of one (theta, ME) at the cell of M_Fdata_3c.f90:393-489, T(a, b) =
// the value at (ix + a, iy + b)
template <int NME>
__device__ __forceinline__ double interim3c(const double *tab, int64_t o, int rs, double px, double py, bool m1) {
#define TV(a, b) tab[o + (int64_t)(b) * rs + (a) * NME]
  if (m1) return ((1.0 - px - py) * TV(0, 0) + px * TV(1, 0)) + py * TV(0, 1);
  double gk[4];
#pragma unroll
  for (int k = -1; k <= 2; k++) {
    double f1m1 = TV(k, -1);
    double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
    double f0p3 = 3.0 * TV(k, 0), f0p6 = 2.0 * f0p3;
    double f1p3 = 3.0 * TV(k, 1), f1p6 = 2.0 * f1p3;
    double f2p1 = TV(k, 2);
    double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
    double bb2 = f1m3 - f0p6 + f1p3;
    double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
    double bb0 = f0p6;
    gk[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
  }
#undef TV
  double f1m1 = gk[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
  double f0p3 = 3.0 * gk[1], f0p6 = 2.0 * f0p3;
  double f1p3 = 3.0 * gk[2], f1p6 = 2.0 * f1p3;
  double f2p1 = gk[3];
  double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
  double bb2 = f1m3 - f0p6 + f1p3;
  double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
  double bb0 = f0p6;
  return (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
}

// one matrix element m of one term: sum over theta of p_theta interim, then
// sin(theta) on the mvalue = 1 elements (getMEs_3c_sp)
template <int NME, int NM0 = (NME == 10) ? 6 : (NME == 3 ? 2 : 1)>
__device__ __forceinline__ double me3c(const DView &v, const TPrm64 &q, int term, int m) {
  const double *tab = v.tab3c;
  const int64_t o = q.base + q.gk[term] + m;
  const double p[5] = {1.0, q.p[0], q.p[1], q.p[2], q.p[3]};
  double F = 0.0;
  for (int th = 0; th < FB_NTHETA; th++)
    F = F + p[th] * interim3c<NME>(tab, o + (int64_t)th * q.tstride, q.rowstride, q.px, q.py, q.m1[term]);
  return m >= NM0 ? F * q.sint : F;
}

template <bool CHARGED>
__global__ __launch_bounds__(CH64) void k_3c_fp64_ph(DView v) {
  __shared__ S3c64<CHARGED> S;
  constexpr int KT = CHARGED ? KT_SP : 1;
  // contributions reuse the parameter and ME storage once phase C has read it
  double(*sC)[KT][NB + 1] = reinterpret_cast<double(*)[KT][NB + 1]>(&S.sq[0]);
  typedef S3c64<CHARGED> S_t;
  static_assert(offsetof(S_t, sF) == sizeof(S.sq) && offsetof(S_t, sx) == sizeof(S.sq) + sizeof(S.sF) &&
                    sizeof(S.sq) + sizeof(S.sF) + sizeof(S.sx) >= sizeof(double) * CH64 * KT * (NB + 1),
                "contribution rows must fit in the parameter, ME and charge rows");
  const int tid = threadIdx.x;
  const int4 td = v.tiles[0][xcd_tile(blockIdx.x, gridDim.x)];
  const int64_t first = td.x, last = td.y, r0 = td.z, r1 = td.w;
  if (first > last) return;
  const int nown = (int)(last - first + 1);
  for (int k = tid; k <= nown; k += CH64) S.sSeg[k] = (int)v.seg_off[first + k];
  for (int k = tid; k < nown; k += CH64) {
    const int p = v.seg_pair[first + k];
    const int64_t b0 = v.boff[p];
    S.sBo[k] = b0;
    S.sNb[k] = (unsigned char)(v.boff[p + 1] - b0);
  }
  double *const vna = v.mat[FB_MAT_VNA];
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += CH64) {
    const int nt = (int)min((int64_t)CH64, r1 - c0);
    // ---------------- phase A
    bool m1n = false;
    if (tid < nt) {
      TPrm64 &q = S.sq[tid];
      const int64_t t = c0 + tid;
      const int4 tr = v.trip[t];
      const int aux = v.trip_aux[t];
      Geo3 G;
      d_geo3(v, tr, aux, G);
      double pm[9];
      d_pmat<double>(G.eps, pm);
#pragma unroll
      for (int k = 0; k < 6; k++) q.pm[k] = pm[k];
      q.cls = (unsigned char)((G.nmu == 4 ? 2 : 0) + (G.nnu == 4 ? 1 : 0));
      double p[5], sint;
      d_legendre(G.cost, p, &sint);
      q.sint = sint;
#pragma unroll
      for (int k = 0; k < 4; k++) q.p[k] = p[k + 1];
      // the neutral group of the triplet, from its list-build word
      int gid = aux & 0xffff;
      if (gid == 0xffff) {
        status |= ST_NO_TABLE;
        gid = 0;
      }
      const DGroup3c g = group3c_scalar(v, gid);
      const int nx = g.nx, ny = g.ny;
      const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
      int ix = (int)(G.x / dx) + 1, iy = (int)(G.z / dy) + 1;
      if (ix < 2) ix = 2;
      if (iy < 2) iy = 2;
      if (ix > nx - 2) ix = nx - 2;
      if (iy > ny - 2) iy = ny - 2;
      q.px = G.x / dx - ix + 1.0;
      q.py = G.z / dy - iy + 1.0;
      q.rowstride = nx * g.nME;
      q.tstride = ny * nx * g.nME;
      q.base = g.off + ((int64_t)(iy - 1) * nx + (ix - 1)) * g.nME;
      const int64_t icell = (int64_t)(iy - 1) * nx + (ix - 1);
      q.m1[0] = v.meth3c[g.moff + icell];
      q.gk[0] = 0;
      m1n = q.m1[0];
      int nterm = 1;
      if (CHARGED) {
        TChg64 &c = S.sx[tid];
        c.z = G.z;
        c.d13 = G.d13;
        c.d23 = G.d23;
        c.stinky = G.stinky;
        c.pair = v.trip_pair[t];
        for (int isorp = 1; isorp <= v.sp.nssh[G.indna - 1] && nterm < KT_SP; isorp++) {
          const double dQ = v.dQ[tr.x * FB_MAXSH + isorp - 1];
          if (dQ == 0.0) continue;   // the reference adds an exact zero here
          const int gq = group3c_id(v, G.in1, G.in2, G.indna, isorp);
          if (gq < 0) {
            status |= ST_NO_TABLE;
            continue;
          }
          const DGroup3c gc = v.grp3c[gq];
          c.dq[nterm] = dQ;
          q.gk[nterm] = (int)(gc.off - g.off);
          q.m1[nterm] = v.meth3c[gc.moff + icell];
          nterm++;
        }
      }
      q.nterms = (unsigned char)nterm;
    }
    // the chunk's triplets in a list with the 3-point cells first, so that the
    // lanes of a wavefront in phase B mostly take the same branch
    {
      const unsigned long long b1 = __ballot(tid < nt && m1n), b3 = __ballot(tid < nt && !m1n);
      const unsigned long long below = (tid == 0) ? 0ull : (~0ull >> (64 - tid));
      if (tid < nt) S.sOrd[m1n ? __popcll(b1 & below) : __popcll(b1) + __popcll(b3 & below)] = (unsigned char)tid;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---------------- phase B: one lane per (triplet, term, ME)
    for (int it = tid; it < nt * KT * NME3; it += CH64) {
      const int k = it / (KT * NME3), r = it - k * (KT * NME3);
      const int term = r / NME3, m = r - term * NME3;
      const int t = S.sOrd[k];
      const TPrm64 &q = S.sq[t];
      if (term >= q.nterms) continue;
      double F;
      switch (q.cls) {
        case 3: F = me3c<10>(v, q, term, m); break;
        case 2:
        case 1:
          if (m >= 3) continue;
          F = me3c<3>(v, q, term, m);
          break;
        default:
          if (m >= 1) continue;
          F = me3c<1>(v, q, term, m);
          break;
      }
      S.sF[t][term][m] = F;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---------------- phase C: one lane per triplet, its terms in turn
    double X[KT][16];
    int nterm = 0, cls = 0;
    TChg64 c{};   // read before the contribution rows overwrite it
    if (tid < nt) {
      if (CHARGED) c = S.sx[tid];
      const TPrm64 &q = S.sq[tid];
      double pm[9];
#pragma unroll
      for (int k = 0; k < 6; k++) pm[k] = q.pm[k];
      // xp = yp x zp of epsilon.f90, the same expression in the (y, z, x) order
      pm[6] = pm[1] * pm[5] - pm[2] * pm[4];
      pm[7] = pm[2] * pm[3] - pm[0] * pm[5];
      pm[8] = pm[0] * pm[4] - pm[1] * pm[3];
      nterm = q.nterms;
      cls = q.cls;
#pragma unroll
      for (int j = 0; j < KT; j++) {
        if (j >= nterm) break;
        double F[NME3];
#pragma unroll
        for (int m = 0; m < NME3; m++) F[m] = S.sF[tid][j][m];
        switch (cls) {
          case 3: rot3c_sp<double>(3, pm, F, X[j]); break;
          case 2: rot3c_sp<double>(2, pm, F, X[j]); break;
          case 1: rot3c_sp<double>(1, pm, F, X[j]); break;
          default: rot3c_sp<double>(0, pm, F, X[j]); break;
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (tid < nt) {
      const int nbl = (cls & 2 ? 4 : 1) * (cls & 1 ? 4 : 1);
#pragma unroll
      for (int i = 0; i < 16; i++)
        if (i < nbl) sC[tid][0][i] = X[0][i] * FB_P_EQ2;
      if (CHARGED && nterm > 1) {
        const double *sb = v.mat[FB_MAT_S] + v.boff[c.pair];
        const double *db = v.mat[FB_MAT_DIP] + v.boff[c.pair];
#pragma unroll
        for (int j = 1; j < KT; j++) {
          if (j >= nterm) break;
#pragma unroll
          for (int i = 0; i < 16; i++) {
            if (i >= nbl) break;
            double sterm = sb[i] / 2.0, dterm = db[i] / c.z;
            double emnpl = (sterm - dterm) / c.d13 + (sterm + dterm) / c.d23;
            sC[tid][j][i] = c.dq[j] * (c.stinky * X[j][i] + (1.0 - c.stinky) * emnpl) * FB_P_EQ2;
          }
        }
      }
      // the term count travels in the row's last slot
      sC[tid][0][NB] = (double)nterm;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    // ---------------- phase D: ordered accumulation (the segment's sum
    // starts from zero in its first chunk and continues from the block)
    for (int sl = tid; sl < nown * NB; sl += CH64) {
      const int sg = sl / NB, e = sl - sg * NB;
      const int s0 = (int)(S.sSeg[sg] - c0), s1 = (int)(S.sSeg[sg + 1] - c0);
      const int k0 = max(s0, 0), k1 = min(s1, nt);
      if (k0 < k1 && e < S.sNb[sg]) {
        double *const o = vna + S.sBo[sg] + e;
        double a = (s0 < 0) ? *o : 0.0;
        for (int k = k0; k < k1; k++) {
          const int n = CHARGED ? (int)sC[k][0][NB] : 1;
          for (int j = 0; j < n; j++) a = a + sC[k][j][e];
        }
        *o = a;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (status) atomicOr(v.status, status);
}

extern "C" int fbk_3c_fp64_ph(const DView &v, int charged, hipStream_t st) {
  if (v.ntiles[0] == 0) return 0;
  const dim3 grid((unsigned)v.ntiles[0]);
  if (charged)
    hipLaunchKernelGGL(k_3c_fp64_ph<true>, grid, dim3(CH64), 0, st, v);
  else
    hipLaunchKernelGGL(k_3c_fp64_ph<false>, grid, dim3(CH64), 0, st, v);
  return (int)hipGetLastError();
}
