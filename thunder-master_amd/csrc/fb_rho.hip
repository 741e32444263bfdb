// McWEDA input densities, j.ASSEMBLERS/M_assemble_rho_McWEDA_DOGS.f90.
//
// rho_in / rho_bond (orbital blocks, block layout) and rho_in_weighted /
// rho_bond_weighted (shell blocks nssh_mu x nssh_nu): sums over the shells
// isorp of the contributing atom of Qin(isorp) times an interpolated table,
// rotated into the crystal frame for the orbital blocks.  The reference adds
// every (neighbour or triplet, isorp) term into its block as the term is formed;
// the kernels keep that order for every block:
//  * the ontop terms of a pair are summed by the pair's lane in sequence;
//  * the atom terms of every neighbour go to the self block of the atom: each
//    lane stages the products Qin * term of its neighbour in LDS, then one lane
//    per block element adds them in neighbour order, isorp inner;
//  * the 3c terms of a target block are added by one owner thread per target
//    block in list order (segments of the triplets sorted by target pair), on
//    top of the two-centre value already in the block.
// Binary64 throughout; no atomics, deterministic.
#include "fb_device.h"
#include "fb_3c_dev.h"

#define NBR 16    // elements of one orbital or shell block of s/p species (at most 4 x 4)

// the generic kernels' sizes: s/p species (MO 4), or with d shells (MO 9,
// blocks of up to 81 elements; fewer lanes keep the staged terms in LDS small)
template <int MO> struct GenR;
template <> struct GenR<FB_MO> {
  static constexpr int W = 64;    // lanes (pairs) per atom and pass (2c)
  static constexpr int T = 64;    // triplets per 3c tile
};
template <> struct GenR<FB_MOD> {
  static constexpr int W = 16;
  static constexpr int T = 16;
};

__device__ __forceinline__ int cell2c_r(const DView &v, int in1, int in2, int slot, int isub) {
  return v.idx2c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * FB_NSLOT2 + slot) * FB_MAXSUB2 + isub];
}

__device__ __forceinline__ int group3c_r(const DView &v, int in1, int in2, int in3, int w, int isub) {
  return v.idx3r[((((in1 - 1) * v.sp.ns + (in2 - 1)) * v.sp.ns + (in3 - 1)) * 2 + w) * FB_MAXSUB3 + isub];
}

// getMEs_Fdata_2c of one cell recovered into hm (column-major nr x nc)
template <int MO>
__device__ inline int me2c_block(const DView &v, int cid, int nr, int nc, double x, double *hm) {
  if (cid < 0) return ST_NO_TABLE;
  const DCell2c c = v.cell2c[cid];
  if (c.nME > MO * MO) return ST_CAPACITY;
  double F[MO * MO];
  int st = d_interp2c(c, v.tab2c, x, F);
  if (st) return st;
  for (int i = 0; i < nr * nc; i++) hm[i] = 0.0;
  for (int k = 0; k < c.nME; k++) hm[(v.nu[c.munu + k] - 1) * nr + (v.mu[c.munu + k] - 1)] = F[k];
  return 0;
}

// one term of a block: the table block (rotated for orbital blocks, W = false;
// dm: the d-shell rotation when MO = 9)
template <bool W, int MO>
__device__ inline int term2c(const DView &v, int cid, int in1, int in3, int nr, int nc, double z, const double *pm,
                             const double *dm, double *X) {
  if (W) return me2c_block<MO>(v, cid, nr, nc, z, X);
  double hm[MO * MO];
  int st = me2c_block<MO>(v, cid, nr, nc, z, hm);
  if (st) return st;
  d_rotate<double>(v.sp, in1, in3, pm, hm, nr, X, MO == FB_MOD ? dm : nullptr);
  return 0;
}

// assemble_rho_2c (W = false, :90-341) and assemble_rho_weighted_2c (W = true,
// :534-762): one workgroup of WR lanes per atom, lane k takes the atom's pair k
// (chunks of WR).  Dynamic LDS: [WR][stride] staged atom terms, isorp major,
// NB elements per term.
template <bool W, int MO>
__global__ __launch_bounds__(GenR<MO>::W) void k_rho_2c(DView v, int stride) {
  constexpr int WR = GenR<MO>::W, NB = MO * MO, NE = (NB + WR - 1) / WR;
  extern __shared__ double s_t[];
  __shared__ int s_n[WR];
  const int ia = blockIdx.x;
  const int lane = threadIdx.x;
  const int s = v.atom_struct[ia];
  const DStruct St = v.st[s];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const int in1 = v.imass[ia];
  const int nsi = v.sp.nssh[in1 - 1];
  const int nmu = W ? nsi : v.sp.norb[in1 - 1];
  const int n = v.neighn[ia];
  const int64_t p0 = v.noff[ia];
  const int kself = v.neigh_self[ia] - 1;
  const int64_t pself = p0 + kself;
  const double r1[3] = {v.ratom[3 * ia], v.ratom[3 * ia + 1], v.ratom[3 * ia + 2]};
  double *out_in = v.mat[W ? FB_MAT_RHOW_IN : FB_MAT_RHO_IN];
  double *out_bond = v.mat[W ? FB_MAT_RHOW_BOND : FB_MAT_RHO_BOND];
  const int64_t *off = W ? v.wboff : v.boff;
  const int nself = nmu * nmu;
  int status = 0;
  // the lane owns elements lane + j WR (j < NE) of the self blocks
  double acc_in[NE], acc_bond[NE];
  for (int j = 0; j < NE; j++) {
    acc_in[j] = 0.0;
    acc_bond[j] = 0.0;
  }

  for (int c0 = 0; c0 < n; c0 += WR) {
    const int k = c0 + lane;
    int nsj = 0;
    if (k < n) {
      const int64_t p = p0 + k;
      const int ja = St.atom_off + v.neigh_j[p] - 1;
      const int mb = v.neigh_b[p];
      const int in2 = v.imass[ja];
      nsj = v.sp.nssh[in2 - 1];
      const int nnu = W ? nsj : v.sp.norb[in2 - 1];
      const int nb = nmu * nnu;
      double r2[3];
      for (int q = 0; q < 3; q++) r2[q] = v.ratom[3 * ja + q] + xls[3 * mb + q];
      const double z = d_distance(r1, r2);
      double pm[9], dm[MO == FB_MOD ? 25 : 1];
      if (!W) {
        double sighat[3], eps[9];
        if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
        else for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / z;
        d_epsilon(r2, sighat, eps);
        d_pmat<double>(eps, pm);
        if (MO == FB_MOD) d_dmat(eps, dm);
      }
      double X[NB];
      if (k != kself) {
        // ontop terms: potential of iatom's shells, then of jatom's
        double acc[NB];
        for (int i = 0; i < nb; i++) acc[i] = 0.0;
        for (int isorp = 1; isorp <= nsi; isorp++) {
          const double q = v.qin[ia * FB_MAXSH + isorp - 1];
          status |= term2c<W, MO>(v, cell2c_r(v, in1, in2, W ? SLOT_RHOS_ONL : SLOT_RHO_ONL, isorp), in1, in2, nmu,
                                  nnu, z, pm, dm, X);
          for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i] * q;
        }
        for (int isorp = 1; isorp <= nsj; isorp++) {
          const double q = v.qin[ja * FB_MAXSH + isorp - 1];
          status |= term2c<W, MO>(v, cell2c_r(v, in1, in2, W ? SLOT_RHOS_ONR : SLOT_RHO_ONR, isorp), in1, in2, nmu,
                                  nnu, z, pm, dm, X);
          for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i] * q;
        }
        const int64_t bo = off[p];
        for (int i = 0; i < nb; i++) {
          out_in[bo + i] = acc[i];
          out_bond[bo + i] = acc[i];
        }
      }
      // atom terms <i|rho_j|i>, staged as products Qin * term
      for (int isorp = 1; isorp <= nsj; isorp++) {
        const double q = v.qin[ja * FB_MAXSH + isorp - 1];
        status |= term2c<W, MO>(v, cell2c_r(v, in1, in2, W ? SLOT_RHOS_ATOM : SLOT_RHO_ATOM, isorp), in1, in1, nmu,
                                nmu, z, pm, dm, X);
        for (int i = 0; i < nself; i++) s_t[lane * stride + (isorp - 1) * NB + i] = X[i] * q;
      }
    }
    s_n[lane] = nsj;
    __syncthreads();
    const int kmax = min(WR, n - c0);
    for (int j = 0; j < NE; j++) {
      const int e = lane + j * WR;
      if (e >= nself) break;
      for (int kk = 0; kk < kmax; kk++) {
        const bool self = (c0 + kk == kself);
        for (int isorp = 0; isorp < s_n[kk]; isorp++) {
          const double t = s_t[kk * stride + isorp * NB + e];
          acc_in[j] = acc_in[j] + t;
          if (self) acc_bond[j] = acc_bond[j] + t;
        }
      }
    }
    __syncthreads();
  }
  for (int j = 0; j < NE; j++) {
    const int e = lane + j * WR;
    if (e >= nself) break;
    out_in[off[pself] + e] = acc_in[j];
    out_bond[off[pself] + e] = acc_bond[j];
  }
  if (status) atomicOr(v.status, status);
}

// triplet geometry of the density assemblers, M_assemble_rho_McWEDA_DOGS.f90:451-468
__device__ inline void rho_geo(const DView &v, int4 tr, double &x, double &z, double &cost, double *pm,
                               double *eps_out = nullptr) {
  const int ial = tr.x, gia = tr.y, gja = tr.z;
  const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
  const int s = v.atom_struct[ial];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  double r1[3], r2[3], rna[3], r21[3], sighat[3], rnabc[3], rhat[3], eps[9];
  for (int k = 0; k < 3; k++) {
    r1[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
    r2[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
    rna[k] = v.ratom[3 * ial + k];
    r21[k] = r2[k] - r1[k];
  }
  z = d_distance(r1, r2);
  if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) sighat[k] = r21[k] / z;
  for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1[k] + r21[k] / 2.0);
  x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
  if (x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
  else for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] / x;
  cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
  d_epsilon(rhat, sighat, eps);
  d_pmat<double>(eps, pm);
  if (eps_out)
    for (int k = 0; k < 9; k++) eps_out[k] = eps[k];
}

// assemble_rho_3c (W = false, :362-511) and assemble_rho_weighted_3c (W = true,
// :783-924).  A workgroup owns the target segments that start in its tile of
// T3R sorted triplets; every thread computes the terms of one triplet per
// chunk into LDS ([T3R][stride], isorp major), and one owner thread per segment
// adds them in order onto the two-centre block.
template <bool W, int MO>
__global__ __launch_bounds__(GenR<MO>::T) void k_rho_3c(DView v, int stride) {
  constexpr int T3R = GenR<MO>::T, NB = MO * MO;
  extern __shared__ double s_c[];
  __shared__ int s_n[T3R];
  const int64_t ts = xcd_tile(blockIdx.x, gridDim.x) * T3R;
  const int64_t te = min(ts + T3R, v.ntrip);
  int64_t first, last, r0, r1;
  if (!tile_segments(v, ts, te, first, last, r0, r1)) return;
  const int64_t my = first + threadIdx.x;
  const bool own = my <= last;
  double *out = v.mat[W ? FB_MAT_RHOW_IN : FB_MAT_RHO_IN];
  const int64_t *off = W ? v.wboff : v.boff;
  int64_t a = 0, b = 0, dst = 0;
  int nb = 0;
  double acc[NB];
  if (own) {
    a = v.seg_off[my];
    b = v.seg_off[my + 1];
    const int p = v.seg_pair[my];
    const int ia = v.pair_atom[p];
    const int ja = v.st[v.atom_struct[ia]].atom_off + v.neigh_j[p] - 1;
    const int in1 = v.imass[ia], in2 = v.imass[ja];
    nb = W ? v.sp.nssh[in1 - 1] * v.sp.nssh[in2 - 1] : v.sp.norb[in1 - 1] * v.sp.norb[in2 - 1];
    dst = off[p];
    for (int i = 0; i < NB; i++) acc[i] = (i < nb) ? out[dst + i] : 0.0;
  }
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += T3R) {
    const int64_t t = c0 + threadIdx.x;
    int nsorp = 0;
    if (t < r1) {
      const int4 tr = v.trip[t];
      const int in1 = v.imass[tr.y], in2 = v.imass[tr.z], indna = v.imass[tr.x];
      const int nr = W ? v.sp.nssh[in1 - 1] : v.sp.norb[in1 - 1];
      const int nc = W ? v.sp.nssh[in2 - 1] : v.sp.norb[in2 - 1];
      double x, z, cost, pm[9], eps[9], dm[MO == FB_MOD ? 25 : 1], hm[NB], X[NB];
      rho_geo(v, tr, x, z, cost, pm, eps);
      if (MO == FB_MOD && !W) d_dmat(eps, dm);
      nsorp = v.sp.nssh[indna - 1];
      for (int isorp = 1; isorp <= nsorp; isorp++) {
        const double q = v.qin[tr.x * FB_MAXSH + isorp - 1];
        status |= d_getMEs_3c<NB>(v, group3c_r(v, in1, in2, indna, W ? 1 : 0, isorp), x, z, cost, nr, nc, hm);
        if (W) {
          for (int i = 0; i < nr * nc; i++) X[i] = hm[i];
        } else {
          d_rotate<double>(v.sp, in1, in2, pm, hm, nr, X, MO == FB_MOD ? dm : nullptr);
        }
        for (int i = 0; i < nr * nc; i++) s_c[threadIdx.x * stride + (isorp - 1) * NB + i] = X[i] * q;
      }
    }
    s_n[threadIdx.x] = nsorp;
    __syncthreads();
    if (own) {
      const int64_t k0 = max(a, c0), k1 = min(b, c0 + T3R);
      for (int64_t kk = k0; kk < k1; kk++) {
        const int row = (int)(kk - c0);
        for (int isorp = 0; isorp < s_n[row]; isorp++)
          for (int i = 0; i < NB; i++)
            if (i < nb) acc[i] = acc[i] + s_c[row * stride + isorp * NB + i];
      }
    }
    __syncthreads();
  }
  if (own)
    for (int i = 0; i < NB; i++)
      if (i < nb) out[dst + i] = acc[i];
  if (status) atomicOr(v.status, status);
}

// ---------------------------------------------------------------------------
// s/p species with canonical tables: the same sums with compile-time block
// shapes, the register-resident rotations of fb_sp.h, the Lagrange products
// with the reciprocals formed in-kernel by the reference's division, and the
// cell headers read through the scalar cache (the operations and their order
// are those of the generic kernels).  Pair class cls = (mu side is s+p) * 2 +
// (nu side is s+p); an s+p species has two shells, an s species one, so the
// class also fixes the shell-block shape.
#include "fb_sp.h"
#include "fb_2c_dev.h"

#define WRS 256   // lanes per 2c workgroup (whole atoms, at most 16)
static_assert(WRS == FB_2C_PCAP, "the 2c lane order (k_pair_perm) is laid out for passes of FB_2C_PCAP pairs");

// shell block from table order (issh major, make_munuS) to column-major
template <int NSI, int NSJ>
__device__ __forceinline__ void shell_block(const double *F, double *X) {
#pragma unroll
  for (int i = 0; i < NSI; i++)
#pragma unroll
    for (int j = 0; j < NSJ; j++) X[j * NSI + i] = F[i * NSJ + j];
}

// one 2c term of class CLS: orbital block (rotated) or shell block (W)
template <bool W, int CLS>
__device__ __forceinline__ int term2c_sp(const DView &v, int sp2, int slot, int isub, double z, const double *pm,
                                         int &grid_cached, int &ileft, double *prod, double *X) {
  DCell2c c;
  if (cell_scalar(v, sp2, slot, isub, c) < 0) return ST_NO_TABLE;
  if (z < 0.0) return ST_X_NEG;
  if (z > c.xmax * (1.0 + 1.0e-5)) return ST_X_RANGE;
  if (c.grid != grid_cached) {
    const DGrid2c gr = grid_scalar(v, c.grid);
    ileft = window7(z, gr.dx, gr.nx);
    lagrange7_div(z, gr.dx, ileft, prod);
    grid_cached = c.grid;
  }
  constexpr int NSI = (CLS & 2) ? 2 : 1, NSJ = (CLS & 1) ? 2 : 1;
  if (W) {
    double F[NSI * NSJ];
    interp_cell<NSI * NSJ>(v, c, ileft, prod, F);
    shell_block<NSI, NSJ>(F, X);
  } else {
    double F[nme2c_of(CLS)];
    interp_cell<nme2c_of(CLS)>(v, c, ileft, prod, F);
    rot2c_sp<double>(CLS, pm, F, X);
  }
  return 0;
}

template <bool W, int CLS>
__device__ __forceinline__ int pair_rho_sp(const DView &v, int ia, int ja, int in1, int in2, bool self, double z,
                                           const double *pm, int64_t bo, double *st_terms) {
  constexpr int NSI = (CLS & 2) ? 2 : 1, NSJ = (CLS & 1) ? 2 : 1;
  constexpr int NR = W ? NSI : ((CLS & 2) ? 4 : 1), NC = W ? NSJ : ((CLS & 1) ? 4 : 1), NB = NR * NC;
  constexpr int CLSA = (CLS & 2) ? CLS_SPSP : CLS_SS;   // atom terms live on iatom's orbitals or shells
  constexpr int NA = W ? NSI * NSI : ((CLS & 2) ? 16 : 1);
  int status = 0, grid_cached = -1, ileft = 1;
  double prod[7], X[16];
  if (!self) {
    double acc[NB];
#pragma unroll
    for (int i = 0; i < NB; i++) acc[i] = 0.0;
#pragma unroll
    for (int isorp = 1; isorp <= NSI; isorp++) {
      const double q = v.qin[ia * FB_MAXSH + isorp - 1];
      status |= term2c_sp<W, CLS>(v, in1 | (in2 << 16), W ? SLOT_RHOS_ONL : SLOT_RHO_ONL, isorp, z, pm,
                                  grid_cached, ileft, prod, X);
#pragma unroll
      for (int i = 0; i < NB; i++) acc[i] = acc[i] + X[i] * q;
    }
#pragma unroll
    for (int isorp = 1; isorp <= NSJ; isorp++) {
      const double q = v.qin[ja * FB_MAXSH + isorp - 1];
      status |= term2c_sp<W, CLS>(v, in1 | (in2 << 16), W ? SLOT_RHOS_ONR : SLOT_RHO_ONR, isorp, z, pm,
                                  grid_cached, ileft, prod, X);
#pragma unroll
      for (int i = 0; i < NB; i++) acc[i] = acc[i] + X[i] * q;
    }
    double *oi = v.mat[W ? FB_MAT_RHOW_IN : FB_MAT_RHO_IN] + bo;
    double *ob = v.mat[W ? FB_MAT_RHOW_BOND : FB_MAT_RHO_BOND] + bo;
#pragma unroll
    for (int i = 0; i < NB; i++) {
      oi[i] = acc[i];
      ob[i] = acc[i];
    }
  }
#pragma unroll
  for (int isorp = 1; isorp <= NSJ; isorp++) {
    const double q = v.qin[ja * FB_MAXSH + isorp - 1];
    status |= term2c_sp<W, CLSA>(v, in1 | (in2 << 16), W ? SLOT_RHOS_ATOM : SLOT_RHO_ATOM, isorp, z, pm,
                                 grid_cached, ileft, prod, X);
#pragma unroll
    for (int i = 0; i < NA; i++) st_terms[(isorp - 1) * NBR + i] = X[i] * q;
  }
  return status;
}

// one lane per pair over workgroups of whole atoms (at most 16 atoms and
// FB_2C_PCAP pairs); 16 owner lanes per atom add the staged atom terms of the
// atom's neighbours into its self blocks.  Dynamic LDS: [WRS][stride].
// two waves per SIMD (the LDS rows of 256 lanes allow no more): the
// unweighted form then fits in VGPRs alone, without AGPR spill slots
template <bool W>
__global__ __launch_bounds__(WRS, 2) void k_rho_2c_sp(DView v, int stride) {
  extern __shared__ double s_t[];
  __shared__ int s_n[WRS];
  const int tid = threadIdx.x;
  const int a0 = v.wg2c[blockIdx.x], a1 = v.wg2c[blockIdx.x + 1];
  const int64_t pb = v.noff[a0], pe = v.noff[a1];
  const int ra = a0 + (tid >> 4), re = tid & 15;
  const bool rown = ra < a1;
  int64_t rp0 = 0, rp1 = 0, rself = -1;
  int rn = 0;
  if (rown) {
    rp0 = v.noff[ra];
    rp1 = v.noff[ra + 1];
    rself = rp0 + v.neigh_self[ra] - 1;
    rn = W ? v.sp.nssh[v.imass[ra] - 1] : v.sp.norb[v.imass[ra] - 1];
  }
  double acc_in = 0.0, acc_bond = 0.0;
  int status = 0;
  for (int64_t c0 = pb; c0 < pe; c0 += WRS) {
    // the lane at slot c0 + tid takes pair p of its wavefront's group (the
    // 2c lane order of v.pair_perm); LDS rows and sums go by p
    const int64_t slot = c0 + tid;
    const int64_t p = (slot < pe && v.pair_perm) ? c0 + (tid & ~63) + v.pair_perm[slot] : slot;
    const int srow = (int)(p - c0);
    int nsj = 0;
    if (slot < pe) {
      const int ia = v.pair_atom[p];
      const int s = v.atom_struct[ia];
      const int ja = v.st[s].atom_off + v.neigh_j[p] - 1;
      const int mb = v.neigh_b[p];
      const int in1 = v.imass[ia], in2 = v.imass[ja];
      const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
      double r1[3], r2[3], pm[9];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        r1[q] = v.ratom[3 * ia + q];
        r2[q] = v.ratom[3 * ja + q] + xls[3 * mb + q];
      }
      const double z = d_distance(r1, r2);
      if (!W) {
        double sighat[3], eps[9];
        if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
        else {
#pragma unroll
          for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / z;
        }
        d_epsilon(r2, sighat, eps);
        d_pmat<double>(eps, pm);
      }
      const bool self = (ja == ia && mb == 0);
      nsj = v.sp.nssh[in2 - 1];
      const int cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
      const int64_t bo = W ? v.wboff[p] : v.boff[p];
      double *mine = s_t + srow * stride;
      switch (cls) {
        case 3: status |= pair_rho_sp<W, 3>(v, ia, ja, in1, in2, self, z, pm, bo, mine); break;
        case 2: status |= pair_rho_sp<W, 2>(v, ia, ja, in1, in2, self, z, pm, bo, mine); break;
        case 1: status |= pair_rho_sp<W, 1>(v, ia, ja, in1, in2, self, z, pm, bo, mine); break;
        default: status |= pair_rho_sp<W, 0>(v, ia, ja, in1, in2, self, z, pm, bo, mine); break;
      }
    }
    s_n[srow] = nsj;
    __syncthreads();
    if (rown && re < rn * rn) {
      const int64_t k0 = max(rp0, c0), k1 = min(rp1, c0 + WRS);
      for (int64_t k = k0; k < k1; k++) {
        const int row = (int)(k - c0);
        for (int isorp = 0; isorp < s_n[row]; isorp++) {
          const double t = s_t[row * stride + isorp * NBR + re];
          acc_in = acc_in + t;
          if (k == rself) acc_bond = acc_bond + t;
        }
      }
    }
    __syncthreads();
  }
  if (rown && re < rn * rn) {
    const int64_t o = W ? v.wboff[rself] : v.boff[rself];
    v.mat[W ? FB_MAT_RHOW_IN : FB_MAT_RHO_IN][o + re] = acc_in;
    v.mat[W ? FB_MAT_RHOW_BOND : FB_MAT_RHO_BOND][o + re] = acc_bond;
  }
  if (status) atomicOr(v.status, status);
}

// assemble_rho_3c / assemble_rho_weighted_3c for s/p species with canonical
// density tables.  A workgroup of RW lanes owns whole target segments, packed
// at list build into tiles of at most RT triplets (the mixed kernel's 64-triplet
// tiles), and works through them in four phases:
//   A  one lane per triplet: binary64 geometry (rho_geo), Legendre P0..P4 and
//      sin(theta), and per shell isorp of the third atom the group, the grid
//      cell, px, py, the method of the cell and Qin, into LDS.
//   B  one lane per (triplet, isorp, matrix element): the interpolation of
//      getMEs_3c_sp for that element -- the five theta cells in order, each
//      with the 16-point form or the 3-point rule, summed with P_theta --
//      times sin(theta) where mvalue = 1.
//   C  one lane per (triplet, isorp): rotation (or the shell-block reorder of
//      the weighted density) times Qin into the triplet's LDS row.
//   D  one lane per (owned segment, block element): the rows added onto the
//      two-centre block in list order, isorp inner, as the reference does.
// Every element goes through the same binary64 operations in the same order
// as the one-lane-per-triplet form, so the blocks are bit-identical to it.
#define RT 64           // triplets per tile
#define RW 256          // lanes per workgroup
#define RSO 2           // shells of an s/p third atom

struct RTri {           // per triplet
  double p[5], sint, pm[9];
  int cls, nsorp;
};
struct RShell {         // per (triplet, isorp)
  double px, py, q;
  int64_t base;         // FP64 slab index of stencil corner (ix-1, iy-1), theta cell 1, ME 1
  int rs, ts, nme, m1;  // row and theta strides (elements), ME count (0: no table), 3-point rule
};

// the unweighted density at three waves per SIMD (140 VGPRs): at four its
// binary64 rotation spills 48 bytes per lane, whose scratch traffic doubles
// the kernel's HBM bytes for a 5 % shorter run; the weighted one fits four
template <bool W>
__global__ __launch_bounds__(RW) __attribute__((amdgpu_waves_per_eu(W ? 4 : 3))) void k_rho_3c_ph(DView v) {
  constexpr int NMX = W ? 4 : 10;          // matrix elements per (triplet, isorp)
  constexpr int NBX = W ? 4 : 16;          // block elements
  constexpr int ROW = NBX + 1;             // LDS row stride (doubles)
  constexpr size_t SZ_A = sizeof(RTri) * RT + sizeof(RShell) * RT * RSO + sizeof(double) * RT * RSO * NMX;
  constexpr size_t SZ_R = sizeof(double) * RT * RSO * ROW;
  __shared__ __align__(16) unsigned char sbuf[SZ_A > SZ_R ? SZ_A : SZ_R];
  __shared__ int64_t sSeg[RT + 1], sBo[RT];
  __shared__ int sNb[RT], sNs[RT];
  RTri *sT = reinterpret_cast<RTri *>(sbuf);
  RShell *sS = reinterpret_cast<RShell *>(sbuf + sizeof(RTri) * RT);
  double *sF = reinterpret_cast<double *>(sbuf + sizeof(RTri) * RT + sizeof(RShell) * RT * RSO);
  double *sR = reinterpret_cast<double *>(sbuf);   // rows, after phase C has read A and B
  const int tid = threadIdx.x;
  const int4 td = v.tiles[0][xcd_tile(blockIdx.x, gridDim.x)];
  const int64_t first = td.x, last = td.y, r0 = td.z, r1 = td.w;
  if (first > last) return;
  const int nown = (int)(last - first + 1);
  double *const out = v.mat[W ? FB_MAT_RHOW_IN : FB_MAT_RHO_IN];
  const int64_t *const off = W ? v.wboff : v.boff;
  for (int k = tid; k <= nown; k += RW) sSeg[k] = v.seg_off[first + k];
  for (int k = tid; k < nown; k += RW) {
    const int p = v.seg_pair[first + k];
    const int64_t b0 = off[p];
    sBo[k] = b0;
    sNb[k] = (int)(off[p + 1] - b0);
  }
  int status = 0;
  for (int64_t c0 = r0; c0 < r1; c0 += RT) {
    const int nt = (int)min((int64_t)RT, r1 - c0);
    // ---------------- phase A
    if (tid < nt) {
      const int4 tr = v.trip[c0 + tid];
      const int in1 = v.imass[tr.y], in2 = v.imass[tr.z], indna = v.imass[tr.x];
      RTri &T = sT[tid];
      double x, z, cost;
      rho_geo(v, tr, x, z, cost, T.pm);
      d_legendre(cost, T.p, &T.sint);
      T.cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
      int nsorp = v.sp.nssh[indna - 1];
      if (nsorp > RSO) {
        status |= ST_NO_TABLE;
        nsorp = RSO;
      }
      T.nsorp = nsorp;
      sNs[tid] = nsorp;
      for (int so = 0; so < nsorp; so++) {
        RShell &S = sS[tid * RSO + so];
        const int gid = group3c_r(v, in1, in2, indna, W ? 1 : 0, so + 1);
        S.q = v.qin[tr.x * FB_MAXSH + so];
        if (gid < 0) {
          status |= ST_NO_TABLE;
          S.nme = 0;
          continue;
        }
        const DGroup3c g = v.grp3c[gid];
        const int nx = g.nx, ny = g.ny;
        // e.FDATA/M_Fdata_3c.f90:367-381, as getMEs_3c_sp
        const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
        int ix = (int)(x / dx) + 1, iy = (int)(z / dy) + 1;
        if (ix < 2) ix = 2;
        if (iy < 2) iy = 2;
        if (ix > nx - 2) ix = nx - 2;
        if (iy > ny - 2) iy = ny - 2;
        S.px = x / dx - ix + 1.0;
        S.py = z / dy - iy + 1.0;
        S.nme = g.nME;
        S.rs = nx * g.nME;
        S.ts = ny * nx * g.nME;
        S.base = g.off + ((int64_t)(iy - 2) * nx + (ix - 2)) * g.nME;
        S.m1 = v.meth3c[g.moff + (int64_t)(iy - 1) * nx + (ix - 1)];
      }
    }
    __syncthreads();
    // ---------------- phase B
    for (int it = tid; it < nt * RSO * NMX; it += RW) {
      const int t = it / (RSO * NMX), r = it - t * (RSO * NMX), so = r / NMX, m = r - so * NMX;
      const RTri &T = sT[t];
      const RShell &S = sS[t * RSO + so];
      if (so >= T.nsorp || m >= S.nme) continue;
      const double px = S.px, py = S.py;
      const int rs = S.rs, nm = S.nme;
      const double *tab = v.tab3c + S.base + m;
      double F = 0.0;
#pragma unroll 1
      for (int th = 0; th < FB_NTHETA; th++) {
        const double *tt = tab + (int64_t)th * S.ts;
        double interim;
        if (S.m1) {
          interim = ((1.0 - px - py) * tt[rs + nm] + px * tt[rs + 2 * nm]) + py * tt[2 * rs + nm];
        } else {
          double gk[4];
#pragma unroll
          for (int k = 0; k < 4; k++) {
            const double *col = tt + k * nm;
            double f1m1 = col[0];
            double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
            double f0p3 = 3.0 * col[rs], f0p6 = 2.0 * f0p3;
            double f1p3 = 3.0 * col[2 * rs], f1p6 = 2.0 * f1p3;
            double f2p1 = col[3 * rs];
            double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
            double bb2 = f1m3 - f0p6 + f1p3;
            double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
            double bb0 = f0p6;
            gk[k] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
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
        F = F + T.p[th] * interim;
      }
      // sin(theta) on the mvalue = 1 elements: all but the leading 6 / 2 / 1
      // of the canonical order (none in the weighted shell blocks)
      if (!W && m >= (nm == 10 ? 6 : (nm == 3 ? 2 : 1))) F = F * T.sint;
      sF[(t * RSO + so) * NMX + m] = F;
    }
    __syncthreads();
    // ---------------- phase C
    if (W) {
      // shell blocks: one lane per (triplet, isorp)
      double X[NBX];
      int nbc = 0;
      double q = 0.0;
      const int tc = tid / RSO, soc = tid - tc * RSO;
      if (tc < nt && soc < sT[tc].nsorp) {
        const RShell &S = sS[tc * RSO + soc];
        double F[NMX];
#pragma unroll
        for (int m = 0; m < NMX; m++) F[m] = (m < S.nme) ? sF[(tc * RSO + soc) * NMX + m] : 0.0;
        q = S.q;
        switch (sT[tc].cls) {
          case 3: shell_block<2, 2>(F, X); nbc = 4; break;
          case 2: shell_block<2, 1>(F, X); nbc = 2; break;
          case 1: shell_block<1, 2>(F, X); nbc = 2; break;
          default: shell_block<1, 1>(F, X); nbc = 1; break;
        }
        if (S.nme == 0) nbc = 0;
      }
      __syncthreads();   // phase A and B data read; the rows overwrite them
      if (tc < nt && soc < RSO) {
        double *row = sR + (tc * RSO + soc) * ROW;
#pragma unroll
        for (int i = 0; i < NBX; i++) row[i] = (i < nbc) ? X[i] * q : 0.0;
      }
    } else {
      // rotated blocks: one lane per (triplet, isorp, block column), the
      // elements of rot3c_sp one at a time (rot3c_el: same terms, same order)
      constexpr int NCI = (RT * RSO * 4 + RW - 1) / RW;
      double Xc[NCI][4];
      int dst[NCI], nr[NCI];
#pragma unroll
      for (int k = 0; k < NCI; k++) {
        const int it = tid + k * RW, pr = it >> 2, col = it & 3, tc = pr / RSO, soc = pr - tc * RSO;
        dst[k] = -1;
        nr[k] = 0;
        if (tc < nt && soc < sT[tc].nsorp) {
          const RTri &T = sT[tc];
          const RShell &S = sS[pr];
          const int cls = T.cls, nmu = (cls & 2) ? 4 : 1, nnu = (cls & 1) ? 4 : 1;
          if (S.nme > 0 && col < nnu) {
            double F[NMX], pm[9];
#pragma unroll
            for (int m = 0; m < NMX; m++) F[m] = (m < S.nme) ? sF[pr * NMX + m] : 0.0;
#pragma unroll
            for (int e = 0; e < 9; e++) pm[e] = T.pm[e];
            const double q = S.q;
#pragma unroll
            for (int r = 0; r < 4; r++) {
              double x = 0.0;
              if (r < nmu) {
                switch (cls) {
                  case 3: x = rot3c_el<3>(r, col, pm, F); break;
                  case 2: x = rot3c_el<2>(r, col, pm, F); break;
                  case 1: x = rot3c_el<1>(r, col, pm, F); break;
                  default: x = rot3c_el<0>(r, col, pm, F); break;
                }
              }
              Xc[k][r] = x * q;
            }
            dst[k] = pr * ROW + nmu * col;
            nr[k] = nmu;
          }
        }
      }
      __syncthreads();   // phase A and B data read; the rows overwrite them
#pragma unroll
      for (int k = 0; k < NCI; k++)
        if (dst[k] >= 0)
#pragma unroll
          for (int r = 0; r < 4; r++)
            if (r < nr[k]) sR[dst[k] + r] = Xc[k][r];
    }
    __syncthreads();
    // ---------------- phase D
#pragma unroll
    for (int j = 0; j < (RT * NBX + RW - 1) / RW; j++) {
      const int sl = tid + j * RW;
      if (sl < nown * NBX) {
        const int sg = sl / NBX, e = sl - sg * NBX;
        const int64_t k0 = max(sSeg[sg], c0), k1 = min(sSeg[sg + 1], c0 + nt);
        if (k0 < k1 && e < sNb[sg]) {
          double *const o = out + sBo[sg] + e;
          double a = *o;
          for (int64_t k = k0; k < k1; k++) {
            const int row = (int)(k - c0);
            for (int so = 0; so < sNs[row]; so++) a = a + sR[(row * RSO + so) * ROW + e];
          }
          *o = a;
        }
      }
    }
    __syncthreads();
  }
  if (status) atomicOr(v.status, status);
}

// per pair: nssh_mu * nssh_nu (sizes of the shell blocks)
__global__ void k_shell_sizes(DView v, int64_t *sz) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= v.npairs) return;
  const int ia = v.pair_atom[p];
  const int ja = v.st[v.atom_struct[ia]].atom_off + v.neigh_j[p] - 1;
  sz[p] = (int64_t)v.sp.nssh[v.imass[ia] - 1] * v.sp.nssh[v.imass[ja] - 1];
}

// generic kernels, sized by the orbital count of the widest species
template <int MO>
static void launch_rho_2c(const DView &v, int weighted, int maxsh, hipStream_t st) {
  constexpr int WR = GenR<MO>::W;
  const int stride = maxsh * MO * MO + 1;
  const size_t lds = sizeof(double) * WR * stride;
  if (weighted)
    hipLaunchKernelGGL((k_rho_2c<true, MO>), dim3(v.natoms), dim3(WR), lds, st, v, stride);
  else
    hipLaunchKernelGGL((k_rho_2c<false, MO>), dim3(v.natoms), dim3(WR), lds, st, v, stride);
}

template <int MO>
static void launch_rho_3c(const DView &v, int weighted, int maxsh, hipStream_t st) {
  constexpr int T3R = GenR<MO>::T;
  const int stride = maxsh * MO * MO + 1;
  const size_t lds = sizeof(double) * T3R * stride;
  const dim3 grid((unsigned)((v.ntrip + T3R - 1) / T3R));
  if (weighted)
    hipLaunchKernelGGL((k_rho_3c<true, MO>), grid, dim3(T3R), lds, st, v, stride);
  else
    hipLaunchKernelGGL((k_rho_3c<false, MO>), grid, dim3(T3R), lds, st, v, stride);
}

extern "C" {
// maxsh: largest shell count of any species (sizes the staged terms); sp: the
// s/p kernels (every species s or s+p, canonical density tables)
int fbk_rho_2c(const DView &v, int weighted, int maxsh, int sp, hipStream_t st) {
  if (v.natoms == 0) return 0;
  const int stride = maxsh * NBR + 1;
  if (sp) {
    if (v.nwg2c == 0) return 0;
    const size_t lds = sizeof(double) * WRS * stride;
    if (weighted)
      hipLaunchKernelGGL(k_rho_2c_sp<true>, dim3(v.nwg2c), dim3(WRS), lds, st, v, stride);
    else
      hipLaunchKernelGGL(k_rho_2c_sp<false>, dim3(v.nwg2c), dim3(WRS), lds, st, v, stride);
    return (int)hipGetLastError();
  }
  if (fb_wide_species(v.sp)) launch_rho_2c<FB_MOD>(v, weighted, maxsh, st);
  else launch_rho_2c<FB_MO>(v, weighted, maxsh, st);
  return (int)hipGetLastError();
}

int fbk_rho_3c(const DView &v, int weighted, int maxsh, int sp, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  if (sp) {
    if (v.ntiles[0] == 0) return 0;
    const dim3 g((unsigned)v.ntiles[0]);
    if (weighted)
      hipLaunchKernelGGL(k_rho_3c_ph<true>, g, dim3(RW), 0, st, v);
    else
      hipLaunchKernelGGL(k_rho_3c_ph<false>, g, dim3(RW), 0, st, v);
    return (int)hipGetLastError();
  }
  if (fb_wide_species(v.sp)) launch_rho_3c<FB_MOD>(v, weighted, maxsh, st);
  else launch_rho_3c<FB_MO>(v, weighted, maxsh, st);
  return (int)hipGetLastError();
}

int fbk_shell_sizes(const DView &v, int64_t *sz, hipStream_t st) {
  if (v.npairs == 0) return 0;
  hipLaunchKernelGGL(k_shell_sizes, dim3((unsigned)((v.npairs + 255) / 256)), dim3(256), 0, st, v, sz);
  return (int)hipGetLastError();
}
}
