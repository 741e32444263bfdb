// Two-centre assembly: S, T, dipole_z and the Vna ontop / atom terms.
//
// One wavefront per central atom: the atom's pairs are contiguous in the CSR
// pair list, so lane l handles pair noff[iatom] + l (chunks of 64 when an
// atom has more neighbours).  Per pair the lane builds the bond frame once
// (binary64, b.FUNCTIONS/epsilon.f90 with the reference's argument choice
// epsilon(r2, sighat), j.ASSEMBLERS/M_assemble_2c_DOGS.f90:177), interpolates
// every needed table cell (7-point Lagrange), rotates and stores the blocks.
// The Vna atom term <i|V_j|i> of every neighbour is summed into the self block
// of iatom in neighbour order (M_assemble_2c_DOGS.f90:710-816) through LDS, so
// the sum is deterministic and in the reference's order.  With shell charges
// a neighbour contributes one term per charged shell besides the neutral one,
// and the reference adds each to the block on its own (:790-812): the terms
// are kept apart in LDS and added one by one.
//
// The kernel is bound by the block stores (HBM); its arithmetic stays binary64.
#include "fb_device.h"
#include "fb_2c_sp.h"

// the generic kernel's sizes: s/p species (MO 4), or with d shells (MO 9)
template <int MO> struct Gen2c;
template <> struct Gen2c<FB_MO> {
  static constexpr int W = 64;        // lanes (pairs) per central atom and pass
  static constexpr int NME = 16;      // non-zero 2c matrix elements
};
template <> struct Gen2c<FB_MOD> {
  static constexpr int W = 16;        // the LDS of the atom terms grows with MO^2
  static constexpr int NME = FB_NME2_WIDE;
};

// interpolate + recover + rotate one interaction into X (column-major nmu x nnu)
template <int MO>
__device__ inline int me2c_rotated(const DView &v, int cid, int in1, int in3, int nmu, double x, const double *pm,
                                   const double *dm, double *X) {
  if (cid < 0) return ST_NO_TABLE;
  const DCell2c c = v.cell2c[cid];
  double F[Gen2c<MO>::NME];
  int st = d_interp2c(c, v.tab2c, x, F);
  if (st) return st;
  double hm[MO * MO];
  int nnu = v.sp.norb[in3 - 1];
  for (int i = 0; i < nmu * nnu; i++) hm[i] = 0.0;
  for (int k = 0; k < c.nME; k++) hm[(v.nu[c.munu + k] - 1) * nmu + (v.mu[c.munu + k] - 1)] = F[k];
  d_rotate<double>(v.sp, in1, in3, pm, hm, nmu, X, dm);
  return 0;
}

// mask bits: 1 = S, 2 = T, 4 = dipole_z, 8 = vna (ontop into blocko, atom into
// block(neigh_self), zero non-self block for the 3c accumulation)
template <int MO>
__global__ __launch_bounds__(Gen2c<MO>::W) void k_2c(DView v, int mask) {
  constexpr int W2C = Gen2c<MO>::W, NB = MO * MO;
  // atom terms of the chunk's neighbours: [0] neutral, [isorp] charged shell isorp
  __shared__ double s_atom[1 + FB_MAXSH][W2C][NB + 1];
  const int ia = blockIdx.x;
  const int lane = threadIdx.x;
  const int s = v.atom_struct[ia];
  const DStruct St = v.st[s];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  const int in1 = v.imass[ia];
  const int nmu = v.sp.norb[in1 - 1];
  const int n = v.neighn[ia];
  const int64_t p0 = v.noff[ia];
  const double r1[3] = {v.ratom[3 * ia], v.ratom[3 * ia + 1], v.ratom[3 * ia + 2]};
  const int64_t pself = p0 + v.neigh_self[ia] - 1;
  int status = 0;
  // lane owns elements lane + j W2C (j < NE) of the self block
  constexpr int NE = (NB + W2C - 1) / W2C;
  double acc_atom[NE];
  for (int j = 0; j < NE; j++) acc_atom[j] = 0.0;
  const double rc1min = v.sp.rcmin[in1 - 1];

  for (int c0 = 0; c0 < n; c0 += W2C) {
    const int k = c0 + lane;
    const bool active = k < n;
    double atom_c[NB];
    for (int i = 0; i < NB; i++) atom_c[i] = 0.0;
    for (int q = 1; q <= FB_MAXSH; q++)
      for (int i = 0; i < NB; i++) s_atom[q][lane][i] = 0.0;
    if (active) {
      const int64_t p = p0 + k;
      const int ja = St.atom_off + v.neigh_j[p] - 1;
      const int mb = v.neigh_b[p];
      const int in2 = v.imass[ja];
      const int nnu = v.sp.norb[in2 - 1];
      const int nb = nmu * nnu;
      double r2[3];
      for (int q = 0; q < 3; q++) r2[q] = v.ratom[3 * ja + q] + xls[3 * mb + q];
      double z = d_distance(r1, r2), sighat[3], eps[9], pm[9], dm[MO == FB_MOD ? 25 : 1];
      if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
      else for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / z;
      d_epsilon(r2, sighat, eps);
      d_pmat<double>(eps, pm);
      if (MO == FB_MOD) d_dmat(eps, dm);
      const int64_t bo = v.boff[p];
      double X[NB];
      if (mask & 1) {
        status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_S, 0), in1, in2, nmu, z, pm, dm, X);
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_S][bo + i] = X[i];
      }
      if (mask & 2) {
        status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_T, 0), in1, in2, nmu, z, pm, dm, X);
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_T][bo + i] = X[i];
      }
      if (mask & 4) {
        status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_DIP, 0), in1, in2, nmu, z, pm, dm, X);
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_DIP][bo + i] = X[i];
      }
      if (mask & 8) {
        // ontop L and R into blocko (M_assemble_2c_DOGS.f90:640-702)
        double acc[NB];
        for (int i = 0; i < nb; i++) acc[i] = 0.0;
        const bool self = (v.neigh_j[p] == ia - St.atom_off + 1 && mb == 0);
        if (!self) {
          status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_ONL, 0), in1, in2, nmu, z, pm, dm, X);
          for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i] * FB_P_EQ2;
          for (int isorp = 1; isorp <= v.sp.nssh[in1 - 1]; isorp++) {
            double dQ = v.dQ[ia * FB_MAXSH + isorp - 1];
            if (dQ == 0.0) continue;
            status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_ONL, isorp), in1, in2, nmu, z, pm, dm, X);
            for (int i = 0; i < nb; i++) acc[i] = acc[i] + dQ * X[i] * FB_P_EQ2;
          }
          status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_ONR, 0), in1, in2, nmu, z, pm, dm, X);
          for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i] * FB_P_EQ2;
          for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) {
            double dQ = v.dQ[ja * FB_MAXSH + isorp - 1];
            if (dQ == 0.0) continue;
            status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_ONR, isorp), in1, in2, nmu, z, pm, dm, X);
            for (int i = 0; i < nb; i++) acc[i] = acc[i] + dQ * X[i] * FB_P_EQ2;
          }
        }
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_VNAO][bo + i] = acc[i];
        if (p != pself && !((mask & FB_2C_SKIP_TGT) && v.pair_tgt[p]))
          for (int i = 0; i < nb; i++) v.mat[FB_MAT_VNA][bo + i] = 0.0;
        // atom term <i|V_j|i> (M_assemble_2c_DOGS.f90:763-812)
        double rend = rc1min + v.sp.rcmin[in2 - 1];
        double smooth = d_smoother(z, rend, 0.8);
        status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_ATOM, 0), in1, in1, nmu, z, pm, dm, X);
        const int na = nmu * nmu;
        for (int i = 0; i < na; i++) atom_c[i] = X[i] * FB_P_EQ2;
        bool charged = false;
        for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++)
          charged |= (v.dQ[ja * FB_MAXSH + isorp - 1] != 0.0);
        if (charged) {
          // the charged terms use the self overlap block, written by an
          // earlier launch (assemble_S precedes assemble_vna_2c)
          const double *sself = v.mat[FB_MAT_S] + v.boff[pself];
          for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) {
            double dQ = v.dQ[ja * FB_MAXSH + isorp - 1];
            status |= me2c_rotated<MO>(v, cell2c_id(v, in1, in2, SLOT_ATOM, isorp), in1, in1, nmu, z, pm, dm, X);
            for (int i = 0; i < na; i++) {
              double emnpl = (z > 1.0e-4) ? sself[i] / z : 0.0;
              s_atom[isorp][lane][i] = FB_P_EQ2 * dQ * ((1.0 - smooth) * emnpl + smooth * X[i]);
            }
          }
        }
      }
    }
    if (mask & 8) {
      // ordered sum over this chunk's neighbours into the self block, each
      // neighbour's terms one at a time (an absent charged term is an exact zero)
      for (int i = 0; i < NB; i++) s_atom[0][lane][i] = atom_c[i];
      __syncthreads();
      const int kmax = min(W2C, n - c0);
      const int nq = v.any_dq ? FB_MAXSH : 0;
      for (int j = 0; j < NE; j++) {
        const int e = lane + j * W2C;
        if (e >= nmu * nmu) break;
        for (int kk = 0; kk < kmax; kk++) {
          acc_atom[j] = acc_atom[j] + s_atom[0][kk][e];
          for (int q = 1; q <= nq; q++) acc_atom[j] = acc_atom[j] + s_atom[q][kk][e];
        }
      }
      __syncthreads();
    }
  }
  if (mask & 8)
    for (int j = 0; j < NE; j++) {
      const int e = lane + j * W2C;
      if (e < nmu * nmu) v.mat[FB_MAT_VNA][v.boff[pself] + e] = acc_atom[j];
    }
  if (status) atomicOr(v.status, status);
}

extern "C" int fbk_2c(const DView &v, int mask, hipStream_t st) {
  if (v.natoms == 0) return 0;
  if (fb_wide_species(v.sp))
    hipLaunchKernelGGL(k_2c<FB_MOD>, dim3(v.natoms), dim3(Gen2c<FB_MOD>::W), 0, st, v, mask);
  else
    hipLaunchKernelGGL(k_2c<FB_MO>, dim3(v.natoms), dim3(Gen2c<FB_MO>::W), 0, st, v, mask);
  return (int)hipGetLastError();
}

// ---------------------------------------------------------------------------
// s/p species with canonical tables (fb_2c_sp.h)
template <int KQ>
__global__ __launch_bounds__(WG2, 3) void k_2c_sp(DView v, int mask) {
  __shared__ S2c<KQ> S;
  wg2c_sp<KQ>(v, mask, blockIdx.x, S);
}

extern "C" int fbk_2c_sp(const DView &v, int mask, hipStream_t st) {
  if (v.nwg2c == 0) return 0;
  if (v.any_dq && (mask & 8))
    hipLaunchKernelGGL(k_2c_sp<2>, dim3(v.nwg2c), dim3(WG2), 0, st, v, mask);
  else
    hipLaunchKernelGGL(k_2c_sp<0>, dim3(v.nwg2c), dim3(WG2), 0, st, v, mask);
  return (int)hipGetLastError();
}


// bond geometry of pair p (M_assemble_2c_DOGS.f90:165-177)
struct Pair2c {
  int ia, ja, in1, in2, cls;
  bool self;
  double z;
  double pm[9];
  int64_t bo;
};

__device__ __forceinline__ void pair_geometry(const DView &v, int64_t p, Pair2c &g) {
  const int ia = v.pair_atom[p];
  const int s = v.atom_struct[ia];
  const int ja = v.st[s].atom_off + v.neigh_j[p] - 1;
  const int mb = v.neigh_b[p];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  double r1[3], r2[3], sighat[3], eps[9];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    r1[q] = v.ratom[3 * ia + q];
    r2[q] = v.ratom[3 * ja + q] + xls[3 * mb + q];
  }
  g.z = d_distance(r1, r2);
  if (g.z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else {
#pragma unroll
    for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / g.z;
  }
  d_epsilon(r2, sighat, eps);
  d_pmat<double>(eps, g.pm);
  g.ia = ia;
  g.ja = ja;
  g.in1 = v.imass[ia];
  g.in2 = v.imass[ja];
  g.self = (ja == ia && mb == 0);
  g.cls = (v.sp.norb[g.in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[g.in2 - 1] == 4 ? 1 : 0);
  g.bo = v.boff[p];
}

// ---------------------------------------------------------------------------
// LDS-staged variant for s/p species (the default): the workgroup walks the
// interactions in the reference's order (S, T, dipole_z, ontop L with its
// charge subtypes, ontop R with its subtypes, atom with its subtypes).  For
// each it stages every distinct table cell its lanes need in LDS with
// contiguous loads and interpolates from there, so the per-lane gathers hit
// LDS instead of the vector memory pipeline.  Lagrange reciprocals are formed
// in-kernel with the reference's division.  Each lane writes its block into
// an LDS row and the workgroup stores the contiguous block range of its pairs
// with consecutive lanes on consecutive addresses.  Arithmetic and its order
// per pair are those of k_2c_sp, so the blocks are the same bits.
#define CELLCAP 4096   // doubles of one staged cell (nx * nME)
#define ROWP 17        // doubles per LDS block row (odd: conflict-free lane rows)

template <int CLS>
__device__ __forceinline__ void interp_rot(const double *base, const double *prod, const double *pm, double *X) {
  constexpr int NME = nme2c_of(CLS);
  double F[NME];
#pragma unroll
  for (int m = 0; m < NME; m++) F[m] = 0.0;
#pragma unroll
  for (int is = 0; is < 7; is++)
#pragma unroll
    for (int m = 0; m < NME; m++) F[m] = F[m] + base[is * NME + m] * prod[is];
  rot2c_sp<double>(CLS, pm, F, X);
}

// One interaction slot: every lane of the workgroup calls it; a lane with
// cell my >= 0 gets its rotated block in X (zeros otherwise).  The loop stages
// one distinct cell per pass; the cell id is uniform, so its header loads are
// scalar.
__device__ inline int slot2c(const DView &v, int my, int cls, const Pair2c &g, double *s_cell, int *s_sel,
                             int &grid_c, int &ileft, double *prod, double *X) {
  int status = 0;
  bool todo = my >= 0;
#pragma unroll
  for (int i = 0; i < 16; i++) X[i] = 0.0;
  while (true) {
    __syncthreads();
    if (threadIdx.x == 0) *s_sel = -1;
    __syncthreads();
    if (todo) *s_sel = my;   // any one of the competing writes wins
    __syncthreads();
    const int c = *s_sel;
    if (c < 0) break;
    const DCell2c cc = v.cell2c[c];
    const int nv = cc.nx * cc.nME;
    const bool lds = nv <= CELLCAP;
    if (lds)
      for (int i = threadIdx.x; i < nv; i += WG2) s_cell[i] = v.tab2c[cc.off + i];
    __syncthreads();
    if (todo && my == c) {
      todo = false;
      const double x = g.z;
      if (x < 0.0) {
        status |= ST_X_NEG;
      } else if (x > cc.xmax * (1.0 + 1.0e-5)) {
        status |= ST_X_RANGE;
      } else {
        if (cc.grid != grid_c) {
          const DGrid2c gr = v.grid2c[cc.grid];
          ileft = window7(x, gr.dx, gr.nx);
          lagrange7_div(x, gr.dx, ileft, prod);
          grid_c = cc.grid;
        }
        // separate branches keep the LDS pointer in its address space (ds_read
        // rather than flat loads)
        if (lds) {
          const double *base = s_cell + (ileft - 1) * cc.nME;
          switch (cls) {
            case 3: interp_rot<3>(base, prod, g.pm, X); break;
            case 2: interp_rot<2>(base, prod, g.pm, X); break;
            case 1: interp_rot<1>(base, prod, g.pm, X); break;
            default: interp_rot<0>(base, prod, g.pm, X); break;
          }
        } else {
          const double *base = v.tab2c + cc.off + (int64_t)(ileft - 1) * cc.nME;
          switch (cls) {
            case 3: interp_rot<3>(base, prod, g.pm, X); break;
            case 2: interp_rot<2>(base, prod, g.pm, X); break;
            case 1: interp_rot<1>(base, prod, g.pm, X); break;
            default: interp_rot<0>(base, prod, g.pm, X); break;
          }
        }
      }
    }
  }
  return status;
}

// contiguous store of the chunk's block range [rb0, rb0 + len) from the LDS rows
__device__ __forceinline__ void store_rows(double *dst, int len, const double *s_row, const unsigned short *s_map) {
  for (int j = threadIdx.x; j < len; j += WG2) dst[j] = s_row[s_map[j]];
}

__global__ __launch_bounds__(WG2) void k_2c_lds(DView v, int mask) {
  __shared__ double s_cell[CELLCAP];
  __shared__ double s_row[WG2 * ROWP];          // lane blocks, then atom terms for the reduction
  __shared__ unsigned short s_map[WG2 * 16];    // block range element -> LDS row element
  __shared__ int s_sel;
  const int tid = threadIdx.x;
  const int a0 = v.wg2c[blockIdx.x], a1 = v.wg2c[blockIdx.x + 1];
  const int64_t pb = v.noff[a0], pe = v.noff[a1];
  // reduction slot of this thread: (atom a0 + tid/16, element tid%16)
  const int ra = a0 + (tid >> 4), re = tid & 15;
  const bool rown = ra < a1;
  int64_t rp0 = 0, rp1 = 0;
  int rna = 0;
  if (rown) {
    rp0 = v.noff[ra];
    rp1 = v.noff[ra + 1];
    rna = v.sp.norb[v.imass[ra] - 1];
  }
  int maxsh = 0;
  for (int s = 0; s < v.sp.ns; s++) maxsh = max(maxsh, v.sp.nssh[s]);
  double racc = 0.0;
  int status = 0;
  for (int64_t c0 = pb; c0 < pe; c0 += WG2) {
    const int nt = (int)min((int64_t)WG2, pe - c0);
    const bool act = tid < nt;
    Pair2c g{};
    int nb = 0, ro = 0;
    const int64_t rb0 = v.boff[c0];
    const int len = (int)(v.boff[c0 + nt] - rb0);
    if (act) {
      pair_geometry(v, c0 + tid, g);
      nb = ((g.cls & 2) ? 4 : 1) * ((g.cls & 1) ? 4 : 1);
      ro = (int)(g.bo - rb0);
      for (int i = 0; i < nb; i++) s_map[ro + i] = (unsigned short)(tid * ROWP + i);
    }
    int grid_c = -1, ileft = 1;
    double prod[7], X[16], acc[16];
    bool charged = false;
    double smooth = 0.0;
    const double *sself = nullptr;
    if (act && (mask & 8)) {
      for (int isorp = 1; isorp <= v.sp.nssh[g.in2 - 1]; isorp++)
        charged |= (v.dQ[g.ja * FB_MAXSH + isorp - 1] != 0.0);
      if (charged) {
        const int64_t pself = v.noff[g.ia] + v.neigh_self[g.ia] - 1;
        sself = v.mat[FB_MAT_S] + v.boff[pself];
        smooth = d_smoother(g.z, v.sp.rcmin[g.in1 - 1] + v.sp.rcmin[g.in2 - 1], 0.8);
      }
    }
    // the slot sequence, uniform over the workgroup: S, T, dipole_z, then
    // ontop L and R with their charge subtypes, then the atom term with its
    // subtypes (per pair, the reference's order of accumulation)
    const int nsl = (maxsh + 1);
    const int nslots = 3 + ((mask & 8) ? 3 * nsl : 0);
    for (int si = 0; si < nslots; si++) {
      const int kind = si < 3 ? si : 3 + (si - 3) / nsl;       // 0 S 1 T 2 DIP 3 ONL 4 ONR 5 ATOM
      const int isorp = si < 3 ? 0 : (si - 3) % nsl;
      if (kind < 3 && !(mask & (1 << kind))) continue;
      int my = -1, cls = g.cls;
      double dQ = 1.0;
      if (act) {
        if (kind < 3) {
          my = cell2c_id(v, g.in1, g.in2, kind == 0 ? SLOT_S : (kind == 1 ? SLOT_T : SLOT_DIP), 0);
          if (my < 0) status |= ST_NO_TABLE;
        } else if (kind < 5) {
          if (!g.self) {
            const int inq = kind == 3 ? g.in1 : g.in2, aq = kind == 3 ? g.ia : g.ja;
            bool need = isorp == 0;
            if (isorp > 0 && isorp <= v.sp.nssh[inq - 1]) {
              dQ = v.dQ[aq * FB_MAXSH + isorp - 1];
              need = dQ != 0.0;
            }
            if (need) {
              my = cell2c_id(v, g.in1, g.in2, kind == 3 ? SLOT_ONL : SLOT_ONR, isorp);
              if (my < 0) status |= ST_NO_TABLE;
            }
          }
        } else if (isorp == 0 || (charged && isorp <= v.sp.nssh[g.in2 - 1])) {
          my = cell2c_id(v, g.in1, g.in2, SLOT_ATOM, isorp);
          if (my < 0) status |= ST_NO_TABLE;
          if (isorp > 0) dQ = v.dQ[g.ja * FB_MAXSH + isorp - 1];
          cls = (g.cls & 2) ? CLS_SPSP : CLS_SS;   // the atom term lives on iatom's orbitals
        }
      }
      if (si == 3) {
#pragma unroll
        for (int i = 0; i < 16; i++) acc[i] = 0.0;
      }
      status |= slot2c(v, my, cls, g, s_cell, &s_sel, grid_c, ileft, prod, X);
      if (kind < 3) {
        if (act)
          for (int i = 0; i < nb; i++) s_row[tid * ROWP + i] = X[i];
        __syncthreads();
        store_rows(v.mat[kind == 0 ? FB_MAT_S : (kind == 1 ? FB_MAT_T : FB_MAT_DIP)] + rb0, len, s_row, s_map);
      } else if (kind < 5) {
        if (my >= 0) {
          if (isorp == 0) {
#pragma unroll
            for (int i = 0; i < 16; i++) acc[i] = acc[i] + X[i] * FB_P_EQ2;
          } else {
#pragma unroll
            for (int i = 0; i < 16; i++) acc[i] = acc[i] + dQ * X[i] * FB_P_EQ2;
          }
        }
        if (si == 3 + 2 * nsl - 1) {
          // blocko complete (M_assemble_2c_DOGS.f90:640-702); the Vna block is
          // zero here, its self blocks receive the atom sums below
          __syncthreads();
          if (act)
            for (int i = 0; i < nb; i++) s_row[tid * ROWP + i] = acc[i];
          __syncthreads();
          store_rows(v.mat[FB_MAT_VNAO] + rb0, len, s_row, s_map);
          if (mask & FB_2C_SKIP_TGT) {
            if (act && !g.self && !v.pair_tgt[c0 + tid])
              for (int i = 0; i < nb; i++) v.mat[FB_MAT_VNA][g.bo + i] = 0.0;
          } else {
            for (int j = tid; j < len; j += WG2) v.mat[FB_MAT_VNA][rb0 + j] = 0.0;
          }
#pragma unroll
          for (int i = 0; i < 16; i++) acc[i] = 0.0;   // reused for the atom term
        }
      } else if (my >= 0) {
        // atom term <i|V_j|i> (M_assemble_2c_DOGS.f90:710-816)
        const int na = (g.cls & 2) ? 16 : 1;
        if (isorp == 0) {
          for (int i = 0; i < na; i++) acc[i] = X[i] * FB_P_EQ2;
        } else {
          for (int i = 0; i < na; i++) {
            double emnpl = (g.z > 1.0e-4) ? sself[i] / g.z : 0.0;
            acc[i] = acc[i] + FB_P_EQ2 * dQ * ((1.0 - smooth) * emnpl + smooth * X[i]);
          }
        }
      }
    }
    if (mask & 8) {
      // ordered sum over this chunk's neighbours into the self block
      __syncthreads();
#pragma unroll
      for (int i = 0; i < 16; i++) s_row[tid * ROWP + i] = act ? acc[i] : 0.0;
      __syncthreads();
      if (rown && re < rna * rna) {
        const int64_t k0 = max(rp0, c0), k1 = min(rp1, c0 + (int64_t)nt);
        for (int64_t k = k0; k < k1; k++) racc = racc + s_row[(int)(k - c0) * ROWP + re];
      }
    }
    __syncthreads();
  }
  if ((mask & 8) && rown && re < rna * rna) {
    const int64_t pself = v.noff[ra] + v.neigh_self[ra] - 1;
    v.mat[FB_MAT_VNA][v.boff[pself] + re] = racc;
  }
  if (status) atomicOr(v.status, status);
}

extern "C" int fbk_2c_lds(const DView &v, int mask, hipStream_t st) {
  if (v.nwg2c == 0) return 0;
  // this variant sums a neighbour's charged atom terms before adding them to
  // the self block; charged Vna builds take the s/p kernel, which adds each
  // term on its own as the reference does
  if (v.any_dq && (mask & 8)) return fbk_2c_sp(v, mask, st);
  hipLaunchKernelGGL(k_2c_lds, dim3(v.nwg2c), dim3(WG2), 0, st, v, mask);
  return (int)hipGetLastError();
}
