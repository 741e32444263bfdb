// Two-centre assembly for s/p species with canonical tables (S, T, dipole_z,
// Vna ontop and atom terms, j.ASSEMBLERS/M_assemble_2c_DOGS.f90:87-829): one
// lane per pair, workgroups of WG2 threads packed by whole atoms (at most 16
// atoms and 256 pairs, or one larger atom), Lagrange reciprocals formed
// in-kernel with the reference's division, rotations from fb_sp.h.
// Bit-identical to the reference in binary64.  Shared by the 2c kernel and
// the fused H+S kernel.
#pragma once

#include "fb_device.h"
#include "fb_sp.h"
#include "fb_2c_dev.h"

#define WG2 256
static_assert(WG2 == FB_2C_PCAP, "the 2c lane order (k_pair_perm) is laid out for passes of FB_2C_PCAP pairs");

__device__ __forceinline__ int cell2c_id(const DView &v, int in1, int in2, int slot, int isub) {
  return v.idx2c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * FB_NSLOT2 + slot) * FB_MAXSUB2 + isub];
}

// rotated block of one interaction; returns status bits
template <int CLS>
__device__ __forceinline__ int block2c(const DView &v, int sp2, int slot, int isub, double x, const double *pm,
                                       int &grid_cached, int &ileft, double *prod, double *X) {
  DCell2c c;
  if (cell_scalar(v, sp2, slot, isub, c) < 0) return ST_NO_TABLE;
  if (x < 0.0) return ST_X_NEG;
  if (x > c.xmax * (1.0 + 1.0e-5)) return ST_X_RANGE;
  if (c.grid != grid_cached) {
    const DGrid2c gr = grid_scalar(v, c.grid);
    ileft = window7(x, gr.dx, gr.nx);
    lagrange7_div(x, gr.dx, ileft, prod);
    grid_cached = c.grid;
  }
  constexpr int NME = nme2c_of(CLS);
  double F[NME];
  interp_cell<NME>(v, c, ileft, prod, F);
  rot2c_sp<double>(CLS, pm, F, X);
  return 0;
}

// Block stores of a wavefront whose active lanes all hold sp-sp pairs (16
// doubles each, so the wavefront's blocks are one contiguous range starting
// at boff of its first pair): each lane writes its block to an LDS row of 17
// doubles, and the lanes then store the range with consecutive lanes on
// consecutive doubles.  A lane storing its own 128-byte block instead touches
// a separate cache line per lane and element group.  Other wavefronts store
// per lane.
struct WStore {
  bool uni;        // every active lane of the wavefront holds an sp-sp pair
  int lane, nact;  // lane, active lanes (a prefix of the wavefront)
  int row;         // LDS row of the lane's pair: its index within the wavefront's group
  int64_t wb0;     // block offset of the wavefront's first pair
  double *ws;      // the wavefront's 64 LDS rows
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// the block already in the lane's LDS row
template <int NBL>
__device__ __forceinline__ void put_row(double *m, int64_t bo, const WStore &W) {
  const double *row = W.ws + W.row * 17;
  if (NBL == 16 && W.uni) {
    wave_sync();
    double *dst = m + W.wb0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int o = k * W.nact + W.lane;
      dst[o] = W.ws[(o >> 4) * 17 + (o & 15)];
    }
    wave_sync();
  } else {
#pragma unroll
    for (int i = 0; i < NBL; i++) m[bo + i] = row[i];
  }
}

template <int NBL>
__device__ __forceinline__ void put_block(double *m, int64_t bo, const double *X, const WStore &W) {
  if (NBL == 16 && W.uni) {
    double *row = W.ws + W.row * 17;
#pragma unroll
    for (int i = 0; i < 16; i++) row[i] = X[i];
    wave_sync();
    double *dst = m + W.wb0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      const int o = k * W.nact + W.lane;
      dst[o] = W.ws[(o >> 4) * 17 + (o & 15)];
    }
    wave_sync();
  } else {
#pragma unroll
    for (int i = 0; i < NBL; i++) m[bo + i] = X[i];
  }
}

// atom_c receives the neutral atom term; with KQ > 0 the charged terms of
// shells 1..KQ go to qrow[(isorp - 1) * QROW + i] (LDS rows that the caller
// cleared), so that the self-block sum can add each term on its own
#define QROW (WG2 * 17)
template <int CLS, int KQ>
__device__ __forceinline__ int pair2c(const DView &v, int mask, int ia, int ja, int in1, int in2, bool self,
                                      bool keep_vna, double z, const double *pm, int64_t bo, double *atom_c,
                                      double *qrow, const WStore &W) {
  constexpr int NMU = (CLS & 2) ? 4 : 1, NNU = (CLS & 1) ? 4 : 1, NBL = NMU * NNU;
  constexpr int CLSA = (CLS & 2) ? CLS_SPSP : CLS_SS;   // atom term lives on iatom's orbitals
  int status = 0, grid_cached = -1, ileft = 1;
  double prod[7], X[16];
  const int sp2 = in1 | (in2 << 16);
  if (mask & 1) {
    status |= block2c<CLS>(v, sp2, SLOT_S, 0, z, pm, grid_cached, ileft, prod, X);
    put_block<NBL>(v.mat[FB_MAT_S], bo, X, W);
  }
  if (mask & 2) {
    status |= block2c<CLS>(v, sp2, SLOT_T, 0, z, pm, grid_cached, ileft, prod, X);
    put_block<NBL>(v.mat[FB_MAT_T], bo, X, W);
  }
  if (mask & 4) {
    status |= block2c<CLS>(v, sp2, SLOT_DIP, 0, z, pm, grid_cached, ileft, prod, X);
    put_block<NBL>(v.mat[FB_MAT_DIP], bo, X, W);
  }
  if (mask & 8) {
    // the ontop sum accumulates in the pair's LDS row
    double *acc = W.ws + W.row * 17;
#pragma unroll
    for (int i = 0; i < NBL; i++) acc[i] = 0.0;
    if (!self) {
      status |= block2c<CLS>(v, sp2, SLOT_ONL, 0, z, pm, grid_cached, ileft, prod, X);
#pragma unroll
      for (int i = 0; i < NBL; i++) acc[i] = acc[i] + X[i] * FB_P_EQ2;
      // (no charged shell in the batch: the subtype loops add nothing)
      for (int isorp = 1; v.any_dq && isorp <= v.sp.nssh[in1 - 1]; isorp++) {
        double dQ = v.dQ[ia * FB_MAXSH + isorp - 1];
        if (dQ == 0.0) continue;
        status |= block2c<CLS>(v, sp2, SLOT_ONL, isorp, z, pm, grid_cached, ileft, prod, X);
#pragma unroll
        for (int i = 0; i < NBL; i++) acc[i] = acc[i] + dQ * X[i] * FB_P_EQ2;
      }
      status |= block2c<CLS>(v, sp2, SLOT_ONR, 0, z, pm, grid_cached, ileft, prod, X);
#pragma unroll
      for (int i = 0; i < NBL; i++) acc[i] = acc[i] + X[i] * FB_P_EQ2;
      for (int isorp = 1; v.any_dq && isorp <= v.sp.nssh[in2 - 1]; isorp++) {
        double dQ = v.dQ[ja * FB_MAXSH + isorp - 1];
        if (dQ == 0.0) continue;
        status |= block2c<CLS>(v, sp2, SLOT_ONR, isorp, z, pm, grid_cached, ileft, prod, X);
#pragma unroll
        for (int i = 0; i < NBL; i++) acc[i] = acc[i] + dQ * X[i] * FB_P_EQ2;
      }
    }
    const bool clr = !self && !keep_vna;
    acc[16] = clr ? 1.0 : 0.0;   // the row's pad element carries the pair's flag
    put_row<NBL>(v.mat[FB_MAT_VNAO], bo, W);
    if (NBL == 16 && W.uni) {
      double *dst = v.mat[FB_MAT_VNA] + W.wb0;
#pragma unroll
      for (int k = 0; k < 16; k++) {
        const int o = k * W.nact + W.lane;
        if (W.ws[(o >> 4) * 17 + 16] != 0.0) dst[o] = 0.0;
      }
    } else if (clr) {
#pragma unroll
      for (int i = 0; i < NBL; i++) v.mat[FB_MAT_VNA][bo + i] = 0.0;
    }
    // atom term <i|V_j|i>
    constexpr int NA = NMU * NMU;
    status |= block2c<CLSA>(v, sp2, SLOT_ATOM, 0, z, pm, grid_cached, ileft, prod, X);
#pragma unroll
    for (int i = 0; i < 16; i++) atom_c[i] = i < NA ? X[i] * FB_P_EQ2 : 0.0;
    bool charged = false;
    for (int isorp = 1; v.any_dq && isorp <= v.sp.nssh[in2 - 1]; isorp++)
      charged |= (v.dQ[ja * FB_MAXSH + isorp - 1] != 0.0);
    if (charged) {
      const int64_t pself = v.noff[ia] + v.neigh_self[ia] - 1;
      const double *sself = v.mat[FB_MAT_S] + v.boff[pself];
      const double smooth = d_smoother(z, v.sp.rcmin[in1 - 1] + v.sp.rcmin[in2 - 1], 0.8);
      for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) {
        double dQ = v.dQ[ja * FB_MAXSH + isorp - 1];
        status |= block2c<CLSA>(v, sp2, SLOT_ATOM, isorp, z, pm, grid_cached, ileft, prod, X);
#pragma unroll
        for (int i = 0; i < NA; i++) {
          double emnpl = (z > 1.0e-4) ? sself[i] / z : 0.0;
          const double t = FB_P_EQ2 * dQ * ((1.0 - smooth) * emnpl + smooth * X[i]);
          if (KQ > 0 && isorp <= KQ) qrow[(isorp - 1) * QROW + i] = t;
          else atom_c[i] = atom_c[i] + t;
        }
      }
    }
  }
  return status;
}

// LDS of one workgroup item: the atom terms of its pairs, [0] neutral and
// [1..KQ] the charged shells.  Before the atom terms are written, the rows of
// [0] serve each wavefront's block stores (WStore).
template <int KQ>
struct S2c {
  double s_at[1 + KQ][WG2][17];
};

// KQ: charged atom-term rows kept apart (0 for neutral builds, 2 = the most
// shells of an s/p species, for charged ones).  item: index into v.wg2c;
// every thread of the WG2-thread workgroup calls it.
template <int KQ>
__device__ __forceinline__ void wg2c_sp(const DView &v, int mask, int item, S2c<KQ> &S) {
  auto &s_at = S.s_at;
  const int tid = threadIdx.x;
  const int a0 = v.wg2c[item], a1 = v.wg2c[item + 1];
  const int64_t pb = v.noff[a0], pe = v.noff[a1];
  // reduction slot of this thread: (atom a0 + tid/16, element tid%16)
  const int ra = a0 + (tid >> 4), re = tid & 15;
  const bool rown = ra < a1;
  double racc = 0.0;
  int status = 0;
  WStore W;
  W.lane = tid & 63;
  W.ws = &s_at[0][0][0] + __builtin_amdgcn_readfirstlane(tid & ~63) * 17;
  for (int64_t c0 = pb; c0 < pe; c0 += WG2) {
    // the lane at slot c0 + tid takes pair p of its wavefront's group (in
    // the order of v.pair_perm); blocks, LDS rows and sums go by p
    const int64_t slot = c0 + tid;
    const int64_t p0 = c0 + __builtin_amdgcn_readfirstlane(tid & ~63);
    const int64_t p = (slot < pe && v.pair_perm) ? p0 + v.pair_perm[slot] : slot;
    const int srow = (int)(p - c0);
    double atom_c[16];
#pragma unroll
    for (int i = 0; i < 16; i++) atom_c[i] = 0.0;
    if (KQ > 0 && (mask & 8))
      for (int q = 1; q <= KQ; q++)
#pragma unroll
        for (int i = 0; i < 16; i++) s_at[q][tid][i] = 0.0;
    if (slot < pe) {
      const int4 pr = v.pair_rec[p];
      const int ia = pr.x, ja = pr.y;
      const int in1 = pr.w & 0xffff, in2 = pr.w >> 16;
      const double *xlp = v.xl + (int64_t)pr.z * 3;   // xl(:, mbeta) of the pair's structure
      double r1[3], r2[3], sighat[3], eps[9], pm[9];
#pragma unroll
      for (int q = 0; q < 3; q++) {
        r1[q] = v.ratom[3 * ia + q];
        r2[q] = v.ratom[3 * ja + q] + xlp[q];
      }
      const double z = d_distance(r1, r2);
      if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
      else {
#pragma unroll
        for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / z;
      }
      d_epsilon(r2, sighat, eps);
      d_pmat<double>(eps, pm);
      const bool self = (ja == ia && pr.z % FB_NIMG == 0);   // jatom = iatom in the home cell
      const int cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
      const int64_t bo = v.boff[p];
      // the active lanes are a prefix of the wavefront
      W.nact = (int)min((int64_t)64, pe - p0);
      W.row = (int)(p - p0);
      W.uni = __ballot(cls != 3) == 0;
      const int64_t wb0 = v.boff[p0];
      W.wb0 = ((int64_t)__builtin_amdgcn_readfirstlane((int)(wb0 >> 32)) << 32) |
              (uint32_t)__builtin_amdgcn_readfirstlane((int)wb0);
      const bool keep = (mask & FB_2C_SKIP_TGT) && v.pair_tgt[p];
      double *qrow = KQ > 0 ? &s_at[KQ > 0 ? 1 : 0][srow][0] : nullptr;
      switch (cls) {
        case 3: status |= pair2c<3, KQ>(v, mask, ia, ja, in1, in2, self, keep, z, pm, bo, atom_c, qrow, W); break;
        case 2: status |= pair2c<2, KQ>(v, mask, ia, ja, in1, in2, self, keep, z, pm, bo, atom_c, qrow, W); break;
        case 1: status |= pair2c<1, KQ>(v, mask, ia, ja, in1, in2, self, keep, z, pm, bo, atom_c, qrow, W); break;
        default: status |= pair2c<0, KQ>(v, mask, ia, ja, in1, in2, self, keep, z, pm, bo, atom_c, qrow, W); break;
      }
    }
    if (mask & 8) {
#pragma unroll
      for (int i = 0; i < 16; i++) s_at[0][srow][i] = atom_c[i];
      __syncthreads();
      const int rna = rown ? v.sp.norb[v.imass[ra] - 1] : 0;
      if (rown && re < rna * rna) {
        // each neighbour's terms one at a time, as the reference adds them
        const int64_t k0 = max(v.noff[ra], c0), k1 = min(v.noff[ra + 1], c0 + WG2);
        for (int64_t k = k0; k < k1; k++) {
          racc = racc + s_at[0][k - c0][re];
#pragma unroll
          for (int q = 1; q <= KQ; q++) racc = racc + s_at[q][k - c0][re];
        }
      }
      __syncthreads();
    }
  }
  if ((mask & 8) && rown && re < v.sp.norb[v.imass[ra] - 1] * v.sp.norb[v.imass[ra] - 1]) {
    const int64_t pself = v.noff[ra] + v.neigh_self[ra] - 1;
    v.mat[FB_MAT_VNA][v.boff[pself] + re] = racc;
  }
  if (status) atomicOr(v.status, status);
}
