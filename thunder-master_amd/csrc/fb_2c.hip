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
// the sum is deterministic and in the reference's order.
//
// The kernel is bound by the block stores (HBM); its arithmetic stays binary64.
#include "fb_device.h"

#define W2C 64
#define NB (FB_MO * FB_MO)
#define NME2 16   // max non-zero 2c matrix elements for s/p species

__device__ __forceinline__ int cell2c_id(const DView &v, int in1, int in2, int slot, int isub) {
  return v.idx2c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * FB_NSLOT2 + slot) * FB_MAXSUB2 + isub];
}

// interpolate + recover + rotate one interaction into X (column-major nmu x nnu)
__device__ inline int me2c_rotated(const DView &v, int cid, int in1, int in3, int nmu, double x, const double *pm,
                                   double *X) {
  if (cid < 0) return ST_NO_TABLE;
  const DCell2c c = v.cell2c[cid];
  double F[NME2];
  int st = d_interp2c(c, v.tab2c, x, F);
  if (st) return st;
  double hm[NB];
  int nnu = v.sp.norb[in3 - 1];
  for (int i = 0; i < nmu * nnu; i++) hm[i] = 0.0;
  for (int k = 0; k < c.nME; k++) hm[(v.nu[c.munu + k] - 1) * nmu + (v.mu[c.munu + k] - 1)] = F[k];
  d_rotate<double>(v.sp, in1, in3, pm, hm, nmu, X);
  return 0;
}

// mask bits: 1 = S, 2 = T, 4 = dipole_z, 8 = vna (ontop into blocko, atom into
// block(neigh_self), zero non-self block for the 3c accumulation)
__global__ __launch_bounds__(W2C) void k_2c(DView v, int mask) {
  __shared__ double s_atom[W2C][NB + 1];
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
  double acc_atom = 0.0;  // lane e < nmu*nmu owns element e of the self block
  const double rc1min = v.sp.rcmin[in1 - 1];

  for (int c0 = 0; c0 < n; c0 += W2C) {
    const int k = c0 + lane;
    const bool active = k < n;
    double atom_c[NB];
    for (int i = 0; i < NB; i++) atom_c[i] = 0.0;
    if (active) {
      const int64_t p = p0 + k;
      const int ja = St.atom_off + v.neigh_j[p] - 1;
      const int mb = v.neigh_b[p];
      const int in2 = v.imass[ja];
      const int nnu = v.sp.norb[in2 - 1];
      const int nb = nmu * nnu;
      double r2[3];
      for (int q = 0; q < 3; q++) r2[q] = v.ratom[3 * ja + q] + xls[3 * mb + q];
      double z = d_distance(r1, r2), sighat[3], eps[9], pm[9];
      if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
      else for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / z;
      d_epsilon(r2, sighat, eps);
      d_pmat<double>(eps, pm);
      const int64_t bo = v.boff[p];
      double X[NB];
      if (mask & 1) {
        status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_S, 0), in1, in2, nmu, z, pm, X);
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_S][bo + i] = X[i];
      }
      if (mask & 2) {
        status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_T, 0), in1, in2, nmu, z, pm, X);
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_T][bo + i] = X[i];
      }
      if (mask & 4) {
        status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_DIP, 0), in1, in2, nmu, z, pm, X);
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_DIP][bo + i] = X[i];
      }
      if (mask & 8) {
        // ontop L and R into blocko (M_assemble_2c_DOGS.f90:640-702)
        double acc[NB];
        for (int i = 0; i < nb; i++) acc[i] = 0.0;
        const bool self = (v.neigh_j[p] == ia - St.atom_off + 1 && mb == 0);
        if (!self) {
          status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_ONL, 0), in1, in2, nmu, z, pm, X);
          for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i] * FB_P_EQ2;
          for (int isorp = 1; isorp <= v.sp.nssh[in1 - 1]; isorp++) {
            double dQ = v.dQ[ia * FB_MAXSH + isorp - 1];
            if (dQ == 0.0) continue;
            status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_ONL, isorp), in1, in2, nmu, z, pm, X);
            for (int i = 0; i < nb; i++) acc[i] = acc[i] + dQ * X[i] * FB_P_EQ2;
          }
          status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_ONR, 0), in1, in2, nmu, z, pm, X);
          for (int i = 0; i < nb; i++) acc[i] = acc[i] + X[i] * FB_P_EQ2;
          for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) {
            double dQ = v.dQ[ja * FB_MAXSH + isorp - 1];
            if (dQ == 0.0) continue;
            status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_ONR, isorp), in1, in2, nmu, z, pm, X);
            for (int i = 0; i < nb; i++) acc[i] = acc[i] + dQ * X[i] * FB_P_EQ2;
          }
        }
        for (int i = 0; i < nb; i++) v.mat[FB_MAT_VNAO][bo + i] = acc[i];
        if (p != pself)
          for (int i = 0; i < nb; i++) v.mat[FB_MAT_VNA][bo + i] = 0.0;
        // atom term <i|V_j|i> (M_assemble_2c_DOGS.f90:763-812)
        double rend = rc1min + v.sp.rcmin[in2 - 1];
        double smooth = d_smoother(z, rend, 0.8);
        status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_ATOM, 0), in1, in1, nmu, z, pm, X);
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
            status |= me2c_rotated(v, cell2c_id(v, in1, in2, SLOT_ATOM, isorp), in1, in1, nmu, z, pm, X);
            for (int i = 0; i < na; i++) {
              double emnpl = (z > 1.0e-4) ? sself[i] / z : 0.0;
              atom_c[i] = atom_c[i] + FB_P_EQ2 * dQ * ((1.0 - smooth) * emnpl + smooth * X[i]);
            }
          }
        }
      }
    }
    if (mask & 8) {
      // ordered sum over this chunk's neighbours into the self block
      for (int i = 0; i < NB; i++) s_atom[lane][i] = atom_c[i];
      __syncthreads();
      if (lane < nmu * nmu) {
        int kmax = min(W2C, n - c0);
        for (int kk = 0; kk < kmax; kk++) acc_atom = acc_atom + s_atom[kk][lane];
      }
      __syncthreads();
    }
  }
  if ((mask & 8) && lane < nmu * nmu) v.mat[FB_MAT_VNA][v.boff[pself] + lane] = acc_atom;
  if (status) atomicOr(v.status, status);
}

extern "C" int fbk_2c(const DView &v, int mask, hipStream_t st) {
  if (v.natoms == 0) return 0;
  hipLaunchKernelGGL(k_2c, dim3(v.natoms), dim3(W2C), 0, st, v, mask);
  return (int)hipGetLastError();
}
