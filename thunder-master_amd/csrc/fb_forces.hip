// Force assemblers of k.DASSEMBLERS in binary64 on gfx950:
//   Dassemble_S / _T / _dipole_z   Dblock(3, norb_mu, norb_nu) per pair
//                                  (M_Dassemble_2c_DOGS.f90:95-655)
//   Dassemble_vna_2c               forces%vna_ontop / vna_atom per pair
//                                  (M_Dassemble_2c_DOGS.f90:679-1199)
//   Dassemble_vna_3c               forces%f3naa / f3nab / f3nac per atom
//                                  (M_Dassemble_3c_DOGS.f90:136-793)
// with the derivative helpers getDMEs_Fdata_2c (e.FDATA/M_Fdata_2c.f90:446-600),
// getDMEs_Fdata_3c (e.FDATA/M_Fdata_3c.f90:549-846), Depsilon_2c / Depsilon_3c
// (b.FUNCTIONS), Dtwister / Drotate (c.SYSTEM/M_Drotations.f90:102-298) and
// Dsmoother (b.FUNCTIONS/Dsmoother.f90).  Expressions keep the reference's
// evaluation order (the library is compiled with -ffp-contract=off).
//
// Layout: 2c work is one wavefront per central atom with one lane per pair
// (the atom's pairs are contiguous), so the per-pair outputs (Dblocks, the two
// force vectors) are stored without any reduction.  3c work is one thread per
// triplet in the list (reference) order; each thread contracts its rotated
// derivative blocks with the density block of its target pair at once, so a
// triplet leaves nine numbers (the forces on ialpha, iatom and jatom), and a
// second kernel sums them per atom in three fixed orders (list order for
// ialpha, target-pair order for iatom, a stable jatom order for jatom) with a
// fixed wavefront tree: deterministic, no atomics.
#include "fb_device.h"
#include "fb_forces.h"

#define WD2 64
#define NME3F 10

// Levi-Civita and Kronecker tensors of include/constants.h (xlevi, delk), 1-based
__device__ __forceinline__ double d_levi(int i, int j, int k) {
  if (i == j || j == k || i == k) return 0.0;
  return ((i == 1 && j == 2) || (i == 2 && j == 3) || (i == 3 && j == 1)) ? 1.0 : -1.0;
}
__device__ __forceinline__ double d_delk(int i, int j) { return i == j ? 1.0 : 0.0; }

// v[k] for a k that is not a compile-time constant, without addressing a
// private array (which would put it in scratch memory): a blend with exact
// 0/1 weights, equal to v[k] (finite values)
__device__ __forceinline__ double sel3(const double *v, int k) {
  const double w0 = (k == 0) ? 1.0 : 0.0, w1 = (k == 1) ? 1.0 : 0.0, w2 = (k == 2) ? 1.0 : 0.0;
  return (v[0] * w0 + v[1] * w1) + v[2] * w2;
}
__device__ __forceinline__ void put3(double *v, int k, double x) {
  v[0] = (k == 0) ? x : v[0];
  v[1] = (k == 1) ? x : v[1];
  v[2] = (k == 2) ? x : v[2];
}

#define EPS_(e, i, j) (e)[3 * ((j) - 1) + (i) - 1]
#define DEP_(d, ix, imu, k) (d)[((ix) - 1) + 3 * ((imu) - 1) + 9 * ((k) - 1)]

__device__ __forceinline__ double d_Dsmoother(double r, double rend, double x) {
  double rbegin = x * rend;
  if (r > rend) return 0.0;
  if (r < rbegin) return 0.0;
  double ratio = (r - rbegin) / (rend - rbegin);
  return -(2.0 * (1.0 - ratio * ratio) * (2.0 * ratio) / (rend - rbegin));
}

// deps(ix, imu, k) = d eps(imu, k) / d r1(ix) of the 2c frame epsilon(r2, sighat)
__device__ inline void d_Depsilon_2c(const double *r1, const double *r2, double z, const double *eps, double *deps) {
#pragma unroll
  for (int i = 0; i < 27; i++) deps[i] = 0.0;
  if (z < 1.0e-5) return;
  double ca0 = r2[1] * r1[2] - r2[2] * r1[1];
  double ca1 = r2[2] * r1[0] - r2[0] * r1[2];
  double ca2 = r2[0] * r1[1] - r2[1] * r1[0];
  double crossmag = sqrt(ca0 * ca0 + ca1 * ca1 + ca2 * ca2);
  if (fabs(crossmag) < 1.0e-3) return;
  double r2mag2 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
  double r1dotr2 = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  double iz = 1.0 / z, ic = 1.0 / crossmag;
#pragma unroll
  for (int ix = 1; ix <= 3; ix++)
#pragma unroll
    for (int imu = 1; imu <= 3; imu++) {
      double sum = d_levi(ix, imu, 1) * r2[0] + d_levi(ix, imu, 2) * r2[1] + d_levi(ix, imu, 3) * r2[2];
      DEP_(deps, ix, imu, 1) = (EPS_(eps, imu, 1) * EPS_(eps, ix, 3) * iz) -
                               (EPS_(eps, imu, 1) * (ic * ic)) * (r2mag2 * r1[ix - 1] - r1dotr2 * r2[ix - 1]) +
                               iz * ic *
                                   (d_delk(imu, ix) * (r2mag2 - r1dotr2) - r1[imu - 1] * r2[ix - 1] -
                                    r2[imu - 1] * r2[ix - 1] + 2.0 * r1[ix - 1] * r2[imu - 1]);
      DEP_(deps, ix, imu, 2) = ic * (sum + (EPS_(eps, imu, 2) * ic) * (r1dotr2 * r2[ix - 1] - r2mag2 * r1[ix - 1]));
      DEP_(deps, ix, imu, 3) = -((d_delk(imu, ix) - EPS_(eps, imu, 3) * EPS_(eps, ix, 3)) * iz);
    }
}

// dera3 = d eps / d ratm and der13 = d eps / d r1 of the 3c frame epsilon(rhat, sighat)
__device__ inline void d_Depsilon_3c(const double *r1, const double *r2, const double *r21, double z,
                                     const double *ratm, const double *rnabc, const double *eps3, double *dera3,
                                     double *der13) {
#pragma unroll
  for (int i = 0; i < 27; i++) { dera3[i] = 0.0; der13[i] = 0.0; }
  double ca0 = r21[1] * rnabc[2] - r21[2] * rnabc[1];
  double ca1 = r21[2] * rnabc[0] - r21[0] * rnabc[2];
  double ca2 = r21[0] * rnabc[1] - r21[1] * rnabc[0];
  double crossmag = sqrt(ca0 * ca0 + ca1 * ca1 + ca2 * ca2);
  if (fabs(crossmag) < 1.0e-3) return;
  double r2mag2 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
  double r1mag2 = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
  double ratmmag2 = ratm[0] * ratm[0] + ratm[1] * ratm[1] + ratm[2] * ratm[2];
  double r1dotr2 = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  double r2dotratm = r2[0] * ratm[0] + r2[1] * ratm[1] + r2[2] * ratm[2];
  double r1dotratm = r1[0] * ratm[0] + r1[1] * ratm[1] + r1[2] * ratm[2];
#pragma unroll
  for (int ix = 1; ix <= 3; ix++)
#pragma unroll
    for (int imu = 1; imu <= 3; imu++) {
      double sum = d_levi(ix, imu, 1) * (r2[0] - r1[0]) + d_levi(ix, imu, 2) * (r2[1] - r1[1]) +
                   d_levi(ix, imu, 3) * (r2[2] - r1[2]);
      DEP_(dera3, ix, imu, 2) =
          (1.0 / crossmag) * (sum - (EPS_(eps3, imu, 2) / crossmag) *
                                        ((r1mag2 + r2mag2 - 2.0 * r1dotr2) * ratm[ix - 1] +
                                         (r1dotr2 - r1mag2 - r2dotratm + r1dotratm) * r2[ix - 1] +
                                         (r1dotr2 - r2mag2 + r2dotratm - r1dotratm) * r1[ix - 1]));
    }
#pragma unroll
  for (int ix = 1; ix <= 3; ix++) {
    DEP_(dera3, ix, 1, 1) = EPS_(eps3, 3, 3) * DEP_(dera3, ix, 2, 2) - EPS_(eps3, 2, 3) * DEP_(dera3, ix, 3, 2);
    DEP_(dera3, ix, 2, 1) = EPS_(eps3, 1, 3) * DEP_(dera3, ix, 3, 2) - EPS_(eps3, 3, 3) * DEP_(dera3, ix, 1, 2);
    DEP_(dera3, ix, 3, 1) = EPS_(eps3, 2, 3) * DEP_(dera3, ix, 1, 2) - EPS_(eps3, 1, 3) * DEP_(dera3, ix, 2, 2);
  }
#pragma unroll
  for (int ix = 1; ix <= 3; ix++)
#pragma unroll
    for (int imu = 1; imu <= 3; imu++) {
      DEP_(der13, ix, imu, 3) = (1.0 / z) * (EPS_(eps3, imu, 3) * EPS_(eps3, ix, 3) - d_delk(imu, ix));
      double sum = d_levi(ix, imu, 1) * (ratm[0] - r2[0]) + d_levi(ix, imu, 2) * (ratm[1] - r2[1]) +
                   d_levi(ix, imu, 3) * (ratm[2] - r2[2]);
      DEP_(der13, ix, imu, 2) =
          (1.0 / crossmag) * (sum - (EPS_(eps3, imu, 2) / crossmag) *
                                        ((r2dotratm - r1dotratm + r1dotr2 - r2mag2) * ratm[ix - 1] +
                                         (r2dotratm + r1dotratm - r1dotr2 - ratmmag2) * r2[ix - 1] +
                                         (ratmmag2 - 2.0 * r2dotratm + r2mag2) * r1[ix - 1]));
    }
#pragma unroll
  for (int ix = 1; ix <= 3; ix++) {
    DEP_(der13, ix, 1, 1) = EPS_(eps3, 3, 3) * DEP_(der13, ix, 2, 2) - EPS_(eps3, 3, 2) * DEP_(der13, ix, 2, 3) -
                            EPS_(eps3, 2, 3) * DEP_(der13, ix, 3, 2) + EPS_(eps3, 2, 2) * DEP_(der13, ix, 3, 3);
    DEP_(der13, ix, 2, 1) = EPS_(eps3, 1, 3) * DEP_(der13, ix, 3, 2) - EPS_(eps3, 1, 2) * DEP_(der13, ix, 3, 3) -
                            EPS_(eps3, 3, 3) * DEP_(der13, ix, 1, 2) + EPS_(eps3, 3, 2) * DEP_(der13, ix, 1, 3);
    DEP_(der13, ix, 3, 1) = EPS_(eps3, 2, 3) * DEP_(der13, ix, 1, 2) - EPS_(eps3, 2, 2) * DEP_(der13, ix, 1, 3) -
                            EPS_(eps3, 1, 3) * DEP_(der13, ix, 2, 2) + EPS_(eps3, 1, 2) * DEP_(der13, ix, 2, 3);
  }
}

// derivative (component ix, 1-based) of the p rotation: Dpmat(ix, :, :) of
// Dtwister, the (y, z, x) permutation of deps(ix, :, :)
__device__ __forceinline__ void d_dpmat(const double *deps, int ix, double *dpm) {
  dpm[0] = DEP_(deps, ix, 2, 2); dpm[3] = DEP_(deps, ix, 2, 3); dpm[6] = DEP_(deps, ix, 2, 1);
  dpm[1] = DEP_(deps, ix, 3, 2); dpm[4] = DEP_(deps, ix, 3, 3); dpm[7] = DEP_(deps, ix, 3, 1);
  dpm[2] = DEP_(deps, ix, 1, 2); dpm[5] = DEP_(deps, ix, 1, 3); dpm[8] = DEP_(deps, ix, 1, 1);
}

// derivative (component ix, 1-based) of the d rotation: Ddmat(ix, :, :) of
// Dtwister (c.SYSTEM/M_Drotations.f90:203-298) with the A matrices of
// include/constants.h:26-35, the sums in the reference's order; column-major
// ddm[5 (m - 1) + (k - 1)] = Ddmat(ix, k, m)
__device__ inline void d_ddmat(const double *eps, const double *deps, int ix, double *ddm) {
  const double r3 = 1.0 / (2.0 * sqrt(3.0));
  // amat(i, j, imu) = amat(j, i, imu): imu = xy, yz, z^2, xz, x^2-y^2
  const double amat[5][3][3] = {
      {{0.0, 0.5, 0.0}, {0.5, 0.0, 0.0}, {0.0, 0.0, 0.0}},
      {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.5}, {0.0, 0.5, 0.0}},
      {{-r3, 0.0, 0.0}, {0.0, -r3, 0.0}, {0.0, 0.0, 2.0 / (2.0 * sqrt(3.0))}},
      {{0.0, 0.0, 0.5}, {0.0, 0.0, 0.0}, {0.5, 0.0, 0.0}},
      {{0.5, 0.0, 0.0}, {0.0, -0.5, 0.0}, {0.0, 0.0, 0.0}}};
  for (int imu = 0; imu < 5; imu++) {
    double a12 = 0.0, a32 = 0.0, a33 = 0.0, a13 = 0.0, a11 = 0.0;
    for (int i = 1; i <= 3; i++)
      for (int j = 1; j <= 3; j++) {
        const double a = amat[imu][j - 1][i - 1];
        if (a != 0.0) {
          a12 = a12 + a * (DEP_(deps, ix, i, 1) * EPS_(eps, j, 2) + EPS_(eps, i, 1) * DEP_(deps, ix, j, 2));
          a32 = a32 + a * (DEP_(deps, ix, i, 3) * EPS_(eps, j, 2) + EPS_(eps, i, 3) * DEP_(deps, ix, j, 2));
          a33 = a33 + a * (DEP_(deps, ix, i, 3) * EPS_(eps, j, 3) + EPS_(eps, i, 3) * DEP_(deps, ix, j, 3));
          a13 = a13 + a * (DEP_(deps, ix, i, 1) * EPS_(eps, j, 3) + EPS_(eps, i, 1) * DEP_(deps, ix, j, 3));
          a11 = a11 + a * (DEP_(deps, ix, i, 1) * EPS_(eps, j, 1) + EPS_(eps, i, 1) * DEP_(deps, ix, j, 1));
        }
      }
    ddm[0 * 5 + imu] = 2.0 * a12;
    ddm[1 * 5 + imu] = 2.0 * a32;
    ddm[2 * 5 + imu] = sqrt(3.0) * a33;
    ddm[3 * 5 + imu] = 2.0 * a13;
    ddm[4 * 5 + imu] = 2.0 * a11 + a33;
  }
}

// Drotate, one derivative component: X = dL M R^T + L M dR^T + L dM R^T shell
// pair by shell pair with the reference's loop order (m2, k2, m1, k1) and the
// s-shell rotation 1, derivative 0 of Dchooser; dm, ddm: the d rotation and
// its derivative (read only for d shells).  M, DM, X column-major, ld nmu.
__device__ inline void d_Drot(const DSpecies &sp, int in1, int in2, const double *pm, const double *dpm,
                              const double *dm, const double *ddm, const double *M, const double *DM, int nmu,
                              double *X) {
  const int nnu = sp.norb[in2 - 1];
  for (int i = 0; i < nmu * nnu; i++) X[i] = 0.0;
  int n1 = 0;
  for (int issh = 0; issh < sp.nssh[in1 - 1]; issh++) {
    const int l1 = sp.lssh[in1 - 1][issh];
    int n2 = 0;
    for (int jssh = 0; jssh < sp.nssh[in2 - 1]; jssh++) {
      const int l2 = sp.lssh[in2 - 1][jssh];
      for (int m2 = 0; m2 < 2 * l2 + 1; m2++)
        for (int k2 = 0; k2 < 2 * l2 + 1; k2++)
          for (int m1 = 0; m1 < 2 * l1 + 1; m1++)
            for (int k1 = 0; k1 < 2 * l1 + 1; k1++) {
              const double lf = (l1 == 0) ? 1.0 : (l1 == 1 ? pm[3 * m1 + k1] : dm[5 * m1 + k1]);
              const double dl = (l1 == 0) ? 0.0 : (l1 == 1 ? dpm[3 * m1 + k1] : ddm[5 * m1 + k1]);
              const double rt = (l2 == 0) ? 1.0 : (l2 == 1 ? pm[3 * m2 + k2] : dm[5 * m2 + k2]);
              const double dr = (l2 == 0) ? 0.0 : (l2 == 1 ? dpm[3 * m2 + k2] : ddm[5 * m2 + k2]);
              const int o = (n2 + m2) * nmu + (n1 + m1);
              double *x = &X[(n2 + k2) * nmu + (n1 + k1)];
              *x = *x + dl * rt * M[o] + lf * dr * M[o] + lf * rt * DM[o];
            }
      n2 += 2 * l2 + 1;
    }
    n1 += 2 * l1 + 1;
  }
}

// getDMEs_Fdata_2c on one cell: values and x-derivatives of the nME elements
__device__ inline int d_getD2c(const DCell2c &c, const double *__restrict__ tab, double x, double *F, double *dF) {
  if (x < 0.0) return ST_X_NEG;
  if (x > c.xmax * (1.0 + 1.0e-5)) return ST_X_RANGE;
  const double *base = tab + c.off;
  if (x < 1.0e-5) {
    for (int m = 0; m < c.nME; m++) { F[m] = base[m]; dF[m] = 0.0; }
    return 0;
  }
  int imid = (int)(x / c.dx) + 1;
  int ileft = imid - 3, iright = imid + 3;
  if (ileft < 1) { ileft = 1; iright = 7; }
  else if (iright > c.nx) { ileft = c.nx - 6; iright = c.nx; }
  double xx[7], pdenom[7];
#pragma unroll
  for (int k = 0; k < 7; k++) xx[k] = (double)(ileft + k - 1) * c.dx;
  for (int m = 0; m < c.nME; m++) { F[m] = 0.0; dF[m] = 0.0; }
#pragma unroll
  for (int is = 0; is < 7; is++) {
    double prod = 1.0;
    pdenom[is] = 1.0;
#pragma unroll
    for (int ip = 0; ip < 7; ip++) {
      if (ip != is) {
        double pden = 1.0 / (xx[is] - xx[ip]);
        pdenom[is] = pdenom[is] * pden;
        prod = prod * (x - xx[ip]) * pden;
      }
    }
    const double *row = base + (int64_t)(ileft - 1 + is) * c.nME;
    for (int m = 0; m < c.nME; m++) F[m] = F[m] + row[m] * prod;
  }
#pragma unroll
  for (int ipt = 0; ipt < 7; ipt++) {
    double xsum = 0.0;
#pragma unroll
    for (int jpt = 0; jpt < 7; jpt++) {
      if (ipt == jpt) continue;
      double xprod = 1.0;
#pragma unroll
      for (int kpt = 0; kpt < 7; kpt++) {
        if (kpt == jpt || kpt == ipt) continue;
        xprod = xprod * (x - xx[kpt]);
      }
      xsum = xsum + xprod;
    }
    const double *row = base + (int64_t)(ileft - 1 + ipt) * c.nME;
    for (int m = 0; m < c.nME; m++) dF[m] = dF[m] + row[m] * pdenom[ipt] * xsum;
  }
  return 0;
}

__device__ __forceinline__ int cell2c_of(const DView &v, int in1, int in2, int slot, int isub) {
  return v.idx2c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * FB_NSLOT2 + slot) * FB_MAXSUB2 + isub];
}

// the generic kernels' sizes: s/p species (MO 4), or with d shells (MO 9)
template <int MO> struct GenF;
template <> struct GenF<FB_MO> {
  static constexpr int NME2 = 16, NME3 = NME3F;
};
template <> struct GenF<FB_MOD> {
  static constexpr int NME2 = FB_NME2_WIDE, NME3 = FB_NME3_WIDE;
};

// value and r1-derivative blocks (molecular frame) of one 2c interaction:
// hm = getDMEs value, vd(ix) = -eta(ix) dF/dx where z > zcut (else 0)
template <int MO>
__device__ inline int d_dblock2c(const DView &v, int cid, int nmu, int nnu, double z, double *hm, double *dhm) {
  if (cid < 0) return ST_NO_TABLE;
  const DCell2c c = v.cell2c[cid];
  if (c.nME > GenF<MO>::NME2) return ST_CAPACITY;
  double F[GenF<MO>::NME2], dF[GenF<MO>::NME2];
  int st = d_getD2c(c, v.tab2c, z, F, dF);
  if (st) return st;
  for (int i = 0; i < nmu * nnu; i++) { hm[i] = 0.0; dhm[i] = 0.0; }
  for (int k = 0; k < c.nME; k++) {
    const int o = (v.nu[c.munu + k] - 1) * nmu + (v.mu[c.munu + k] - 1);
    hm[o] = F[k];
    dhm[o] = dF[k];
  }
  return 0;
}

// f(ix) -= rho(o) [dq] X_ix(o) P_eq2 over o, with X_ix the Drotate of (hm, -eta dhm)
template <int MO>
__device__ inline void d_contract_2c(const DView &v, int in1, int in2, int nmu, double z, double zcut,
                                     const double *pm, const double *dm, const double *eps, const double *deps,
                                     const double *eta, const double *hm, const double *dhm, const double *rho,
                                     double dq, bool use_dq, double *f) {
  const int nb = nmu * v.sp.norb[in2 - 1];
  double DM[MO * MO], X[MO * MO], dpm[9], ddm[MO == FB_MOD ? 25 : 1];
#pragma unroll
  for (int ix = 0; ix < 3; ix++) {
    for (int i = 0; i < nb; i++) DM[i] = (z > zcut) ? -eta[ix] * dhm[i] : 0.0;
    d_dpmat(deps, ix + 1, dpm);
    if (MO == FB_MOD) d_ddmat(eps, deps, ix + 1, ddm);
    d_Drot(v.sp, in1, in2, pm, dpm, dm, ddm, hm, DM, nmu, X);
    for (int i = 0; i < nb; i++) {
      if (use_dq) f[ix] = f[ix] - rho[i] * dq * X[i] * FB_P_EQ2;
      else f[ix] = f[ix] - rho[i] * X[i] * FB_P_EQ2;
    }
  }
}

// mask bits: 1 Dblock S, 2 Dblock T, 4 Dblock dipole_z, 8 vna_ontop / vna_atom
template <int MO>
__global__ __launch_bounds__(WD2) void k_d2c(DView v, DForce F, int mask) {
  constexpr int NB = MO * MO;
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
  for (int k = lane; k < n; k += WD2) {
    const int64_t p = p0 + k;
    const int ja = St.atom_off + v.neigh_j[p] - 1;
    const int mb = v.neigh_b[p];
    const int in2 = v.imass[ja];
    const int nnu = v.sp.norb[in2 - 1];
    const int nb = nmu * nnu;
    double r2[3];
    for (int q = 0; q < 3; q++) r2[q] = v.ratom[3 * ja + q] + xls[3 * mb + q];
    double z = d_distance(r1, r2), sighat[3], eps[9], pm[9], deps[27], eta[3];
    if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
    else for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / z;
    d_epsilon(r2, sighat, eps);
    d_Depsilon_2c(r1, r2, z, eps, deps);
    d_pmat<double>(eps, pm);
    double dm[MO == FB_MOD ? 25 : 1], ddm[MO == FB_MOD ? 25 : 1];
    if (MO == FB_MOD) d_dmat(eps, dm);
    for (int q = 0; q < 3; q++) eta[q] = eps[6 + q];
    const int64_t bo = v.boff[p];
    double hm[NB], dhm[NB], DM[NB], X[NB], dpm[9];
    const int slots[3] = {SLOT_S, SLOT_T, SLOT_DIP};
    for (int d = 0; d < 3; d++) {
      if (!(mask & (1 << d))) continue;
      int st = d_dblock2c<MO>(v, cell2c_of(v, in1, in2, slots[d], 0), nmu, nnu, z, hm, dhm);
      status |= st;
      if (st) continue;
      double *D = F.dmat[d] + 3 * bo;
      for (int ix = 0; ix < 3; ix++) {
        for (int i = 0; i < nb; i++) DM[i] = (z > 1.0e-3) ? -eta[ix] * dhm[i] : 0.0;
        d_dpmat(deps, ix + 1, dpm);
        if (MO == FB_MOD) d_ddmat(eps, deps, ix + 1, ddm);
        d_Drot(v.sp, in1, in2, pm, dpm, dm, ddm, hm, DM, nmu, X);
        for (int i = 0; i < nb; i++) D[ix + 3 * i] = X[i];
      }
    }
    if (mask & 8) {
      // ontop L and R, contracted with the pair's density block (:760-999)
      double fo[3] = {0.0, 0.0, 0.0};
      const bool self = (v.neigh_j[p] == ia - St.atom_off + 1 && mb == 0);
      if (!self) {
        const double *rho = F.rho + bo;
        for (int side = 0; side < 2; side++) {
          const int slot = side ? SLOT_ONR : SLOT_ONL;
          const int qa = side ? ja : ia;
          const int inq = side ? in2 : in1;
          int st = d_dblock2c<MO>(v, cell2c_of(v, in1, in2, slot, 0), nmu, nnu, z, hm, dhm);
          status |= st;
          if (!st) d_contract_2c<MO>(v, in1, in2, nmu, z, 1.0e-3, pm, dm, eps, deps, eta, hm, dhm, rho, 0.0, false, fo);
          for (int isorp = 1; isorp <= v.sp.nssh[inq - 1]; isorp++) {
            const double dQ = v.dQ[qa * FB_MAXSH + isorp - 1];
            if (dQ == 0.0) continue;
            st = d_dblock2c<MO>(v, cell2c_of(v, in1, in2, slot, isorp), nmu, nnu, z, hm, dhm);
            status |= st;
            if (!st) d_contract_2c<MO>(v, in1, in2, nmu, z, 1.0e-3, pm, dm, eps, deps, eta, hm, dhm, rho, dQ, true, fo);
          }
        }
      }
      // atom term <i|V_j|i>, contracted with the density block of neigh_self (:1005-1190)
      double fa[3] = {0.0, 0.0, 0.0};
      const double *rho_s = F.rho + v.boff[pself];
      int st = d_dblock2c<MO>(v, cell2c_of(v, in1, in2, SLOT_ATOM, 0), nmu, nmu, z, hm, dhm);
      status |= st;
      if (!st) d_contract_2c<MO>(v, in1, in1, nmu, z, 1.0e-3, pm, dm, eps, deps, eta, hm, dhm, rho_s, 0.0, false, fa);
      bool charged = false;
      for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) charged |= (v.dQ[ja * FB_MAXSH + isorp - 1] != 0.0);
      if (charged) {
        const double rend = v.sp.rcmin[in1 - 1] + v.sp.rcmin[in2 - 1];
        const double smooth = d_smoother(z, rend, 0.8), Dsmooth = d_Dsmoother(z, rend, 0.8);
        const double *sself = v.mat[FB_MAT_S] + v.boff[pself];
        const int na = nmu * nmu;
        for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) {
          const double dQ = v.dQ[ja * FB_MAXSH + isorp - 1];
          if (dQ == 0.0) continue;
          st = d_dblock2c<MO>(v, cell2c_of(v, in1, in2, SLOT_ATOM, isorp), nmu, nmu, z, hm, dhm);
          status |= st;
          if (st) continue;
          double bx[NB];
          d_rotate<double>(v.sp, in1, in1, pm, hm, nmu, bx, MO == FB_MOD ? dm : nullptr);
          for (int ix = 0; ix < 3; ix++) {
            for (int i = 0; i < na; i++) DM[i] = (z > 1.0e-4) ? -eta[ix] * dhm[i] : 0.0;
            d_dpmat(deps, ix + 1, dpm);
            if (MO == FB_MOD) d_ddmat(eps, deps, ix + 1, ddm);
            d_Drot(v.sp, in1, in1, pm, dpm, dm, ddm, hm, DM, nmu, X);
            for (int i = 0; i < na; i++) {
              const double emnpl = (z > 1.0e-4) ? sself[i] / z : 0.0;
              const double vdem = (z > 1.0e-4) ? eta[ix] * sself[i] / (z * z) : 0.0;
              fa[ix] = fa[ix] - rho_s[i] * FB_P_EQ2 * dQ *
                                    (smooth * X[i] - eta[ix] * Dsmooth * bx[i] + (1.0 - smooth) * vdem +
                                     eta[ix] * Dsmooth * emnpl);
            }
          }
        }
      }
      for (int q = 0; q < 3; q++) {
        F.fontop[3 * p + q] = fo[q];
        F.fatom[3 * p + q] = fa[q];
      }
    }
  }
  if (status) atomicOr(v.status, status);
}

// ---------------------------------------------------------------- 3c

__device__ __forceinline__ int group3c_of(const DView &v, int in1, int in2, int in3, int isub) {
  return v.idx3c[(((in1 - 1) * v.sp.ns + (in2 - 1)) * v.sp.ns + (in3 - 1)) * FB_MAXSUB3 + isub];
}

// getDMEs_Fdata_3c: value, d/dcos, d/dx, d/dy blocks (molecular frame).  The
// 3-point rule takes the slopes of ME 1 of theta cell 1 for every element.
template <int MO>
__device__ inline int d_getD3c(const DView &v, int gid, double x, double y, double cost, int nmu, int nnu,
                               double *hm, double *dpm_, double *dxm, double *dym) {
  if (gid < 0) return ST_NO_TABLE;
  const DGroup3c g = v.grp3c[gid];
  const int nME = g.nME, nx = g.nx, ny = g.ny;
  if (nME > GenF<MO>::NME3) return ST_CAPACITY;
  double p[5], Dp[5];
  const double cost2 = cost * cost;
  const double sint = sqrt(fmax(1.0e-5, 1.0 - cost2));
  p[0] = 1.0;
  p[1] = cost;
  p[2] = 0.5 * (3.0 * cost2 - 1.0);
  p[3] = 0.5 * cost * (5.0 * cost2 - 3.0);
  p[4] = 0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2));
  Dp[0] = 0.0;
  Dp[1] = 1.0;
  Dp[2] = 3.0 * cost;
  Dp[3] = (15.0 * cost2 - 3.0) / 2.0;
  Dp[4] = 0.5 * cost * (35.0 * cost2 - 15.0);
  const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  const double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
  const double *tab = v.tab3c + g.off;
  const int64_t tstride = (int64_t)ny * nx * nME;
#define T3F(th, i, j, m) tab[(th) * tstride + ((int64_t)((j) - 1) * nx + ((i) - 1)) * nME + (m)]
  const double gx = (T3F(0, ix + 1, iy, 0) - T3F(0, ix, iy, 0)) / dx;
  const double gy = (T3F(0, ix, iy + 1, 0) - T3F(0, ix, iy, 0)) / dy;
  const double gg = gx * gx + gy * gy;
  double Fv[GenF<MO>::NME3], DpF[GenF<MO>::NME3], DxF[GenF<MO>::NME3], DyF[GenF<MO>::NME3];
  for (int m = 0; m < nME; m++) { Fv[m] = 0.0; DpF[m] = 0.0; DxF[m] = 0.0; DyF[m] = 0.0; }
  for (int th = 0; th < FB_NTHETA; th++) {
    const double pt = p[th], dpt = Dp[th];
    for (int m = 0; m < nME; m++) {
      if (gg < 1.0e-6) {
        const double interim =
            ((1.0 - px - py) * T3F(th, ix, iy, m) + px * T3F(th, ix + 1, iy, m)) + py * T3F(th, ix, iy + 1, m);
        Fv[m] = Fv[m] + pt * interim;
        DpF[m] = DpF[m] + dpt * interim;
        DxF[m] = DxF[m] + pt * gx;
        DyF[m] = DyF[m] + pt * gy;
      } else {
        double gk[4], gp[4];
        for (int k = -1; k <= 2; k++) {
          double f1m1 = T3F(th, ix + k, iy - 1, m);
          double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
          double f0p3 = 3.0 * T3F(th, ix + k, iy, m), f0p6 = 2.0 * f0p3;
          double f1p3 = 3.0 * T3F(th, ix + k, iy + 1, m), f1p6 = 2.0 * f1p3;
          double f2p1 = T3F(th, ix + k, iy + 2, m);
          double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
          double bb2 = f1m3 - f0p6 + f1p3;
          double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
          double bb0 = f0p6;
          gk[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
          gp[k + 1] = ((3.0 * bb3 * py + 2.0 * bb2) * py + bb1);
        }
        double f1m1 = gk[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
        double f0p3 = 3.0 * gk[1], f0p6 = 2.0 * f0p3;
        double f1p3 = 3.0 * gk[2], f1p6 = 2.0 * f1p3;
        double f2p1 = gk[3];
        double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        double bb2 = f1m3 - f0p6 + f1p3;
        double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        double bb0 = f0p6;
        const double interim = (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
        Fv[m] = Fv[m] + pt * interim;
        DpF[m] = DpF[m] + dpt * interim;
        DxF[m] = DxF[m] + pt * ((3.0 * bb3 * px + 2.0 * bb2) * px + bb1) / (36.0 * dx);
        f1m1 = gp[0]; f1m2 = 2.0 * f1m1; f1m3 = 3.0 * f1m1;
        f0p3 = 3.0 * gp[1]; f0p6 = 2.0 * f0p3;
        f1p3 = 3.0 * gp[2]; f1p6 = 2.0 * f1p3;
        f2p1 = gp[3];
        bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        bb2 = f1m3 - f0p6 + f1p3;
        bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        bb0 = f0p6;
        DyF[m] = DyF[m] + pt * (((bb3 * px + bb2) * px + bb1) * px + bb0) / (36.0 * dy);
      }
    }
  }
#undef T3F
  for (int i = 0; i < nmu * nnu; i++) { hm[i] = 0.0; dpm_[i] = 0.0; dxm[i] = 0.0; dym[i] = 0.0; }
  for (int k = 0; k < nME; k++) {
    const int o = (v.nu[g.munu + k] - 1) * nmu + (v.mu[g.munu + k] - 1);
    if (v.mv[g.munu + k] != 1) {
      hm[o] = Fv[k]; dpm_[o] = DpF[k]; dxm[o] = DxF[k]; dym[o] = DyF[k];
    } else {
      hm[o] = Fv[k] * sint;
      dpm_[o] = DpF[k] * sint - cost * Fv[k] / sint;
      dxm[o] = DxF[k] * sint;
      dym[o] = DyF[k] * sint;
    }
  }
  return 0;
}

struct DGeo3 {
  double r1[3], r2[3], rna[3], sighat[3], rhat[3], amt[3], bmt[3];
  double z, x, cost, d13, d23;
  double eps[9], depsA[27], depsB[27];
};

// triplet frame and its derivatives; charged = the rhat expression of the
// charged loop (M_Dassemble_3c_DOGS.f90:526), rnabc / x otherwise (:336)
__device__ inline void d_dgeo3(const DView &v, int4 tr, bool charged, DGeo3 &G) {
  const int ial = tr.x, gia = tr.y, gja = tr.z;
  const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
  const int s = v.atom_struct[ial];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  double r21[3], rnabc[3];
  for (int k = 0; k < 3; k++) {
    G.r1[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
    G.r2[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
    G.rna[k] = v.ratom[3 * ial + k];
    r21[k] = G.r2[k] - G.r1[k];
  }
  G.z = d_distance(G.r1, G.r2);
  if (G.z < 1.0e-5) { G.sighat[0] = 0.0; G.sighat[1] = 0.0; G.sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) G.sighat[k] = r21[k] / G.z;
  for (int k = 0; k < 3; k++) rnabc[k] = G.rna[k] - (G.r1[k] + r21[k] / 2.0);
  G.x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
  G.d13 = d_distance(G.rna, G.r1);
  G.d23 = d_distance(G.rna, G.r2);
  if (G.x < 1.0e-5) { G.rhat[0] = 0.0; G.rhat[1] = 0.0; G.rhat[2] = 0.0; }
  else if (charged) for (int k = 0; k < 3; k++) G.rhat[k] = (G.rna[k] - 0.5 * (G.r1[k] + G.r2[k])) / G.x;
  else for (int k = 0; k < 3; k++) G.rhat[k] = rnabc[k] / G.x;
  G.cost = G.sighat[0] * G.rhat[0] + G.sighat[1] * G.rhat[1] + G.sighat[2] * G.rhat[2];
  d_epsilon(G.rhat, G.sighat, G.eps);
  d_Depsilon_3c(G.r1, G.r2, r21, G.z, G.rna, rnabc, G.eps, G.depsA, G.depsB);
  for (int k = 0; k < 3; k++) {
    G.amt[k] = (G.x > 1.0e-5) ? (G.sighat[k] - G.cost * G.rhat[k]) / G.x : 0.0;
    G.bmt[k] = (G.cost * G.sighat[k] - G.rhat[k]) / G.z;
  }
}

// f3naMa / f3naMb of component ix (M_Dassemble_3c_DOGS.f90:388-399) on the
// elements of the group's cell, rotated into Xa / Xb
template <int MO>
__device__ inline void d_f3_rot(const DView &v, const DGroup3c &g, int in1, int in2, int nmu, const DGeo3 &G,
                                const double *pm, const double *dm, int ix, const double *hm, const double *dp,
                                const double *dxm, const double *dym, double *Xa, double *Xb) {
  const int nb = nmu * v.sp.norb[in2 - 1];
  double Ma[MO * MO], Mb[MO * MO], dpm[9], ddm[MO == FB_MOD ? 25 : 1];
  for (int i = 0; i < nb; i++) { Ma[i] = 0.0; Mb[i] = 0.0; }
  for (int k = 0; k < g.nME; k++) {
    const int o = (v.nu[g.munu + k] - 1) * nmu + (v.mu[g.munu + k] - 1);
    Ma[o] = G.rhat[ix] * dxm[o] + G.amt[ix] * dp[o];
    Mb[o] = -G.sighat[ix] * dym[o] + G.bmt[ix] * dp[o] - Ma[o] / 2.0;
  }
  d_dpmat(G.depsA, ix + 1, dpm);
  if (MO == FB_MOD) d_ddmat(G.eps, G.depsA, ix + 1, ddm);
  d_Drot(v.sp, in1, in2, pm, dpm, dm, ddm, hm, Ma, nmu, Xa);
  d_dpmat(G.depsB, ix + 1, dpm);
  if (MO == FB_MOD) d_ddmat(G.eps, G.depsB, ix + 1, ddm);
  d_Drot(v.sp, in1, in2, pm, dpm, dm, ddm, hm, Mb, nmu, Xb);
}

#define T3D 64

template <int MO>
__global__ __launch_bounds__(T3D) void k_d3c(DView v, DForce F) {
  constexpr int NB = MO * MO;
  const int64_t t = xcd_tile(blockIdx.x, gridDim.x) * T3D + threadIdx.x;
  if (t >= v.ntrip) return;
  const int4 tr = F.trip[t];
  const int tp = F.tpair[t];
  const int in1 = v.imass[tr.y], in2 = v.imass[tr.z], indna = v.imass[tr.x];
  const int nmu = v.sp.norb[in1 - 1], nnu = v.sp.norb[in2 - 1], nb = nmu * nnu;
  const int64_t bo = v.boff[tp];
  const double *rho = F.rho + bo;
  int status = 0;
  double fa[3] = {0.0, 0.0, 0.0}, fb[3] = {0.0, 0.0, 0.0}, fc[3] = {0.0, 0.0, 0.0};
  double hm[NB], dp[NB], dxm[NB], dym[NB], Xa[NB], Xb[NB], pm[9], dm[MO == FB_MOD ? 25 : 1];
  {
    DGeo3 G;
    d_dgeo3(v, tr, false, G);
    d_pmat<double>(G.eps, pm);
    if (MO == FB_MOD) d_dmat(G.eps, dm);
    const int gid = group3c_of(v, in1, in2, indna, 0);
    int st = d_getD3c<MO>(v, gid, G.x, G.z, G.cost, nmu, nnu, hm, dp, dxm, dym);
    status |= st;
    if (!st) {
      const DGroup3c g = v.grp3c[gid];
      for (int ix = 0; ix < 3; ix++) {
        d_f3_rot<MO>(v, g, in1, in2, nmu, G, pm, dm, ix, hm, dp, dxm, dym, Xa, Xb);
        for (int i = 0; i < nb; i++) {
          const double xc = -Xa[i] - Xb[i];
          fa[ix] = fa[ix] - rho[i] * Xa[i] * FB_P_EQ2;
          fb[ix] = fb[ix] - rho[i] * Xb[i] * FB_P_EQ2;
          fc[ix] = fc[ix] - rho[i] * xc * FB_P_EQ2;
        }
      }
    }
  }
  bool charged = false;
  for (int isorp = 1; isorp <= v.sp.nssh[indna - 1]; isorp++) charged |= (v.dQ[tr.x * FB_MAXSH + isorp - 1] != 0.0);
  if (charged) {
    // smoothed monopole terms (M_Dassemble_3c_DOGS.f90:528-776)
    DGeo3 G;
    d_dgeo3(v, tr, true, G);
    d_pmat<double>(G.eps, pm);
    if (MO == FB_MOD) d_dmat(G.eps, dm);
    double rA1[3] = {0.0, 0.0, 0.0}, rA2[3] = {0.0, 0.0, 0.0};
    if (G.d13 > 1.0e-5) for (int k = 0; k < 3; k++) rA1[k] = (G.rna[k] - G.r1[k]) / G.d13;
    if (G.d23 > 1.0e-5) for (int k = 0; k < 3; k++) rA2[k] = (G.rna[k] - G.r2[k]) / G.d23;
    const double rc3 = v.sp.rcmin[indna - 1];
    const double rend1 = v.sp.rcmin[in1 - 1] + rc3, rend2 = v.sp.rcmin[in2 - 1] + rc3;
    const double st13 = d_smoother(G.d13, rend1, 0.8), dst13 = d_Dsmoother(G.d13, rend1, 0.8);
    const double st23 = d_smoother(G.d23, rend2, 0.8), dst23 = d_Dsmoother(G.d23, rend2, 0.8);
    const double stinky = st13 * st23;
    double dstnA[3], dstnB[3], dstnC[3];
    for (int k = 0; k < 3; k++) {
      dstnB[k] = -dst13 * st23 * rA1[k];
      dstnC[k] = -st13 * dst23 * rA2[k];
      dstnA[k] = -dstnB[k] - dstnC[k];
    }
    const double *sb = v.mat[FB_MAT_S] + bo, *db = v.mat[FB_MAT_DIP] + bo;
    const double *sD = F.dmat[0] + 3 * bo, *dD = F.dmat[2] + 3 * bo;
    const double d13sq = G.d13 * G.d13, d23sq = G.d23 * G.d23;
    for (int isorp = 1; isorp <= v.sp.nssh[indna - 1]; isorp++) {
      const double dQ = v.dQ[tr.x * FB_MAXSH + isorp - 1];
      if (dQ == 0.0) continue;
      const int gid = group3c_of(v, in1, in2, indna, isorp);
      int st = d_getD3c<MO>(v, gid, G.x, G.z, G.cost, nmu, nnu, hm, dp, dxm, dym);
      status |= st;
      if (st) continue;
      const DGroup3c g = v.grp3c[gid];
      double bx[NB];
      d_rotate<double>(v.sp, in1, in2, pm, hm, nmu, bx, MO == FB_MOD ? dm : nullptr);
      for (int ix = 0; ix < 3; ix++) {
        d_f3_rot<MO>(v, g, in1, in2, nmu, G, pm, dm, ix, hm, dp, dxm, dym, Xa, Xb);
        for (int i = 0; i < nb; i++) {
          const double sterm = sb[i] / 2.0, dterm = db[i] / G.z;
          const double emnpl = (sterm - dterm) / G.d13 + (sterm + dterm) / G.d23;
          const double spterm = sD[ix + 3 * i] / 2.0;
          double dpterm = dD[ix + 3 * i] / G.z;
          dpterm = dpterm + dterm * G.sighat[ix] / G.z;
          const double dA = -((sterm - dterm) * rA1[ix] / d13sq) - (sterm + dterm) * rA2[ix] / d23sq;
          const double dB = (sterm - dterm) * rA1[ix] / d13sq + (spterm - dpterm) / G.d13 + (spterm + dpterm) / G.d23;
          const double dC = -dA - dB;
          const double xc = -Xa[i] - Xb[i];
          const double w = FB_P_EQ2 * dQ * rho[i];
          fa[ix] = fa[ix] - w * (stinky * Xa[i] + dstnA[ix] * bx[i] - dstnA[ix] * emnpl + (1.0 - stinky) * dA);
          fb[ix] = fb[ix] - w * (stinky * Xb[i] + dstnB[ix] * bx[i] - dstnB[ix] * emnpl + (1.0 - stinky) * dB);
          fc[ix] = fc[ix] - w * (stinky * xc + dstnC[ix] * bx[i] - dstnC[ix] * emnpl + (1.0 - stinky) * dC);
        }
      }
    }
  }
  const int64_t ob = 3 * (int64_t)F.inv_perm[t], oc = 3 * (int64_t)F.inv_perm_j[t];
  for (int k = 0; k < 3; k++) {
    F.f3a[3 * t + k] = fa[k];
    F.f3b[ob + k] = fb[k];
    F.f3c[oc + k] = fc[k];
  }
  if (status) atomicOr(v.status, status);
}

// per-atom sums of the triplet forces: one wavefront per atom, lane-strided
// partial sums in a fixed order and a fixed butterfly, so results repeat bit
// for bit from run to run
__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x = x + __shfl_xor(x, off, 64);
  return x;
}

__global__ __launch_bounds__(64) void k_f3_reduce(DView v, DForce F) {
  const int a = blockIdx.x;
  const int lane = threadIdx.x;
  double s[9];
#pragma unroll
  for (int k = 0; k < 9; k++) s[k] = 0.0;
  // each role's records of this atom are contiguous in its own array
  for (int64_t t = F.toff[a] + lane; t < F.toff[a + 1]; t += 64)
#pragma unroll
    for (int k = 0; k < 3; k++) s[k] = s[k] + F.f3a[3 * t + k];
  for (int64_t q = F.ioff[a] + lane; q < F.ioff[a + 1]; q += 64)
#pragma unroll
    for (int k = 0; k < 3; k++) s[3 + k] = s[3 + k] + F.f3b[3 * q + k];
  for (int64_t q = F.joff[a] + lane; q < F.joff[a + 1]; q += 64)
#pragma unroll
    for (int k = 0; k < 3; k++) s[6 + k] = s[6 + k] + F.f3c[3 * q + k];
  for (int k = 0; k < 9; k++) {
    const double r = wave_sum(s[k]);
    if (lane == 0) F.f3[9 * a + k] = r;
  }
}

// triplet ranges per atom in the target-pair order (iatom) and in the jatom
// order: lower bounds of the atom's first pair / id in the sorted keys
__device__ __forceinline__ int64_t lower_bound_i32(const int *a, int64_t n, int64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = (lo + hi) >> 1;
    if ((int64_t)a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__global__ void k_f3_offsets(int natoms, int64_t ntrip, const int64_t *noff, const int *tpair_sorted,
                             const int *jkey_sorted, int64_t *ioff, int64_t *joff) {
  const int a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a > natoms) return;
  ioff[a] = lower_bound_i32(tpair_sorted, ntrip, noff[a]);
  joff[a] = lower_bound_i32(jkey_sorted, ntrip, a);
}

__global__ void k_invert_perm(int64_t n, const int *perm, int *inv) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) inv[perm[s]] = (int)s;
}

__global__ void k_trip_jkey(int64_t ntrip, const int4 *trip, int *key) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntrip) key[t] = trip[t].z;
}

extern "C" int fbk_d2c(const DView &v, const DForce &f, int mask, hipStream_t st) {
  if (v.natoms == 0) return 0;
  if (fb_wide_species(v.sp))
    hipLaunchKernelGGL(k_d2c<FB_MOD>, dim3(v.natoms), dim3(WD2), 0, st, v, f, mask);
  else
    hipLaunchKernelGGL(k_d2c<FB_MO>, dim3(v.natoms), dim3(WD2), 0, st, v, f, mask);
  return (int)hipGetLastError();
}

extern "C" int fbk_d3c(const DView &v, const DForce &f, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  const dim3 grid((unsigned)((v.ntrip + T3D - 1) / T3D));
  if (fb_wide_species(v.sp))
    hipLaunchKernelGGL(k_d3c<FB_MOD>, grid, dim3(T3D), 0, st, v, f);
  else
    hipLaunchKernelGGL(k_d3c<FB_MO>, grid, dim3(T3D), 0, st, v, f);
  return (int)hipGetLastError();
}

extern "C" int fbk_f3_reduce(const DView &v, const DForce &f, hipStream_t st) {
  if (v.natoms == 0) return 0;
  hipLaunchKernelGGL(k_f3_reduce, dim3(v.natoms), dim3(64), 0, st, v, f);
  return (int)hipGetLastError();
}

extern "C" int fbk_f3_offsets(int natoms, int64_t ntrip, const int64_t *noff, const int *tpair_sorted,
                              const int *jkey_sorted, int64_t *ioff, int64_t *joff, hipStream_t st) {
  hipLaunchKernelGGL(k_f3_offsets, dim3((unsigned)((natoms + 1 + 255) / 256)), dim3(256), 0, st, natoms, ntrip,
                     noff, tpair_sorted, jkey_sorted, ioff, joff);
  return (int)hipGetLastError();
}

extern "C" int fbk_invert_perm(int64_t n, const int *perm, int *inv, hipStream_t st) {
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_invert_perm, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, perm, inv);
  return (int)hipGetLastError();
}

extern "C" int fbk_trip_jkey(int64_t ntrip, const int4 *trip, int *key, hipStream_t st) {
  if (ntrip == 0) return 0;
  hipLaunchKernelGGL(k_trip_jkey, dim3((unsigned)((ntrip + 255) / 256)), dim3(256), 0, st, ntrip, trip, key);
  return (int)hipGetLastError();
}

// ===========================================================================
// s/p species with canonical tables (the usual case): compile-time block
// shapes, closed-form Drotate, the Lagrange denominators of the host table,
// and the frame derivatives of one Cartesian component at a time, so that
// everything stays in registers.  The operations and their order are those of
// the generic kernels above (terms that the generic loops add as exact zeros
// are skipped), so the two paths agree to the last bit apart from signs of
// zero.
#include "fb_sp.h"
#include "fb_2c_dev.h"

#define PP_(k, m) pm[3 * (m) + (k)]
#define DD_(k, m) dpm[3 * (m) + (k)]

// one nonzero molecular element (m1, m2) of the p-p block: the three product-rule terms
__device__ __forceinline__ double drot_pp_term(double x, const double *pm, const double *dpm, int k, int l, int m1,
                                               int m2, double Mv, double DMv) {
  x = x + (DD_(k, m1) * PP_(l, m2)) * Mv;
  x = x + (PP_(k, m1) * DD_(l, m2)) * Mv;
  x = x + (PP_(k, m1) * PP_(l, m2)) * DMv;
  return x;
}

// Drotate of a 2c block, canonical ME order {ss, s-pz, pz-s, py-py, pz-pz, px-px}
template <int CLS>
__device__ __forceinline__ void drot2c_sp(const double *pm, const double *dpm, const double *F, const double *DM,
                                          double *X) {
  X[0] = DM[0];
  if (CLS == CLS_SPSP) {
#pragma unroll
    for (int l = 0; l < 3; l++) X[(1 + l) * 4] = DD_(l, 1) * F[1] + PP_(l, 1) * DM[1];
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = DD_(k, 1) * F[2] + PP_(k, 1) * DM[2];
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double x = 0.0;
        x = drot_pp_term(x, pm, dpm, k, l, 0, 0, F[3], DM[3]);
        x = drot_pp_term(x, pm, dpm, k, l, 1, 1, F[4], DM[4]);
        x = drot_pp_term(x, pm, dpm, k, l, 2, 2, F[5], DM[5]);
        X[(1 + l) * 4 + 1 + k] = x;
      }
  } else if (CLS == CLS_SSP) {
#pragma unroll
    for (int l = 0; l < 3; l++) X[1 + l] = DD_(l, 1) * F[1] + PP_(l, 1) * DM[1];
  } else if (CLS == CLS_SPS) {
#pragma unroll
    for (int k = 0; k < 3; k++) X[1 + k] = DD_(k, 1) * F[1] + PP_(k, 1) * DM[1];
  }
}

// Drotate of a 3c block, canonical ME order
// {ss, s-pz, pz-s, py-py, pz-pz, px-px, s-px, px-s, px-pz, pz-px} (s-sp: {ss, s-pz, s-px}; sp-s: {ss, pz-s, px-s})
template <int CLS>
__device__ __forceinline__ void drot3c_sp(const double *pm, const double *dpm, const double *F, const double *DM,
                                          double *X) {
  X[0] = DM[0];
  if (CLS == CLS_SPSP) {
#pragma unroll
    for (int l = 0; l < 3; l++) {
      double x = DD_(l, 1) * F[1];
      x = x + PP_(l, 1) * DM[1];
      x = x + DD_(l, 2) * F[6];
      x = x + PP_(l, 2) * DM[6];
      X[(1 + l) * 4] = x;
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double x = DD_(k, 1) * F[2];
      x = x + PP_(k, 1) * DM[2];
      x = x + DD_(k, 2) * F[7];
      x = x + PP_(k, 2) * DM[7];
      X[1 + k] = x;
    }
#pragma unroll
    for (int l = 0; l < 3; l++)
#pragma unroll
      for (int k = 0; k < 3; k++) {
        double x = 0.0;
        x = drot_pp_term(x, pm, dpm, k, l, 0, 0, F[3], DM[3]);
        x = drot_pp_term(x, pm, dpm, k, l, 1, 1, F[4], DM[4]);
        x = drot_pp_term(x, pm, dpm, k, l, 2, 1, F[8], DM[8]);
        x = drot_pp_term(x, pm, dpm, k, l, 1, 2, F[9], DM[9]);
        x = drot_pp_term(x, pm, dpm, k, l, 2, 2, F[5], DM[5]);
        X[(1 + l) * 4 + 1 + k] = x;
      }
  } else if (CLS == CLS_SSP) {
#pragma unroll
    for (int l = 0; l < 3; l++) {
      double x = DD_(l, 1) * F[1];
      x = x + PP_(l, 1) * DM[1];
      x = x + DD_(l, 2) * F[2];
      x = x + PP_(l, 2) * DM[2];
      X[1 + l] = x;
    }
  } else if (CLS == CLS_SPS) {
#pragma unroll
    for (int k = 0; k < 3; k++) {
      double x = DD_(k, 1) * F[1];
      x = x + PP_(k, 1) * DM[1];
      x = x + DD_(k, 2) * F[2];
      x = x + PP_(k, 2) * DM[2];
      X[1 + k] = x;
    }
  }
}
#undef PP_
#undef DD_

// 7-point window, products and derivative weights of one grid (M_Fdata_2c.f90:470-540)
struct DLag7 {
  int grid, ileft;
  double prod[7], pdenom[7], xsum[7];
};

__device__ __forceinline__ void dlag7_build(const DView &v, int grid, double x, DLag7 &L) {
  const DGrid2c gr = grid_scalar(v, grid);
  int imid = (int)(x / gr.dx) + 1;
  int ileft = imid - 3;
  if (ileft < 1) ileft = 1;
  else if (imid + 3 > gr.nx) ileft = gr.nx - 6;
  double xx[7], r[7][7];
#pragma unroll
  for (int k = 0; k < 7; k++) xx[k] = (double)(ileft + k - 1) * gr.dx;
  // pden = 1/(xx_is - xx_ip) with the reference's division; the pairs ip < is
  // by exact negation (fl(a - b) = -fl(b - a))
#pragma unroll
  for (int is = 0; is < 7; is++)
#pragma unroll
    for (int ip = is + 1; ip < 7; ip++) {
      r[is][ip] = 1.0 / (xx[is] - xx[ip]);
      r[ip][is] = -r[is][ip];
    }
#pragma unroll
  for (int is = 0; is < 7; is++) {
    double p = 1.0, q = 1.0;
#pragma unroll
    for (int ip = 0; ip < 7; ip++)
      if (ip != is) {
        q = q * r[is][ip];
        p = p * (x - xx[ip]) * r[is][ip];
      }
    L.prod[is] = p;
    L.pdenom[is] = q;
  }
#pragma unroll
  for (int ipt = 0; ipt < 7; ipt++) {
    double xs = 0.0;
#pragma unroll
    for (int jpt = 0; jpt < 7; jpt++) {
      if (ipt == jpt) continue;
      double xp = 1.0;
#pragma unroll
      for (int kpt = 0; kpt < 7; kpt++)
        if (kpt != jpt && kpt != ipt) xp = xp * (x - xx[kpt]);
      xs = xs + xp;
    }
    L.xsum[ipt] = xs;
  }
  L.grid = grid;
  L.ileft = ileft;
}

// values and x-derivatives of one canonical 2c cell (getDMEs_Fdata_2c)
template <int NME>
__device__ __forceinline__ int dinterp2c_sp(const DView &v, int sp2, int slot, int isub, double x, DLag7 &L, double *F,
                                            double *dF) {
  DCell2c c;
  if (cell_scalar(v, sp2, slot, isub, c) < 0) return ST_NO_TABLE;
  if (x < 0.0) return ST_X_NEG;
  if (x > c.xmax * (1.0 + 1.0e-5)) return ST_X_RANGE;
  const double *tab = v.tab2c + c.off;
  if (x < 1.0e-5) {
#pragma unroll
    for (int m = 0; m < NME; m++) { F[m] = tab[m]; dF[m] = 0.0; }
    return 0;
  }
  if (c.grid != L.grid) dlag7_build(v, c.grid, x, L);
  const double *base = tab + (int64_t)(L.ileft - 1) * NME;
#pragma unroll
  for (int m = 0; m < NME; m++) { F[m] = 0.0; dF[m] = 0.0; }
#pragma unroll
  for (int is = 0; is < 7; is++)
#pragma unroll
    for (int m = 0; m < NME; m++) F[m] = F[m] + base[is * NME + m] * L.prod[is];
#pragma unroll
  for (int is = 0; is < 7; is++)
#pragma unroll
    for (int m = 0; m < NME; m++) dF[m] = dF[m] + base[is * NME + m] * L.pdenom[is] * L.xsum[is];
  return 0;
}

// d/dr1 of one rotated 2c interaction, component ix, contracted: f -= rho [dq] X P_eq2
template <int CLS>
__device__ __forceinline__ void dcontract2c_sp(const double *pm, const double *deps, const double *eta, double z,
                                               double zcut, const double *F, const double *dF, const double *rho,
                                               double dq, bool use_dq, double *f) {
  constexpr int NBL = ((CLS & 2) ? 4 : 1) * ((CLS & 1) ? 4 : 1);
  constexpr int NME = nme2c_of(CLS);
  double DM[NME], X[16], dpm[9];
#pragma unroll
  for (int ix = 0; ix < 3; ix++) {
#pragma unroll
    for (int m = 0; m < NME; m++) DM[m] = (z > zcut) ? -eta[ix] * dF[m] : 0.0;
    d_dpmat(deps, ix + 1, dpm);
    drot2c_sp<CLS>(pm, dpm, F, DM, X);
#pragma unroll
    for (int i = 0; i < NBL; i++) {
      if (use_dq) f[ix] = f[ix] - rho[i] * dq * X[i] * FB_P_EQ2;
      else f[ix] = f[ix] - rho[i] * X[i] * FB_P_EQ2;
    }
  }
}

// PART 0: the Dblock of interaction d (0 S, 1 T, 2 dipole_z) into dst
// (Dblock(3, norb_mu, norb_nu) column-major); PART 1: the vna_ontop force of
// the pair; PART 2: the neutral vna_atom term; PART 3: the charged vna_atom
// terms, added to what PART 2 stored
template <int CLS, int PART>
__device__ __forceinline__ int pair_d2c_sp(const DView &v, const DForce &Fo, int mask, int64_t p, int ia, int ja,
                                           int in1, int in2, bool self, double z, const double *pm,
                                           const double *deps, const double *eta, int d = 0, double *dst = nullptr) {
  constexpr int NMU = (CLS & 2) ? 4 : 1, NNU = (CLS & 1) ? 4 : 1, NBL = NMU * NNU;
  constexpr int NME = nme2c_of(CLS);
  constexpr int CLSA = (CLS & 2) ? CLS_SPSP : CLS_SS;   // atom term on iatom's orbitals
  constexpr int NMEA = nme2c_of(CLSA), NA = NMU * NMU;
  int status = 0;
  DLag7 L;
  L.grid = -1;
  L.ileft = 1;
  double F[6], dF[6], X[16], DM[6], dpm[9];
  const int64_t bo = v.boff[p];
  if (PART == 0) {
    const int slot = d == 0 ? SLOT_S : (d == 1 ? SLOT_T : SLOT_DIP);
    int st = dinterp2c_sp<NME>(v, in1 | (in2 << 16), slot, 0, z, L, F, dF);
    status |= st;
    if (st) return status;
#pragma unroll
    for (int ix = 0; ix < 3; ix++) {
#pragma unroll
      for (int m = 0; m < NME; m++) DM[m] = (z > 1.0e-3) ? -eta[ix] * dF[m] : 0.0;
      d_dpmat(deps, ix + 1, dpm);
      drot2c_sp<CLS>(pm, dpm, F, DM, X);
#pragma unroll
      for (int i = 0; i < NBL; i++) dst[ix + 3 * i] = X[i];
    }
    return status;
  }
  if (PART == 1) {
    // vna_ontop (M_Dassemble_2c_DOGS.f90:679-863)
    double fo[3] = {0.0, 0.0, 0.0};
    if (!self) {
      double rho[NBL];
#pragma unroll
      for (int i = 0; i < NBL; i++) rho[i] = Fo.rho[bo + i];
#pragma unroll
      for (int side = 0; side < 2; side++) {
        const int slot = side ? SLOT_ONR : SLOT_ONL;
        const int qa = side ? ja : ia;
        const int inq = side ? in2 : in1;
        int st = dinterp2c_sp<NME>(v, in1 | (in2 << 16), slot, 0, z, L, F, dF);
        status |= st;
        if (!st) dcontract2c_sp<CLS>(pm, deps, eta, z, 1.0e-3, F, dF, rho, 0.0, false, fo);
        for (int isorp = 1; isorp <= v.sp.nssh[inq - 1]; isorp++) {
          const double dQ = v.dQ[qa * FB_MAXSH + isorp - 1];
          if (dQ == 0.0) continue;
          st = dinterp2c_sp<NME>(v, in1 | (in2 << 16), slot, isorp, z, L, F, dF);
          status |= st;
          if (!st) dcontract2c_sp<CLS>(pm, deps, eta, z, 1.0e-3, F, dF, rho, dQ, true, fo);
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 3; q++) Fo.fontop[3 * p + q] = fo[q];
    return status;
  }
  {
    // vna_atom (M_Dassemble_2c_DOGS.f90:866-1199): PART 2 the neutral term,
    // PART 3 continues the sum with the charged shells of jatom
    double fa[3] = {0.0, 0.0, 0.0};
    const int64_t pself = v.noff[ia] + v.neigh_self[ia] - 1;
    double rs[NA];
#pragma unroll
    for (int i = 0; i < NA; i++) rs[i] = Fo.rho[v.boff[pself] + i];
    int st = 0;
    if (PART == 2) {
      st = dinterp2c_sp<NMEA>(v, in1 | (in2 << 16), SLOT_ATOM, 0, z, L, F, dF);
      status |= st;
      if (!st) dcontract2c_sp<CLSA>(pm, deps, eta, z, 1.0e-3, F, dF, rs, 0.0, false, fa);
    }
    bool charged = false;
    if (PART == 3)
      for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) charged |= (v.dQ[ja * FB_MAXSH + isorp - 1] != 0.0);
    if (charged) {
#pragma unroll
      for (int q = 0; q < 3; q++) fa[q] = Fo.fatom[3 * p + q];
      const double rend = v.sp.rcmin[in1 - 1] + v.sp.rcmin[in2 - 1];
      const double smooth = d_smoother(z, rend, 0.8), Dsmooth = d_Dsmoother(z, rend, 0.8);
      const double *sself = v.mat[FB_MAT_S] + v.boff[pself];
      for (int isorp = 1; isorp <= v.sp.nssh[in2 - 1]; isorp++) {
        const double dQ = v.dQ[ja * FB_MAXSH + isorp - 1];
        if (dQ == 0.0) continue;
        st = dinterp2c_sp<NMEA>(v, in1 | (in2 << 16), SLOT_ATOM, isorp, z, L, F, dF);
        status |= st;
        if (st) continue;
        double bx[16];
        rot2c_sp<double>(CLSA, pm, F, bx);
#pragma unroll
        for (int ix = 0; ix < 3; ix++) {
#pragma unroll
          for (int m = 0; m < NMEA; m++) DM[m] = (z > 1.0e-4) ? -eta[ix] * dF[m] : 0.0;
          d_dpmat(deps, ix + 1, dpm);
          drot2c_sp<CLSA>(pm, dpm, F, DM, X);
#pragma unroll
          for (int i = 0; i < NA; i++) {
            const double emnpl = (z > 1.0e-4) ? sself[i] / z : 0.0;
            const double vdem = (z > 1.0e-4) ? eta[ix] * sself[i] / (z * z) : 0.0;
            fa[ix] = fa[ix] - rs[i] * FB_P_EQ2 * dQ *
                                  (smooth * X[i] - eta[ix] * Dsmooth * bx[i] + (1.0 - smooth) * vdem +
                                   eta[ix] * Dsmooth * emnpl);
          }
        }
      }
    }
    if (PART == 2 || charged)
#pragma unroll
      for (int q = 0; q < 3; q++) Fo.fatom[3 * p + q] = fa[q];
  }
  return status;
}

#define WD2S 128

// Three kernels over the pairs, one lane per pair.  PART 0 forms the Dblocks
// of S, T and dipole_z (mask bits 1, 2, 4) one interaction at a time into
// LDS, packed as in the output, and the workgroup stores the contiguous block
// range of its pairs with consecutive lanes on consecutive addresses (a lane
// writing its own 48-value block would leave partly written cache lines
// behind).  PARTs 1 to 3 form the vna_ontop force and the neutral and charged
// vna_atom terms (bit 8; PART 3 only for builds with shell charges).
// Splitting the work keeps each kernel's registers below the spill point.
template <int PART>
__global__ __launch_bounds__(WD2S) void k_d2c_sp(DView v, DForce Fo, int mask) {
  __shared__ double sD[PART == 0 ? 3 * 16 * WD2S : 1];
  const int64_t p0 = (int64_t)blockIdx.x * WD2S, pe = min(p0 + (int64_t)WD2S, v.npairs);
  const int64_t p = p0 + threadIdx.x;
  const bool act = p < pe;
  int ia = 0, ja = 0, in1 = 1, in2 = 1, cls = 0;
  bool self = false;
  double z = 0.0, pm[9], deps[27], eta[3];
  if (act) {
    ia = v.pair_atom[p];
    const int s = v.atom_struct[ia];
    ja = v.st[s].atom_off + v.neigh_j[p] - 1;
    const int mb = v.neigh_b[p];
    in1 = v.imass[ia];
    in2 = v.imass[ja];
    const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
    double r1[3], r2[3], sighat[3], eps[9];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      r1[q] = v.ratom[3 * ia + q];
      r2[q] = v.ratom[3 * ja + q] + xls[3 * mb + q];
    }
    z = d_distance(r1, r2);
    if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
    else {
#pragma unroll
      for (int q = 0; q < 3; q++) sighat[q] = (r2[q] - r1[q]) / z;
    }
    d_epsilon(r2, sighat, eps);
    d_Depsilon_2c(r1, r2, z, eps, deps);
    d_pmat<double>(eps, pm);
#pragma unroll
    for (int q = 0; q < 3; q++) eta[q] = eps[6 + q];
    self = (ja == ia && mb == 0);
    cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
  }
  int status = 0;
  if (PART == 0) {
    const int64_t b0 = v.boff[p0], nb = 3 * (v.boff[pe] - b0);
    double *const dst = act ? sD + 3 * (v.boff[p] - b0) : nullptr;
    for (int d = 0; d < 3; d++) {
      if (!(mask & (1 << d))) continue;
      if (act) {
        switch (cls) {
          case 3: status |= pair_d2c_sp<3, 0>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta, d, dst); break;
          case 2: status |= pair_d2c_sp<2, 0>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta, d, dst); break;
          case 1: status |= pair_d2c_sp<1, 0>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta, d, dst); break;
          default: status |= pair_d2c_sp<0, 0>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta, d, dst); break;
        }
      }
      __syncthreads();
      double *const out = Fo.dmat[d] + 3 * b0;
      for (int64_t k = threadIdx.x; k < nb; k += WD2S) out[k] = sD[k];
      __syncthreads();
    }
  } else if (act) {
    switch (cls) {
      case 3: status = pair_d2c_sp<3, PART>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta); break;
      case 2: status = pair_d2c_sp<2, PART>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta); break;
      case 1: status = pair_d2c_sp<1, PART>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta); break;
      default: status = pair_d2c_sp<0, PART>(v, Fo, mask, p, ia, ja, in1, in2, self, z, pm, deps, eta); break;
    }
  }
  if (status) atomicOr(v.status, status);
}

// ---------------------------------------------------------------- 3c, s/p

// Depsilon_3c restricted to one Cartesian component ix (1-based): the
// (ix, :, :) slices of dera3 and der13 depend on that component only
struct D3Scal {
  double crossmag, r2mag2, r1mag2, ratmmag2, r1dotr2, r2dotratm, r1dotratm;
  bool zero;
};

__device__ __forceinline__ void d3_scal(const double *r1, const double *r2, const double *r21, const double *ratm,
                                        const double *rnabc, D3Scal &S) {
  double ca0 = r21[1] * rnabc[2] - r21[2] * rnabc[1];
  double ca1 = r21[2] * rnabc[0] - r21[0] * rnabc[2];
  double ca2 = r21[0] * rnabc[1] - r21[1] * rnabc[0];
  S.crossmag = sqrt(ca0 * ca0 + ca1 * ca1 + ca2 * ca2);
  S.zero = fabs(S.crossmag) < 1.0e-3;
  S.r2mag2 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
  S.r1mag2 = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
  S.ratmmag2 = ratm[0] * ratm[0] + ratm[1] * ratm[1] + ratm[2] * ratm[2];
  S.r1dotr2 = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  S.r2dotratm = r2[0] * ratm[0] + r2[1] * ratm[1] + r2[2] * ratm[2];
  S.r1dotratm = r1[0] * ratm[0] + r1[1] * ratm[1] + r1[2] * ratm[2];
}

// dA[imu + 3*(k-1)] = dera3(ix, imu, k), dB = der13(ix, imu, k)
__device__ __forceinline__ void d3_slice(const double *r1, const double *r2, double z, const double *ratm,
                                         const double *eps3, const D3Scal &S, int ix, double *dA, double *dB) {
#define E3(i, j) eps3[3 * ((j) - 1) + (i) - 1]
#define SA(imu, k) dA[((imu) - 1) + 3 * ((k) - 1)]
#define SB(imu, k) dB[((imu) - 1) + 3 * ((k) - 1)]
#pragma unroll
  for (int i = 0; i < 9; i++) { dA[i] = 0.0; dB[i] = 0.0; }
  if (S.zero) return;
  const double cm = S.crossmag;
  const double ra = sel3(ratm, ix - 1), r1x = sel3(r1, ix - 1), r2x = sel3(r2, ix - 1);
  const double e3x = ix == 1 ? E3(1, 3) : (ix == 2 ? E3(2, 3) : E3(3, 3));
#pragma unroll
  for (int imu = 1; imu <= 3; imu++) {
    double sum = d_levi(ix, imu, 1) * (r2[0] - r1[0]) + d_levi(ix, imu, 2) * (r2[1] - r1[1]) +
                 d_levi(ix, imu, 3) * (r2[2] - r1[2]);
    SA(imu, 2) = (1.0 / cm) * (sum - (E3(imu, 2) / cm) *
                                         ((S.r1mag2 + S.r2mag2 - 2.0 * S.r1dotr2) * ra +
                                          (S.r1dotr2 - S.r1mag2 - S.r2dotratm + S.r1dotratm) * r2x +
                                          (S.r1dotr2 - S.r2mag2 + S.r2dotratm - S.r1dotratm) * r1x));
  }
  SA(1, 1) = E3(3, 3) * SA(2, 2) - E3(2, 3) * SA(3, 2);
  SA(2, 1) = E3(1, 3) * SA(3, 2) - E3(3, 3) * SA(1, 2);
  SA(3, 1) = E3(2, 3) * SA(1, 2) - E3(1, 3) * SA(2, 2);
#pragma unroll
  for (int imu = 1; imu <= 3; imu++) {
    SB(imu, 3) = (1.0 / z) * (E3(imu, 3) * e3x - d_delk(imu, ix));
    double sum = d_levi(ix, imu, 1) * (ratm[0] - r2[0]) + d_levi(ix, imu, 2) * (ratm[1] - r2[1]) +
                 d_levi(ix, imu, 3) * (ratm[2] - r2[2]);
    SB(imu, 2) = (1.0 / cm) * (sum - (E3(imu, 2) / cm) *
                                         ((S.r2dotratm - S.r1dotratm + S.r1dotr2 - S.r2mag2) * ra +
                                          (S.r2dotratm + S.r1dotratm - S.r1dotr2 - S.ratmmag2) * r2x +
                                          (S.ratmmag2 - 2.0 * S.r2dotratm + S.r2mag2) * r1x));
  }
  SB(1, 1) = E3(3, 3) * SB(2, 2) - E3(3, 2) * SB(2, 3) - E3(2, 3) * SB(3, 2) + E3(2, 2) * SB(3, 3);
  SB(2, 1) = E3(1, 3) * SB(3, 2) - E3(1, 2) * SB(3, 3) - E3(3, 3) * SB(1, 2) + E3(3, 2) * SB(1, 3);
  SB(3, 1) = E3(2, 3) * SB(1, 2) - E3(2, 2) * SB(1, 3) - E3(1, 3) * SB(2, 2) + E3(1, 2) * SB(2, 3);
#undef E3
#undef SA
#undef SB
}

// Dtwister p part from a slice s[imu + 3*(k-1)] = deps(ix, imu, k)
__device__ __forceinline__ void dpmat_slice(const double *s, double *dpm) {
#define SL(imu, k) s[((imu) - 1) + 3 * ((k) - 1)]
  dpm[0] = SL(2, 2); dpm[3] = SL(2, 3); dpm[6] = SL(2, 1);
  dpm[1] = SL(3, 2); dpm[4] = SL(3, 3); dpm[7] = SL(3, 1);
  dpm[2] = SL(1, 2); dpm[5] = SL(1, 3); dpm[8] = SL(1, 1);
#undef SL
}

// getDMEs_Fdata_3c of one canonical group: values and d/dcos, d/dx, d/dy
template <int NME>
__device__ __forceinline__ int dgetD3c_sp(const DView &v, int gid, double x, double y, double cost, double *F,
                                          double *DpF, double *DxF, double *DyF) {
  if (gid < 0) return ST_NO_TABLE;
  const DGroup3c g = v.grp3c[gid];
  const int nx = g.nx, ny = g.ny;
  // Legendre P0..P4 and derivatives, selected per theta below (no private array)
  const double cost2 = cost * cost;
  const double sint = sqrt(fmax(1.0e-5, 1.0 - cost2));
  const double p2 = 0.5 * (3.0 * cost2 - 1.0);
  const double p3 = 0.5 * cost * (5.0 * cost2 - 3.0);
  const double p4 = 0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2));
  const double dp2 = 3.0 * cost;
  const double dp3 = (15.0 * cost2 - 3.0) / 2.0;
  const double dp4 = 0.5 * cost * (35.0 * cost2 - 15.0);
  const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  const double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
  const double *tab = v.tab3c + g.off;
  const int64_t tstride = (int64_t)ny * nx * NME;
#define T3S(th, i, j, m) tab[(th) * tstride + ((int64_t)((j) - 1) * nx + ((i) - 1)) * NME + (m)]
  const double gx = (T3S(0, ix + 1, iy, 0) - T3S(0, ix, iy, 0)) / dx;
  const double gy = (T3S(0, ix, iy + 1, 0) - T3S(0, ix, iy, 0)) / dy;
  const bool three = (gx * gx + gy * gy) < 1.0e-6;
#pragma unroll
  for (int m = 0; m < NME; m++) { F[m] = 0.0; DpF[m] = 0.0; DxF[m] = 0.0; DyF[m] = 0.0; }
  for (int th = 0; th < FB_NTHETA; th++) {
    const double pt = th == 0 ? 1.0 : (th == 1 ? cost : (th == 2 ? p2 : (th == 3 ? p3 : p4)));
    const double dpt = th == 0 ? 0.0 : (th == 1 ? 1.0 : (th == 2 ? dp2 : (th == 3 ? dp3 : dp4)));
#pragma unroll
    for (int m = 0; m < NME; m++) {
      if (three) {
        const double interim =
            ((1.0 - px - py) * T3S(th, ix, iy, m) + px * T3S(th, ix + 1, iy, m)) + py * T3S(th, ix, iy + 1, m);
        F[m] = F[m] + pt * interim;
        DpF[m] = DpF[m] + dpt * interim;
        DxF[m] = DxF[m] + pt * gx;
        DyF[m] = DyF[m] + pt * gy;
      } else {
        double gk[4], gp[4];
#pragma unroll
        for (int k = -1; k <= 2; k++) {
          double f1m1 = T3S(th, ix + k, iy - 1, m);
          double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
          double f0p3 = 3.0 * T3S(th, ix + k, iy, m), f0p6 = 2.0 * f0p3;
          double f1p3 = 3.0 * T3S(th, ix + k, iy + 1, m), f1p6 = 2.0 * f1p3;
          double f2p1 = T3S(th, ix + k, iy + 2, m);
          double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
          double bb2 = f1m3 - f0p6 + f1p3;
          double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
          double bb0 = f0p6;
          gk[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
          gp[k + 1] = ((3.0 * bb3 * py + 2.0 * bb2) * py + bb1);
        }
        double f1m1 = gk[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
        double f0p3 = 3.0 * gk[1], f0p6 = 2.0 * f0p3;
        double f1p3 = 3.0 * gk[2], f1p6 = 2.0 * f1p3;
        double f2p1 = gk[3];
        double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        double bb2 = f1m3 - f0p6 + f1p3;
        double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        double bb0 = f0p6;
        const double interim = (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
        F[m] = F[m] + pt * interim;
        DpF[m] = DpF[m] + dpt * interim;
        DxF[m] = DxF[m] + pt * ((3.0 * bb3 * px + 2.0 * bb2) * px + bb1) / (36.0 * dx);
        f1m1 = gp[0]; f1m2 = 2.0 * f1m1; f1m3 = 3.0 * f1m1;
        f0p3 = 3.0 * gp[1]; f0p6 = 2.0 * f0p3;
        f1p3 = 3.0 * gp[2]; f1p6 = 2.0 * f1p3;
        f2p1 = gp[3];
        bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        bb2 = f1m3 - f0p6 + f1p3;
        bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        bb0 = f0p6;
        DyF[m] = DyF[m] + pt * (((bb3 * px + bb2) * px + bb1) * px + bb0) / (36.0 * dy);
      }
    }
  }
#undef T3S
  constexpr int NM0 = (NME == 10) ? 6 : (NME == 3 ? 2 : 1);
#pragma unroll
  for (int m = NM0; m < NME; m++) {
    DpF[m] = DpF[m] * sint - cost * F[m] / sint;
    F[m] = F[m] * sint;
    DxF[m] = DxF[m] * sint;
    DyF[m] = DyF[m] * sint;
  }
  return 0;
}

struct DGeo3s {
  double r1[3], r2[3], rna[3], r21[3], rnabc[3], sighat[3], rhat[3], amt[3], bmt[3];
  double z, x, cost, d13, d23;
  double eps[9];
};

__device__ __forceinline__ void d_dgeo3s(const DView &v, int4 tr, bool charged, DGeo3s &G) {
  const int ial = tr.x, gia = tr.y, gja = tr.z;
  const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
  const int s = v.atom_struct[ial];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
#pragma unroll
  for (int k = 0; k < 3; k++) {
    G.r1[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
    G.r2[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
    G.rna[k] = v.ratom[3 * ial + k];
    G.r21[k] = G.r2[k] - G.r1[k];
  }
  G.z = d_distance(G.r1, G.r2);
  if (G.z < 1.0e-5) { G.sighat[0] = 0.0; G.sighat[1] = 0.0; G.sighat[2] = 1.0; }
  else {
#pragma unroll
    for (int k = 0; k < 3; k++) G.sighat[k] = G.r21[k] / G.z;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) G.rnabc[k] = G.rna[k] - (G.r1[k] + G.r21[k] / 2.0);
  G.x = sqrt(G.rnabc[0] * G.rnabc[0] + G.rnabc[1] * G.rnabc[1] + G.rnabc[2] * G.rnabc[2]);
  G.d13 = d_distance(G.rna, G.r1);
  G.d23 = d_distance(G.rna, G.r2);
  if (G.x < 1.0e-5) { G.rhat[0] = 0.0; G.rhat[1] = 0.0; G.rhat[2] = 0.0; }
  else if (charged) {
#pragma unroll
    for (int k = 0; k < 3; k++) G.rhat[k] = (G.rna[k] - 0.5 * (G.r1[k] + G.r2[k])) / G.x;
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) G.rhat[k] = G.rnabc[k] / G.x;
  }
  G.cost = G.sighat[0] * G.rhat[0] + G.sighat[1] * G.rhat[1] + G.sighat[2] * G.rhat[2];
  d_epsilon(G.rhat, G.sighat, G.eps);
#pragma unroll
  for (int k = 0; k < 3; k++) {
    G.amt[k] = (G.x > 1.0e-5) ? (G.sighat[k] - G.cost * G.rhat[k]) / G.x : 0.0;
    G.bmt[k] = (G.cost * G.sighat[k] - G.rhat[k]) / G.z;
  }
}

// element (row, col) of the rotated derivative of a 3c block: drot3c_sp one
// element at a time (same terms, same order), so that a triplet never holds a
// whole rotated block
template <int CLS>
__device__ __forceinline__ double drot3c_el(int row, int col, const double *pm, const double *dpm, const double *F,
                                            const double *DM) {
#define PP_(k, m) pm[3 * (m) + (k)]
#define DD_(k, m) dpm[3 * (m) + (k)]
  if (row == 0 && col == 0) return DM[0];
  if (CLS == CLS_SPSP) {
    if (row == 0) {
      const int l = col - 1;
      double x = DD_(l, 1) * F[1];
      x = x + PP_(l, 1) * DM[1];
      x = x + DD_(l, 2) * F[6];
      return x + PP_(l, 2) * DM[6];
    }
    if (col == 0) {
      const int k = row - 1;
      double x = DD_(k, 1) * F[2];
      x = x + PP_(k, 1) * DM[2];
      x = x + DD_(k, 2) * F[7];
      return x + PP_(k, 2) * DM[7];
    }
    const int k = row - 1, l = col - 1;
    double x = 0.0;
    x = drot_pp_term(x, pm, dpm, k, l, 0, 0, F[3], DM[3]);
    x = drot_pp_term(x, pm, dpm, k, l, 1, 1, F[4], DM[4]);
    x = drot_pp_term(x, pm, dpm, k, l, 2, 1, F[8], DM[8]);
    x = drot_pp_term(x, pm, dpm, k, l, 1, 2, F[9], DM[9]);
    return drot_pp_term(x, pm, dpm, k, l, 2, 2, F[5], DM[5]);
  }
  if (CLS == CLS_SSP) {
    const int l = col - 1;
    double x = DD_(l, 1) * F[1];
    x = x + PP_(l, 1) * DM[1];
    x = x + DD_(l, 2) * F[2];
    return x + PP_(l, 2) * DM[2];
  }
  if (CLS == CLS_SPS) {
    const int k = row - 1;
    double x = DD_(k, 1) * F[1];
    x = x + PP_(k, 1) * DM[1];
    x = x + DD_(k, 2) * F[2];
    return x + PP_(k, 2) * DM[2];
  }
  return 0.0;
}

#undef PP_
#undef DD_

// ---------------------------------------------------------------------------
// Four lanes per triplet.  Phase A: lane q interpolates matrix elements q, q+4,
// q+8 of the group (value and its cos, x and y derivatives) into the
// triplet's LDS row.  Phase B: lane q < 3 takes Cartesian component q: the
// frame-derivative slices of that component, f3naMa / f3naMb, the two rotated
// derivative blocks element by element, and the contraction with the density
// block, so that no lane loops over components and every private value has a
// compile-time index.  Lane 3 idles in phase B.
#define T3Q 64                 // lanes per workgroup: 16 triplets
#define SIROW (4 * NME3F + 1)  // LDS row per triplet (odd stride: no bank sharing)

// getDMEs_Fdata_3c of elements q, q+4, q+8 of one canonical group into row
__device__ inline int stage_q(const DView &v, int gid, int nme, double x, double y, double cost, double *row, int q) {
  if (gid < 0) return ST_NO_TABLE;
  const DGroup3c g = v.grp3c[gid];
  const int nx = g.nx, ny = g.ny;
  const double cost2 = cost * cost;
  const double sint = sqrt(fmax(1.0e-5, 1.0 - cost2));
  const double p2 = 0.5 * (3.0 * cost2 - 1.0);
  const double p3 = 0.5 * cost * (5.0 * cost2 - 3.0);
  const double p4 = 0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2));
  const double dp2 = 3.0 * cost;
  const double dp3 = (15.0 * cost2 - 3.0) / 2.0;
  const double dp4 = 0.5 * cost * (35.0 * cost2 - 15.0);
  const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  const double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
  const double *tab = v.tab3c + g.off;
  const int64_t tstride = (int64_t)ny * nx * nme;
#define T3Q_(th, i, j, m) tab[(th) * tstride + ((int64_t)((j) - 1) * nx + ((i) - 1)) * nme + (m)]
  const double gx = (T3Q_(0, ix + 1, iy, 0) - T3Q_(0, ix, iy, 0)) / dx;
  const double gy = (T3Q_(0, ix, iy + 1, 0) - T3Q_(0, ix, iy, 0)) / dy;
  const bool three = (gx * gx + gy * gy) < 1.0e-6;
  const int nm0 = (nme == 10) ? 6 : (nme == 3 ? 2 : 1);   // canonical order: mvalue = 1 elements last
  for (int m = q; m < nme; m += 4) {
    double F = 0.0, DpF = 0.0, DxF = 0.0, DyF = 0.0;
    for (int th = 0; th < FB_NTHETA; th++) {
      const double pt = th == 0 ? 1.0 : (th == 1 ? cost : (th == 2 ? p2 : (th == 3 ? p3 : p4)));
      const double dpt = th == 0 ? 0.0 : (th == 1 ? 1.0 : (th == 2 ? dp2 : (th == 3 ? dp3 : dp4)));
      if (three) {
        const double interim =
            ((1.0 - px - py) * T3Q_(th, ix, iy, m) + px * T3Q_(th, ix + 1, iy, m)) + py * T3Q_(th, ix, iy + 1, m);
        F = F + pt * interim;
        DpF = DpF + dpt * interim;
        DxF = DxF + pt * gx;
        DyF = DyF + pt * gy;
      } else {
        double gk[4], gp[4];
#pragma unroll
        for (int k = -1; k <= 2; k++) {
          double f1m1 = T3Q_(th, ix + k, iy - 1, m);
          double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
          double f0p3 = 3.0 * T3Q_(th, ix + k, iy, m), f0p6 = 2.0 * f0p3;
          double f1p3 = 3.0 * T3Q_(th, ix + k, iy + 1, m), f1p6 = 2.0 * f1p3;
          double f2p1 = T3Q_(th, ix + k, iy + 2, m);
          double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
          double bb2 = f1m3 - f0p6 + f1p3;
          double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
          double bb0 = f0p6;
          gk[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
          gp[k + 1] = ((3.0 * bb3 * py + 2.0 * bb2) * py + bb1);
        }
        double f1m1 = gk[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
        double f0p3 = 3.0 * gk[1], f0p6 = 2.0 * f0p3;
        double f1p3 = 3.0 * gk[2], f1p6 = 2.0 * f1p3;
        double f2p1 = gk[3];
        double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        double bb2 = f1m3 - f0p6 + f1p3;
        double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        double bb0 = f0p6;
        const double interim = (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
        F = F + pt * interim;
        DpF = DpF + dpt * interim;
        DxF = DxF + pt * ((3.0 * bb3 * px + 2.0 * bb2) * px + bb1) / (36.0 * dx);
        f1m1 = gp[0]; f1m2 = 2.0 * f1m1; f1m3 = 3.0 * f1m1;
        f0p3 = 3.0 * gp[1]; f0p6 = 2.0 * f0p3;
        f1p3 = 3.0 * gp[2]; f1p6 = 2.0 * f1p3;
        f2p1 = gp[3];
        bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        bb2 = f1m3 - f0p6 + f1p3;
        bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        bb0 = f0p6;
        DyF = DyF + pt * (((bb3 * px + bb2) * px + bb1) * px + bb0) / (36.0 * dy);
      }
    }
#undef T3Q_
    if (m >= nm0) {
      DpF = DpF * sint - cost * F / sint;
      F = F * sint;
      DxF = DxF * sint;
      DyF = DyF * sint;
    }
    row[m] = F;
    row[nme + m] = DpF;
    row[2 * nme + m] = DxF;
    row[3 * nme + m] = DyF;
  }
  return 0;
}

// (x, z, cos theta) of a triplet with the expressions of d_dgeo3s, without the
// frame: what the interpolation needs
__device__ __forceinline__ void xzc3(const DView &v, int4 tr, bool charged, double &x, double &z, double &cost) {
  const int ial = tr.x, gia = tr.y, gja = tr.z;
  const int ib = tr.w & 0xffff, jb = (tr.w >> 16) & 0xffff;
  const int s = v.atom_struct[ial];
  const double *xls = v.xl + (int64_t)s * FB_NIMG * 3;
  double r1[3], r2[3], rna[3], r21[3], sighat[3], rnabc[3], rhat[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    r1[k] = v.ratom[3 * gia + k] + xls[3 * ib + k];
    r2[k] = v.ratom[3 * gja + k] + xls[3 * jb + k];
    rna[k] = v.ratom[3 * ial + k];
    r21[k] = r2[k] - r1[k];
  }
  z = d_distance(r1, r2);
  if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else {
#pragma unroll
    for (int k = 0; k < 3; k++) sighat[k] = r21[k] / z;
  }
#pragma unroll
  for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1[k] + r21[k] / 2.0);
  x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
  if (x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
  else if (charged) {
#pragma unroll
    for (int k = 0; k < 3; k++) rhat[k] = (rna[k] - 0.5 * (r1[k] + r2[k])) / x;
  } else {
#pragma unroll
    for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] / x;
  }
  cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
}

// ---------------------------------------------------------------------------
// Mixed precision (fb_set_precision FB_PREC_MIXED): the interpolation of phase
// A in FP32 over the FP32 slab [iy][ix][ME][theta] -- cubic Lagrange weights in
// x and y and their derivatives (the Horner form of M_Fdata_3c.f90:636-713 is
// this polynomial times 36), the Legendre fold of the value and of its cos theta
// derivative at every stencil point.  The cell, the method, the interpolation
// arguments and everything of phase B stay binary64.

__device__ __forceinline__ __amdgpu_buffer_rsrc_t slab32(const float *tab, int64_t nfloats) {
  return __builtin_amdgcn_make_buffer_rsrc((void *)tab, (short)0, (int)(nfloats * 4), 0x00020000);
}

// sum_theta P_theta T_theta and sum_theta dP_theta T_theta of one (point, ME); off in bytes
__device__ __forceinline__ void fold2(__amdgpu_buffer_rsrc_t r, unsigned off, const float *P, const float *dP,
                                      float &tv, float &tp) {
  const float4 a = __builtin_bit_cast(float4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0));
  const float b = __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(off + 16u), 0, 0));
  float v = P[0] * a.x;
  v = __builtin_fmaf(P[1], a.y, v);
  v = __builtin_fmaf(P[2], a.z, v);
  v = __builtin_fmaf(P[3], a.w, v);
  tv = __builtin_fmaf(P[4], b, v);
  float u = dP[1] * a.y;
  u = __builtin_fmaf(dP[2], a.z, u);
  u = __builtin_fmaf(dP[3], a.w, u);
  tp = __builtin_fmaf(dP[4], b, u);
}

// weights of the 4-point cubic on nodes -1, 0, 1, 2 and their p-derivatives
__device__ __forceinline__ void cubic_wd(float p, float *w, float *d) {
  const float pm1 = p - 1.f, pm2 = p - 2.f, pp1 = p + 1.f, s6 = 1.f / 6.f, p2 = p * p;
  w[0] = -p * pm1 * pm2 * s6;
  w[1] = pp1 * pm1 * pm2 * 0.5f;
  w[2] = -pp1 * p * pm2 * 0.5f;
  w[3] = pp1 * p * pm1 * s6;
  d[0] = -(3.f * p2 - 6.f * p + 2.f) * s6;
  d[1] = (3.f * p2 - 4.f * p - 1.f) * 0.5f;
  d[2] = -(3.f * p2 - 2.f * p - 2.f) * 0.5f;
  d[3] = (3.f * p2 - 1.f) * s6;
}

// elements q, q+4, q+8 of one canonical group into row (binary64 values)
__device__ inline int stage_q_mx(const DView &v, int gid, int nme, double x, double y, double cost, double *row,
                                 int q) {
  if (gid < 0) return ST_NO_TABLE;
  const DGroup3c g = v.grp3c[gid];
  const int nx = g.nx, ny = g.ny;
  const double dx = g.xmax / (nx - 1.0), dy = g.ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  const double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
  const int64_t icell = (int64_t)(iy - 1) * nx + (ix - 1);
  const bool three = v.meth3c[g.moff + icell] != 0;
  const double cost2 = cost * cost;
  const float sint = (float)sqrt(fmax(1.0e-5, 1.0 - cost2)), cf = (float)cost;
  float P[5], dP[5];
  P[0] = 1.f;
  P[1] = cf;
  P[2] = (float)(0.5 * (3.0 * cost2 - 1.0));
  P[3] = (float)(0.5 * cost * (5.0 * cost2 - 3.0));
  P[4] = (float)(0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2)));
  dP[0] = 0.f;
  dP[1] = 1.f;
  dP[2] = (float)(3.0 * cost);
  dP[3] = (float)((15.0 * cost2 - 3.0) / 2.0);
  dP[4] = (float)(0.5 * cost * (35.0 * cost2 - 15.0));
  const __amdgpu_buffer_rsrc_t r = slab32(v.tab3c32, v.ntab3c32);
  const unsigned PT = (unsigned)(nme * FB_TH32 * 4);            // bytes per ix step
  const unsigned RB = (unsigned)(nx * nme * FB_TH32 * 4);       // bytes per iy step
  const unsigned base = (unsigned)(g.off32 + ((int64_t)(iy - 2) * nx + (ix - 2)) * nme * FB_TH32) * 4u;
  const int nm0 = (nme == 10) ? 6 : (nme == 3 ? 2 : 1);
  float wx[4], wy[4], dwx[4], dwy[4];
  cubic_wd((float)px, wx, dwx);
  cubic_wd((float)py, wy, dwy);
  float gxs = 0.f, gys = 0.f;
  if (three) {
    // the 3-point rule's slopes come from ME 1 of theta cell 1 (binary64 table)
    const double *tab = v.tab3c + g.off;
#define T3X(i, j) tab[((int64_t)((j) - 1) * nx + ((i) - 1)) * nme]
    const double gx = (T3X(ix + 1, iy) - T3X(ix, iy)) / dx;
    const double gy = (T3X(ix, iy + 1) - T3X(ix, iy)) / dy;
#undef T3X
    const float ps = (((P[0] + P[1]) + P[2]) + P[3]) + P[4];
    gxs = (float)gx * ps;
    gys = (float)gy * ps;
  }
  const float idx = (float)(1.0 / dx), idy = (float)(1.0 / dy);
  for (int m = q; m < nme; m += 4) {
    const unsigned o = base + (unsigned)(m * FB_TH32 * 4);
    float fv, fp, fx, fy;
    if (three) {
      const unsigned c = o + RB + PT;   // (ix, iy)
      float t0, p0, t1, p1, t2, p2;
      fold2(r, c, P, dP, t0, p0);
      fold2(r, c + PT, P, dP, t1, p1);
      fold2(r, c + RB, P, dP, t2, p2);
      const float w0 = (1.f - (float)px) - (float)py;
      fv = __builtin_fmaf((float)py, t2, __builtin_fmaf((float)px, t1, w0 * t0));
      fp = __builtin_fmaf((float)py, p2, __builtin_fmaf((float)px, p1, w0 * p0));
      fx = gxs;
      fy = gys;
    } else {
      fv = 0.f;
      fp = 0.f;
      fx = 0.f;
      fy = 0.f;
#pragma unroll
      for (int l = 0; l < 4; l++) {
        float gv = 0.f, gp = 0.f, gd = 0.f;
#pragma unroll
        for (int k = 0; k < 4; k++) {
          float tv, tp;
          fold2(r, o + (unsigned)l * RB + (unsigned)k * PT, P, dP, tv, tp);
          gv = __builtin_fmaf(wx[k], tv, gv);
          gp = __builtin_fmaf(wx[k], tp, gp);
          gd = __builtin_fmaf(dwx[k], tv, gd);
        }
        fv = __builtin_fmaf(wy[l], gv, fv);
        fp = __builtin_fmaf(wy[l], gp, fp);
        fx = __builtin_fmaf(wy[l], gd, fx);
        fy = __builtin_fmaf(dwy[l], gv, fy);
      }
      fx = fx * idx;
      fy = fy * idy;
    }
    if (m >= nm0) {
      fp = fp * sint - cf * fv / sint;
      fv = fv * sint;
      fx = fx * sint;
      fy = fy * sint;
    }
    row[m] = fv;
    row[nme + m] = fp;
    row[2 * nme + m] = fx;
    row[3 * nme + m] = fy;
  }
  return 0;
}

// the scalars of one Cartesian component q (0..2) of a triplet's geometry
struct DGeoQ {
  double rh, am, sg, bm;              // rhat, amt, sighat, bmt components
  double r1q, r2q, raq, e3q;          // r1, r2, ratm components and eps(q+1, 3)
  double rA1, rA2;                    // (rna - r1)/d13, (rna - r2)/d23 components (charged terms)
};

__device__ __forceinline__ double pick3(double a, double b, double c, int q) { return q == 0 ? a : (q == 1 ? b : c); }

__device__ __forceinline__ void geo_q(const DGeo3s &G, int q, DGeoQ &Q) {
  Q.rh = pick3(G.rhat[0], G.rhat[1], G.rhat[2], q);
  Q.am = pick3(G.amt[0], G.amt[1], G.amt[2], q);
  Q.sg = pick3(G.sighat[0], G.sighat[1], G.sighat[2], q);
  Q.bm = pick3(G.bmt[0], G.bmt[1], G.bmt[2], q);
  Q.r1q = pick3(G.r1[0], G.r1[1], G.r1[2], q);
  Q.r2q = pick3(G.r2[0], G.r2[1], G.r2[2], q);
  Q.raq = pick3(G.rna[0], G.rna[1], G.rna[2], q);
  Q.e3q = pick3(G.eps[6], G.eps[7], G.eps[8], q);
  const double a1x = G.d13 > 1.0e-5 ? (G.rna[0] - G.r1[0]) / G.d13 : 0.0;
  const double a1y = G.d13 > 1.0e-5 ? (G.rna[1] - G.r1[1]) / G.d13 : 0.0;
  const double a1z = G.d13 > 1.0e-5 ? (G.rna[2] - G.r1[2]) / G.d13 : 0.0;
  const double a2x = G.d23 > 1.0e-5 ? (G.rna[0] - G.r2[0]) / G.d23 : 0.0;
  const double a2y = G.d23 > 1.0e-5 ? (G.rna[1] - G.r2[1]) / G.d23 : 0.0;
  const double a2z = G.d23 > 1.0e-5 ? (G.rna[2] - G.r2[2]) / G.d23 : 0.0;
  Q.rA1 = pick3(a1x, a1y, a1z, q);
  Q.rA2 = pick3(a2x, a2y, a2z, q);
}

// Depsilon_3c slices of component ix = q + 1 from component scalars
__device__ __forceinline__ void d3_slice_q(const DGeo3s &G, const D3Scal &S, const DGeoQ &Q, int ix, double *dA,
                                           double *dB) {
#define E3(i, j) G.eps[3 * ((j) - 1) + (i) - 1]
#define SA(imu, k) dA[((imu) - 1) + 3 * ((k) - 1)]
#define SB(imu, k) dB[((imu) - 1) + 3 * ((k) - 1)]
#pragma unroll
  for (int i = 0; i < 9; i++) { dA[i] = 0.0; dB[i] = 0.0; }
  if (S.zero) return;
  const double cm = S.crossmag;
#pragma unroll
  for (int imu = 1; imu <= 3; imu++) {
    double sum = d_levi(ix, imu, 1) * (G.r2[0] - G.r1[0]) + d_levi(ix, imu, 2) * (G.r2[1] - G.r1[1]) +
                 d_levi(ix, imu, 3) * (G.r2[2] - G.r1[2]);
    SA(imu, 2) = (1.0 / cm) * (sum - (E3(imu, 2) / cm) *
                                         ((S.r1mag2 + S.r2mag2 - 2.0 * S.r1dotr2) * Q.raq +
                                          (S.r1dotr2 - S.r1mag2 - S.r2dotratm + S.r1dotratm) * Q.r2q +
                                          (S.r1dotr2 - S.r2mag2 + S.r2dotratm - S.r1dotratm) * Q.r1q));
  }
  SA(1, 1) = E3(3, 3) * SA(2, 2) - E3(2, 3) * SA(3, 2);
  SA(2, 1) = E3(1, 3) * SA(3, 2) - E3(3, 3) * SA(1, 2);
  SA(3, 1) = E3(2, 3) * SA(1, 2) - E3(1, 3) * SA(2, 2);
#pragma unroll
  for (int imu = 1; imu <= 3; imu++) {
    SB(imu, 3) = (1.0 / G.z) * (E3(imu, 3) * Q.e3q - d_delk(imu, ix));
    double sum = d_levi(ix, imu, 1) * (G.rna[0] - G.r2[0]) + d_levi(ix, imu, 2) * (G.rna[1] - G.r2[1]) +
                 d_levi(ix, imu, 3) * (G.rna[2] - G.r2[2]);
    SB(imu, 2) = (1.0 / cm) * (sum - (E3(imu, 2) / cm) *
                                         ((S.r2dotratm - S.r1dotratm + S.r1dotr2 - S.r2mag2) * Q.raq +
                                          (S.r2dotratm + S.r1dotratm - S.r1dotr2 - S.ratmmag2) * Q.r2q +
                                          (S.ratmmag2 - 2.0 * S.r2dotratm + S.r2mag2) * Q.r1q));
  }
  SB(1, 1) = E3(3, 3) * SB(2, 2) - E3(3, 2) * SB(2, 3) - E3(2, 3) * SB(3, 2) + E3(2, 2) * SB(3, 3);
  SB(2, 1) = E3(1, 3) * SB(3, 2) - E3(1, 2) * SB(3, 3) - E3(3, 3) * SB(1, 2) + E3(3, 2) * SB(1, 3);
  SB(3, 1) = E3(2, 3) * SB(1, 2) - E3(2, 2) * SB(1, 3) - E3(1, 3) * SB(2, 2) + E3(1, 2) * SB(2, 3);
#undef E3
#undef SA
#undef SB
}

// component q of the neutral (charged = false) or one charged shell's forces of
// a triplet, continuing (a, b, c): f3naa, f3nab, f3nac sums
template <int CLS, bool CHG>
__device__ __forceinline__ void force3c_q(const DView &v, const DForce &Fo, const DGeo3s &G, const D3Scal &S,
                                          const DGeoQ &Q, int q, int64_t bo, const double *row, double dQ,
                                          double stinky, double dstA, double dstB, double dstC, double &a, double &b,
                                          double &c) {
  constexpr int NME = nme3c_of(CLS);
  constexpr int NMU = (CLS & 2) ? 4 : 1, NNU = (CLS & 1) ? 4 : 1;
  double pm[9], sA[9], sB[9], dpA[9], dpB[9], F[NME], Ma[NME], Mb[NME];
  d_pmat<double>(G.eps, pm);
  d3_slice_q(G, S, Q, q + 1, sA, sB);
  dpmat_slice(sA, dpA);
  dpmat_slice(sB, dpB);
#pragma unroll
  for (int m = 0; m < NME; m++) {
    F[m] = row[m];
    const double dp = row[NME + m];
    Ma[m] = Q.rh * row[2 * NME + m] + Q.am * dp;
    Mb[m] = -Q.sg * row[3 * NME + m] + Q.bm * dp - Ma[m] / 2.0;
  }
  const double *sb = v.mat[FB_MAT_S] + bo, *db = v.mat[FB_MAT_DIP] + bo;
  const double *sD = Fo.dmat[0] + 3 * bo, *dD = Fo.dmat[2] + 3 * bo;
  const double d13sq = G.d13 * G.d13, d23sq = G.d23 * G.d23;
#pragma unroll
  for (int col = 0; col < NNU; col++)
#pragma unroll
    for (int row_ = 0; row_ < NMU; row_++) {
      const int i = row_ + NMU * col;
      const double xa = drot3c_el<CLS>(row_, col, pm, dpA, F, Ma);
      const double xb = drot3c_el<CLS>(row_, col, pm, dpB, F, Mb);
      const double xc = -xa - xb;
      if (!CHG) {
        const double r = Fo.rho[bo + i];
        a = a - r * xa * FB_P_EQ2;
        b = b - r * xb * FB_P_EQ2;
        c = c - r * xc * FB_P_EQ2;
      } else {
        const double bx = rot3c_el<CLS>(row_, col, pm, F);
        const double sterm = sb[i] / 2.0, dterm = db[i] / G.z;
        const double emnpl = (sterm - dterm) / G.d13 + (sterm + dterm) / G.d23;
        const double spterm = sD[q + 3 * i] / 2.0;
        double dpterm = dD[q + 3 * i] / G.z;
        dpterm = dpterm + dterm * Q.sg / G.z;
        const double dA = -((sterm - dterm) * Q.rA1 / d13sq) - (sterm + dterm) * Q.rA2 / d23sq;
        const double dB = (sterm - dterm) * Q.rA1 / d13sq + (spterm - dpterm) / G.d13 + (spterm + dpterm) / G.d23;
        const double dC = -dA - dB;
        const double w = FB_P_EQ2 * dQ * Fo.rho[bo + i];
        a = a - w * (stinky * xa + dstA * bx - dstA * emnpl + (1.0 - stinky) * dA);
        b = b - w * (stinky * xb + dstB * bx - dstB * emnpl + (1.0 - stinky) * dB);
        c = c - w * (stinky * xc + dstC * bx - dstC * emnpl + (1.0 - stinky) * dC);
      }
    }
}

template <bool CHG>
__device__ __forceinline__ void force3c_dispatch(int cls, const DView &v, const DForce &Fo, const DGeo3s &G,
                                                 const D3Scal &S, const DGeoQ &Q, int q, int64_t bo, const double *row,
                                                 double dQ, double st, double dA, double dB, double dC, double &a,
                                                 double &b, double &c) {
  switch (cls) {
    case 3: force3c_q<3, CHG>(v, Fo, G, S, Q, q, bo, row, dQ, st, dA, dB, dC, a, b, c); break;
    case 2: force3c_q<2, CHG>(v, Fo, G, S, Q, q, bo, row, dQ, st, dA, dB, dC, a, b, c); break;
    case 1: force3c_q<1, CHG>(v, Fo, G, S, Q, q, bo, row, dQ, st, dA, dB, dC, a, b, c); break;
    default: force3c_q<0, CHG>(v, Fo, G, S, Q, q, bo, row, dQ, st, dA, dB, dC, a, b, c); break;
  }
}

// The per-triplet forces go to three arrays in the order in which they are
// summed per atom: f3a (on ialpha) in list order, f3b (on iatom) in target-pair
// order, f3c (on jatom) in jatom order; component q of each.
template <bool MX>
__global__ __launch_bounds__(T3Q) void k_d3c_sp(DView v, DForce Fo) {
  __shared__ double sI[T3Q / 4][SIROW];
  const int slot = threadIdx.x >> 2, q = threadIdx.x & 3;
  const int64_t t = xcd_tile(blockIdx.x, gridDim.x) * (T3Q / 4) + slot;
  if (t >= v.ntrip) return;
  const int4 tr = Fo.trip[t];
  const int in1 = v.imass[tr.y], in2 = v.imass[tr.z], indna = v.imass[tr.x];
  const int cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
  const int nme = cls == 3 ? 10 : (cls == 0 ? 1 : 3);
  double x, z, cost;
  xzc3(v, tr, false, x, z, cost);
  double *row = sI[slot];
  const int gid0 = group3c_of(v, in1, in2, indna, 0);
  int status = MX ? stage_q_mx(v, gid0, nme, x, z, cost, row, q) : stage_q(v, gid0, nme, x, z, cost, row, q);
  // the four lanes of a triplet sit in one wavefront: a wavefront barrier
  // orders phase A's LDS writes before phase B's reads
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  status |= __shfl_xor(status, 1, 4) | __shfl_xor(status, 2, 4);
  if (q < 3) {
    double a = 0.0, b = 0.0, c = 0.0;
    if (!status) {
      DGeo3s G;
      d_dgeo3s(v, tr, false, G);
      D3Scal S;
      d3_scal(G.r1, G.r2, G.r21, G.rna, G.rnabc, S);
      DGeoQ Q;
      geo_q(G, q, Q);
      force3c_dispatch<false>(cls, v, Fo, G, S, Q, q, v.boff[Fo.tpair[t]], row, 0.0, 0.0, 0.0, 0.0, 0.0, a, b, c);
    }
    Fo.f3a[3 * t + q] = a;
    Fo.f3b[3 * (int64_t)Fo.inv_perm[t] + q] = b;
    Fo.f3c[3 * (int64_t)Fo.inv_perm_j[t] + q] = c;
  }
  if (status && q == 0) atomicOr(v.status, status);
}

// charged shells (M_Dassemble_3c_DOGS.f90:460-776): continues each triplet's
// neutral sums with every isorp whose dQ is not zero
__global__ __launch_bounds__(T3Q) void k_d3c_sp_q(DView v, DForce Fo) {
  __shared__ double sI[T3Q / 4][SIROW];
  const int slot = threadIdx.x >> 2, q = threadIdx.x & 3;
  const int64_t t = xcd_tile(blockIdx.x, gridDim.x) * (T3Q / 4) + slot;
  if (t >= v.ntrip) return;
  const int4 tr = Fo.trip[t];
  const int indna = v.imass[tr.x];
  bool charged = false;
  for (int isorp = 1; isorp <= v.sp.nssh[indna - 1]; isorp++) charged |= (v.dQ[tr.x * FB_MAXSH + isorp - 1] != 0.0);
  if (!charged) return;
  const int in1 = v.imass[tr.y], in2 = v.imass[tr.z];
  const int cls = (v.sp.norb[in1 - 1] == 4 ? 2 : 0) + (v.sp.norb[in2 - 1] == 4 ? 1 : 0);
  const int nme = cls == 3 ? 10 : (cls == 0 ? 1 : 3);
  const int64_t bo = v.boff[Fo.tpair[t]];
  DGeo3s G;
  d_dgeo3s(v, tr, true, G);
  D3Scal S;
  d3_scal(G.r1, G.r2, G.r21, G.rna, G.rnabc, S);
  DGeoQ Q;
  geo_q(G, q < 3 ? q : 0, Q);
  const double rc3 = v.sp.rcmin[indna - 1];
  const double rend1 = v.sp.rcmin[in1 - 1] + rc3, rend2 = v.sp.rcmin[in2 - 1] + rc3;
  const double st13 = d_smoother(G.d13, rend1, 0.8), dst13 = d_Dsmoother(G.d13, rend1, 0.8);
  const double st23 = d_smoother(G.d23, rend2, 0.8), dst23 = d_Dsmoother(G.d23, rend2, 0.8);
  const double stinky = st13 * st23;
  const double dstB = -dst13 * st23 * Q.rA1;
  const double dstC = -st13 * dst23 * Q.rA2;
  const double dstA = -dstB - dstC;
  double *row = sI[slot];
  double a = 0.0, b = 0.0, c = 0.0;
  const int64_t ob = 3 * (int64_t)Fo.inv_perm[t] + q, oc = 3 * (int64_t)Fo.inv_perm_j[t] + q;
  if (q < 3) {
    a = Fo.f3a[3 * t + q];
    b = Fo.f3b[ob];
    c = Fo.f3c[oc];
  }
  int status = 0;
  for (int isorp = 1; isorp <= v.sp.nssh[indna - 1]; isorp++) {
    const double dQ = v.dQ[tr.x * FB_MAXSH + isorp - 1];
    if (dQ == 0.0) continue;
    int st = stage_q(v, group3c_of(v, in1, in2, indna, isorp), nme, G.x, G.z, G.cost, row, q);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    st |= __shfl_xor(st, 1, 4) | __shfl_xor(st, 2, 4);
    status |= st;
    if (!st && q < 3)
      force3c_dispatch<true>(cls, v, Fo, G, S, Q, q, bo, row, dQ, stinky, dstA, dstB, dstC, a, b, c);
    // the row is rewritten by the next shell: reads of this one complete first
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  if (q < 3) {
    Fo.f3a[3 * t + q] = a;
    Fo.f3b[ob] = b;
    Fo.f3c[oc] = c;
  }
  if (status && q == 0) atomicOr(v.status, status);
}

extern "C" int fbk_d2c_sp(const DView &v, const DForce &f, int mask, hipStream_t st) {
  if (v.npairs == 0) return 0;
  const dim3 grid((unsigned)((v.npairs + WD2S - 1) / WD2S));
  if (mask & 7) hipLaunchKernelGGL(k_d2c_sp<0>, grid, dim3(WD2S), 0, st, v, f, mask);
  if (mask & 8) {
    hipLaunchKernelGGL(k_d2c_sp<1>, grid, dim3(WD2S), 0, st, v, f, mask);
    hipLaunchKernelGGL(k_d2c_sp<2>, grid, dim3(WD2S), 0, st, v, f, mask);
    if (v.any_dq) hipLaunchKernelGGL(k_d2c_sp<3>, grid, dim3(WD2S), 0, st, v, f, mask);
  }
  return (int)hipGetLastError();
}

// charged = 0: the neutral pass only (every shell charge of the batch is zero);
// mixed = 1: the FP32 interpolation of the neutral pass
extern "C" int fbk_d3c_sp(const DView &v, const DForce &f, int charged, int mixed, hipStream_t st) {
  if (v.ntrip == 0) return 0;
  const dim3 grid((unsigned)((v.ntrip + T3Q / 4 - 1) / (T3Q / 4)));
  if (mixed) hipLaunchKernelGGL(k_d3c_sp<true>, grid, dim3(T3Q), 0, st, v, f);
  else hipLaunchKernelGGL(k_d3c_sp<false>, grid, dim3(T3Q), 0, st, v, f);
  if (charged) hipLaunchKernelGGL(k_d3c_sp_q, grid, dim3(T3Q), 0, st, v, f);
  return (int)hipGetLastError();
}
