// Device pointers of the force assemblers (inputs, outputs and the triplet
// orderings of the per-atom sums), passed by value beside DView.
#pragma once

#include "fb_internal.h"

struct DForce {
  const double *rho;        // [blen] density blocks, block layout of the H/S blocks
  double *dmat[3];          // [3*blen] Dblock(3, norb_mu, norb_nu) of S, T, dipole_z
  double *fatom, *fontop;   // [npairs][3] forces%vna_atom / vna_ontop
  double *f3a;              // [ntrip][3] per-triplet force on ialpha, list order
  double *f3b;              // [ntrip][3] per-triplet force on iatom, target-pair order
  double *f3c;              // [ntrip][3] per-triplet force on jatom, jatom order
  double *f3;               // [natoms][9] f3naa, f3nab, f3nac
  const int4 *trip;         // triplets in list order
  const int *tpair;         // target pair of each triplet, list order
  const int64_t *toff;      // [natoms+1] list-order range of each ialpha
  const int *inv_perm;      // list order -> target-pair order
  const int64_t *ioff;      // [natoms+1] range of each iatom in target-pair order
  const int *inv_perm_j;    // list order -> jatom order
  const int64_t *joff;      // [natoms+1] range of each jatom in jatom order
};

extern "C" {
int fbk_d2c(const DView &v, const DForce &f, int mask, hipStream_t st);
int fbk_d3c(const DView &v, const DForce &f, hipStream_t st);
int fbk_d2c_sp(const DView &v, const DForce &f, int mask, hipStream_t st);
int fbk_d3c_sp(const DView &v, const DForce &f, int charged, int mixed, hipStream_t st);
int fbk_f3_reduce(const DView &v, const DForce &f, hipStream_t st);
int fbk_f3_offsets(int natoms, int64_t ntrip, const int64_t *noff, const int *tpair_sorted, const int *jkey_sorted,
                   int64_t *ioff, int64_t *joff, hipStream_t st);
int fbk_trip_jkey(int64_t ntrip, const int4 *trip, int *key, hipStream_t st);
int fbk_invert_perm(int64_t n, const int *perm, int *inv, hipStream_t st);
}
