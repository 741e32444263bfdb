// Device view of the pseudopotential path: PP shells of the species, the three
// PP neighbour lists of driver_neighbors_PP (c.SYSTEM/M_neighbors_PP.f90) with
// their maps, the PP triplets, and the sVNL / vnl block slabs.  Passed by value
// beside DView.
#pragma once

#include "fb_internal.h"

enum { PPL_PP = 0, PPL_PPX = 1, PPL_PPP = 2 };   // neighbors_PP, neighbors_PPx, neighbors_PPp

struct DPP {
  // species(:)%shell_PP(:)%lssh, %cl (per projector shell), norb_PP_max, rcutoff_PP
  int nssh_pp[FB_MAXSPECIES];
  int lssh_pp[FB_MAXSPECIES][FB_MAXSH];
  int norb_pp[FB_MAXSPECIES];
  double rcpp[FB_MAXSPECIES];
  double cl[FB_MAXSPECIES][FB_MAXSH];
  double clo[FB_MAXSPECIES][4];   // cl per projector orbital (b.FUNCTIONS/cl.f90), s/p PP shells
  int canon;                      // s/p species and PP shells, P_vnl tables in make_munu order
  // lists, CSR over global atoms; neighbour atoms 1-based per structure, images 0..mbeta_max(+1)
  int *n[3];
  int64_t *off[3];        // [natoms+1]
  int *j[3], *b[3];
  int *atom[3];           // owning (global) atom of each pair
  int *self[3];           // [natoms] 1-based self entry
  int *pp_map, *px_map, *px_point;
  int *inv_pp, *inv_px;   // per PPp pair: the PP / PPx pair mapped onto it, or -1
  int64_t *cand_off;      // [natoms+1] PPp candidate ranges
  int *cand_j, *cand_b;   // PPp construction scratch
  // PP triplets: list order (common-neighbour order of each ialpha) and sorted by target
  double *ppp_disp;       // [PPp pairs][3] xl(kbeta) + r_k - r_i; NaN where kbeta lies beyond the images
  // per-atom open-addressing tables of the PPp keys (jatom * 1024 + image):
  // slots pk_off[i]..pk_off[i+1] (a power of two), key and list index per slot
  unsigned *pk_key;
  int *pk_idx;
  int64_t *pk_off;
  int *pc_n;
  int64_t *pc_off;        // [natoms+1]
  int4 *trip;             // (ialpha, iatom, jatom global; ibeta | jbeta << 16)
  int *tpair;             // target PPp pair (global)
  int4 *trip_s;
  int *tpair_s;
  int64_t *seg_off;
  int *seg_pair, *trip_seg;
  int2 *mp;               // per sorted triplet: mpairnay of (iatom, ialpha) and (jatom, ialpha), 0 when not unique
  int4 *tiles;            // whole segments packed into tiles of at most 64 triplets: first, last segment, r0, r1
  int64_t ntrip, nseg;
  int ntiles;
  // blocks
  int64_t *sv_off, *vl_off;   // [npairs+1] of PP / PPp
  double *svnl, *vnl;
  int64_t npairs[3];
};

extern "C" {
int fbk_pp_neighbors(const DView &v, const DPP &q, int list, int fill, int *status, hipStream_t st);
int fbk_ppp_candidates(const DView &v, const DPP &q, int64_t *cnt, hipStream_t st);
int fbk_ppp_build(const DView &v, const DPP &q, int fill, int *status, hipStream_t st);
int fbk_pp_maps(const DView &v, const DPP &q, int *status, hipStream_t st);
int fbk_pp_common(const DView &v, const DPP &q, int fill, hipStream_t st);
int fbk_ppp_disp(const DView &v, const DPP &q, hipStream_t st);
int fbk_ppk_size(const DView &v, const DPP &q, int *sz, hipStream_t st);
int fbk_ppk_insert(const DView &v, const DPP &q, hipStream_t st);
int fbk_pp_block_sizes(const DView &v, const DPP &q, int which, int64_t *sz, hipStream_t st);
int fbk_svnl(const DView &v, const DPP &q, hipStream_t st);
int fbk_vnl_2c(const DView &v, const DPP &q, hipStream_t st);
int fbk_vnl_3c(const DView &v, const DPP &q, hipStream_t st);
int fbk_pp_mpair(const DView &v, const DPP &q, hipStream_t st);
}
