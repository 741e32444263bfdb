// Host side of the C ABI declared in include/fireball_hip.h.
//
// Owns the device copies of species, tables, structures, lists and blocks,
// orders every launch on one HIP stream and turns kernel status bits into the
// FB_E_* codes that replace the reference's `stop` statements.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "fb_internal.h"
#include "fb_forces.h"
#include "fb_pp.h"
#include "fb_mlp.h"
#include "fb_fdata.h"
#include <array>
#include <random>

// kernel launchers (fb_lists.hip, fb_2c.hip, fb_3c.hip, fb_3c_mixed.hip)
extern "C" {
int fbk_neighbors_count(const DView &v, int *neighn, int *status, hipStream_t st);
int fbk_pair_disp(const DView &v, double *disp, hipStream_t st);
int fbk_tiles_atom(int natoms, const int64_t *pair_off, const int64_t *seg_off, const int *seg_pair, int64_t nseg,
                   int ch, int fill, int *cnt, const int64_t *tile_off, int4 *tiles, hipStream_t st);
int fbk_gather_i64(int n, const int64_t *src, const int64_t *idx, int64_t *dst, hipStream_t st);
int fbk_cell_count(const DView &v, int *cnt, int *atom_cell, hipStream_t st);
int fbk_cell_fill(const DView &v, const int *atom_cell, int *cursor, int *cell_atoms, hipStream_t st);
int fbk_neighbors_fill(const DView &v, const int64_t *noff, int *neigh_j, int *neigh_b, int *neigh_self,
                       int *pair_atom, int *status, hipStream_t st);
int fbk_neigh_back(const DView &v, int *neigh_back, int *status, hipStream_t st);
int fbk_common_count(const DView &v, int *ncommon, int *status, hipStream_t st);
int fbk_common_fill(const DView &v, const int64_t *toff, int4 *trip, int *trip_pair, int *status, hipStream_t st);
int fbk_block_sizes(const DView &v, int64_t *sz, int4 *rec, hipStream_t st);
int fbk_pair_perm(const DView &v, unsigned char *perm, hipStream_t st);
int fbk_scan_i32(const int *in, int64_t *out, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t st);
int fbk_scan_i64(const int64_t *in, int64_t *out, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t st);
int fbk_sort_pairs(const int *keys_in, int *keys_out, const int *vals_in, int *vals_out, int64_t n, int end_bit,
                   void *tmp, size_t *tmp_bytes, hipStream_t st);
int fbk_scan_incl_i32(const int *in, int *out, int64_t n, void *tmp, size_t *tmp_bytes, hipStream_t st);
int fbk_iota(int *a, int64_t n, hipStream_t st);
int fbk_seg_mark(int64_t ntrip, const int *sorted_pair, int *flag, hipStream_t st);
int fbk_seg_fill(int64_t ntrip, const int *sorted_pair, const int *segid_incl, int64_t *seg_off, int *seg_pair,
                 int *trip_seg, int *pair_tgt, hipStream_t st);
int fbk_gather_trip(int64_t ntrip, const int *perm, const int4 *in, int4 *out, hipStream_t st);
int fbk_trip_aux(const DView &v, int *aux, hipStream_t st);
int fbk_fill_i32(int *a, int64_t n, int val, hipStream_t st);
int fbk_2c(const DView &v, int mask, hipStream_t st);
int fbk_3c_fp64(const DView &v, hipStream_t st);
int fbk_3c_mixed(const DView &v, int chunk, int charged, hipStream_t st);
int fbk_2c_sp(const DView &v, int mask, hipStream_t st);
int fbk_2c_lds(const DView &v, int mask, hipStream_t st);
int fbk_3c_fp64_ph(const DView &v, int charged, hipStream_t st);
int fbk_gamma_fold(const DView &v, int a0, int natoms, int which, const int *orb_off, int n, double *A,
                   hipStream_t st);
int fbk_3c_mlp(const DView &v, int *queue, int nwg, hipStream_t st);
int fbk_hs_fused(const DView &v, int waves, int *queue, int ncu, hipStream_t st);
int fbk_mlp_samples(const DView &v, int gid, int64_t n, unsigned long long seed, float *X, float *T, hipStream_t st);
int fbk_mlp_colmaxabs(int64_t n, int ncol, const float *A, float *out, hipStream_t st);
int fbk_mlp_rowmax(int64_t n, const float *T, float *M, hipStream_t st);
int fbk_mlp_unenvelope(int64_t n, const float *X, float *T, float c, float kx, float ky, float xh, float yh,
                       hipStream_t st);
int fbk_mlp_scale_cols(int64_t n, int ncol, float *A, const float *inv, hipStream_t st);
int fbk_mlp_gemm(int M, int N, int K, const float *A, int64_t sai, int64_t sak, const float *B, int64_t sbk,
                 int64_t sbj, float *C, int64_t sci, int64_t scj, int nz, float *partial, hipStream_t st);
int fbk_mlp_bias_act(int64_t n, float *Z, float *A, const float *b, hipStream_t st);
int fbk_mlp_out(int64_t n, int nme, float *Y, const float *b3, const float *T, float *dY, float *rowloss,
                float inv_count, hipStream_t st);
int fbk_mlp_act_back(int64_t n, const float *dA, const float *Z, float *dZ, hipStream_t st);
int fbk_mlp_adam(int np, float *p, const float *g, float *m, float *v, float lr, float b1, float b2, float c1,
                 float c2, float eps, hipStream_t st);
int fbk_mlp_infer(const DMlp *w, int64_t n, const float *X, float *Y, hipStream_t st);
int fbk_mlp_loss(int64_t n, const float *rowloss, float scale, float *loss, hipStream_t st);
int fbk_mlp_calib(const DView &v, int gid, const DMlp *w, int64_t n, unsigned long long seed, unsigned *err,
                  hipStream_t st);
}

namespace {

struct DevBuf {
  void *p = nullptr;
  size_t cap = 0;
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    size_t b = bytes ? bytes : 16;
    hipError_t e = hipMalloc(&p, b);
    if (e == hipSuccess) cap = b;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
  template <typename T> T *as() const { return (T *)p; }
};

struct HCell2c {
  int is, js, itype, isub, nME, nx;
  double xmax;
  std::vector<int> mu, nu, mv;
  std::vector<double> data;  // [nx][nME]
};

struct HCell3c {
  int is, js, ks, itype, isub, itheta, nME, nx, ny;
  double xmax, ymax;
  std::vector<int> mu, nu, mv;
  std::vector<double> data;  // [ny][nx][nME]
};

struct HStruct {
  int natoms, icluster;
  std::vector<double> ratom;
  std::vector<int> imass;
  double lattice[9];
  std::vector<double> dQ;  // [natoms][FB_MAXSH]
  std::vector<double> qin; // [natoms][FB_MAXSH] input shell charges Qin (McWEDA densities)
  bool has_qin = false;
};

int slot_of(int itype) {
  switch (itype) {
    case 1: return SLOT_S;
    case 2: return SLOT_T;
    case 3: return SLOT_ONL;
    case 4: return SLOT_ONR;
    case 5: return SLOT_ATOM;
    case 10: return SLOT_DIP;
    case 15: return SLOT_RHO_ONL;
    case 16: return SLOT_RHO_ONR;
    case 17: return SLOT_RHO_ATOM;
    case 19: return SLOT_RHOS_ONL;
    case 20: return SLOT_RHOS_ONR;
    case 21: return SLOT_RHOS_ATOM;
    case 6: return SLOT_VNL;
  }
  return -1;
}

// canonical (mu, nu, mvalue) lists of make_munu (c.SYSTEM/M_integrals_2c.f90:297-335)
// and make_munu_3c (c.SYSTEM/M_integrals_3c.f90:299-477), s and p shells
struct MeList {
  std::vector<int> mu, nu, mv;
};

MeList munu2c(const std::vector<int> &li, const std::vector<int> &lj) {
  MeList r;
  int n1 = 0;
  for (int l1 : li) {
    n1 += l1 + 1;
    int n2 = 0;
    for (int l2 : lj) {
      n2 += l2 + 1;
      for (int m = -std::min(l1, l2); m <= std::min(l1, l2); m++) {
        r.mu.push_back(n1 + m);
        r.nu.push_back(n2 + m);
        r.mv.push_back(0);
      }
      n2 += l2;
    }
    n1 += l1;
  }
  return r;
}

MeList munu3c_sp(const std::vector<int> &li, const std::vector<int> &lj) {
  MeList r = munu2c(li, lj);
  int n1 = 0;
  for (int l1 : li) {
    n1 += l1 + 1;
    int n2 = 0;
    for (int l2 : lj) {
      n2 += l2 + 1;
      if (l1 == 0 && l2 != 0) { r.mu.push_back(n1); r.nu.push_back(n2 + 1); r.mv.push_back(1); }
      if (l1 == 1) {
        r.mu.push_back(n1 + 1); r.nu.push_back(n2); r.mv.push_back(1);
        if (l2 != 0) { r.mu.push_back(n1); r.nu.push_back(n2 + 1); r.mv.push_back(1); }
      }
      n2 += l2;
    }
    n1 += l1;
  }
  return r;
}

bool same_list(const MeList &a, const std::vector<int> &mu, const std::vector<int> &nu, const std::vector<int> &mv) {
  return a.mu == mu && a.nu == nu && a.mv == mv;
}

}  // namespace

struct fb_ctx {
  int device = 0;
  float *scratch = nullptr;   // this device's scratch array (fb_mlp_gradients_noalloc)
  int ncu = 256;              // compute units of the device
  int fused = 0;              // FB_TUNE_FUSED: 0 separate 2c/3c launches, 3 or 4 one launch (waves per SIMD)
  int persist = 0;            // FB_TUNE_PERSIST: the one launch's 3c waves take tiles from a queue
  hipStream_t own_stream = nullptr, stream = nullptr;
  hipStream_t side = nullptr;          // second stream: 2c beside 3c in a build
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  std::string err;
  int precision = FB_PREC_FP64;
  int method3c = FB_3C_TABLE;
  int chunk3c = 128;      // triplets per workgroup chunk of the mixed 3c kernel
  int k2c = 1;            // 2c kernel for s/p species: 0 LDS-staged cells, 1 register-resident

  DSpecies sp{};
  int nZ[FB_MAXSPECIES] = {0};
  bool species_set = false;
  std::vector<HCell2c> c2;
  std::vector<HCell3c> c3;
  bool tables_ok = false;
  DevBuf d_idx3r;   // McWEDA density groups (rho_3c, rhoS_3c)
  int maxsh = 1;    // largest shell count of any species
  DevBuf d_cell2c, d_idx2c, d_grp3c, d_idx3c, d_tab2c, d_tab2c32, d_tab3c, d_tab3c32, d_tab3c32q, d_mu, d_nu, d_mv,
      d_meth3c;
  DevBuf d_grid2c, d_pden, d_wg2c;
  DevBuf d_cell_cnt, d_cell_start, d_cell_atoms, d_atom_cell;   // cell lists of the neighbour search
  DevBuf d_pair_disp;   // [npairs][3] neighbour displacements of the common-neighbour search
  DevBuf d_tile_cnt, d_tile_off, d_gidx, d_gout;   // tile packing per atom; gathers of per-structure offsets
  int64_t ncells = 0;
  int cell_bits = 0;
  int canon_sp = 0;
  int canon_rho = 0;      // s/p species with density tables in make_munu / make_munuS order
  int ncell2c = 0, ngrid2c = 0;
  int64_t ntab3c32 = 0, ntab3c32q = 0;
  int nwg2c = 0;

  std::vector<HStruct> structs;
  bool structs_dirty = true;
  std::vector<DStruct> h_st;
  std::vector<int> h_imass;
  int natoms = 0;
  DevBuf d_st, d_xl, d_ratom, d_imass, d_atom_struct, d_dQ, d_qin;

  DevBuf d_neighn, d_noff, d_neigh_j, d_neigh_b, d_neigh_back, d_neigh_self, d_pair_atom, d_pair_rec, d_pair_perm, d_bsz, d_boff;
  DevBuf d_ncommon, d_toff, d_trip_orig, d_tpair_orig, d_trip, d_tpair, d_perm_in, d_perm, d_flag, d_segincl,
      d_seg_off, d_seg_pair, d_trip_seg, d_pair_tgt, d_trip_aux, d_tiles[3];
  int ntiles[3] = {0, 0, 0};
  DevBuf d_tmp, d_status;
  DevBuf d_mat[FB_NMAT];
  int64_t npairs = 0, ntrip = 0, nseg = 0, blen = 0;
  std::vector<int64_t> h_noff, h_toff, h_boff_struct;
  bool lists_ok = false;
  int done[5] = {0, 0, 0, 0, 0};

  // forces (k.DASSEMBLERS): density matrix in, Dblocks and force pieces out
  DevBuf d_rho, d_dmat[3], d_fatom, d_fontop, d_f3a, d_f3b, d_f3c, d_f3, d_jkey, d_jkey_s, d_iota_t, d_perm_j,
      d_inv_perm, d_inv_perm_j, d_ioff, d_joff;
  long forces_gen = -1;          // lists_gen the force buffers and orderings belong to
  // McWEDA densities: shell-block offsets, per-structure ranges, progress
  DevBuf d_wbsz, d_wboff;
  std::vector<int64_t> h_wboff_struct;
  int64_t wlen = 0;
  long rho_gen = -1;             // lists_gen the density buffers belong to
  int rdone[4] = {0, 0, 0, 0};   // rho_2c, rho_3c, rho_weighted_2c, rho_weighted_3c

  // pseudopotential: PP shells of the species, driver_neighbors_PP lists, sVNL / vnl blocks
  bool pp_species = false;
  DPP hpp{};                     // species part; pview() adds the device pointers
  DevBuf d_ppn[3], d_ppoff[3], d_ppj[3], d_ppb[3], d_ppatom[3], d_ppself[3];
  DevBuf d_pp_map, d_px_map, d_px_point, d_inv_pp, d_inv_px, d_cand_cnt, d_cand_off, d_cand_j, d_cand_b;
  DevBuf d_pc_n, d_pc_off, d_pc_trip, d_pc_tpair, d_pc_trip_s, d_pc_tpair_s, d_pc_perm_in, d_pc_perm, d_pc_flag,
      d_pc_segincl, d_pc_seg_off, d_pc_seg_pair, d_pc_trip_seg, d_pc_tgt, d_pc_mp, d_pc_tiles, d_ppp_disp, d_pk_size,
      d_pk_off, d_pk_key, d_pk_idx;
  int pp_ntiles = 0;
  DevBuf d_sv_bsz, d_sv_off, d_vl_bsz, d_vl_off, d_svnl, d_vnl;
  int64_t pp_npairs[3] = {0, 0, 0}, pp_ntrip = 0, pp_nseg = 0, sv_len = 0, vl_len = 0;
  std::vector<int64_t> h_ppoff[3], h_pcoff, h_sv_struct, h_vl_struct;
  bool pp_lists_ok = false;
  int ppdone[3] = {0, 0, 0};     // svnl, vnl_2c, vnl_3c

  DPP pview() const {
    DPP q = hpp;
    q.canon = hpp.canon && force_kernel;   // FB_TUNE_FORCE_KERNEL 0 selects the generic kernels
    for (int l = 0; l < 3; l++) {
      q.n[l] = d_ppn[l].as<int>();
      q.off[l] = d_ppoff[l].as<int64_t>();
      q.j[l] = d_ppj[l].as<int>();
      q.b[l] = d_ppb[l].as<int>();
      q.atom[l] = d_ppatom[l].as<int>();
      q.self[l] = d_ppself[l].as<int>();
      q.npairs[l] = pp_npairs[l];
    }
    q.pp_map = d_pp_map.as<int>();
    q.px_map = d_px_map.as<int>();
    q.px_point = d_px_point.as<int>();
    q.inv_pp = d_inv_pp.as<int>();
    q.inv_px = d_inv_px.as<int>();
    q.cand_off = d_cand_off.as<int64_t>();
    q.cand_j = d_cand_j.as<int>();
    q.cand_b = d_cand_b.as<int>();
    q.pc_n = d_pc_n.as<int>();
    q.pc_off = d_pc_off.as<int64_t>();
    q.trip = d_pc_trip.as<int4>();
    q.tpair = d_pc_tpair.as<int>();
    q.trip_s = d_pc_trip_s.as<int4>();
    q.tpair_s = d_pc_tpair_s.as<int>();
    q.seg_off = d_pc_seg_off.as<int64_t>();
    q.seg_pair = d_pc_seg_pair.as<int>();
    q.trip_seg = d_pc_trip_seg.as<int>();
    q.mp = d_pc_mp.as<int2>();
    q.tiles = d_pc_tiles.as<int4>();
    q.ppp_disp = d_ppp_disp.as<double>();
    q.pk_key = d_pk_key.as<unsigned>();
    q.pk_idx = d_pk_idx.as<int>();
    q.pk_off = d_pk_off.as<int64_t>();
    q.ntiles = pp_ntiles;
    q.ntrip = pp_ntrip;
    q.nseg = pp_nseg;
    q.sv_off = d_sv_off.as<int64_t>();
    q.vl_off = d_vl_off.as<int64_t>();
    q.svnl = d_svnl.as<double>();
    q.vnl = d_vnl.as<double>();
    return q;
  }
  std::vector<char> rho_ok;      // per structure: density blocks uploaded
  int ddone[5] = {0, 0, 0, 0, 0};  // Dblock S, T, dipole_z; vna_2c forces; vna_3c forces
  int force_kernel = 1;            // s/p species: 1 register-resident kernels, 0 the generic ones

  // deferred status of asynchronous H+S builds
  int *h_status = nullptr;      // pinned
  bool pending = false;
  // captured H+S build
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  long graph_key = -1;
  long lists_gen = 0;
  bool use_graph = true;

  // learned 3c surrogate
  std::vector<DGroup3c> h_grp3c;
  std::vector<std::array<int, 4>> h_grpkey;   // (is, js, ks, isub) of each group
  DevBuf d_mlp, d_mlp_err, d_mlp_count, d_queue;
  std::vector<DMlp> h_mlp;
  std::vector<std::vector<float>> h_mlp_params, h_mlp_err;
  std::vector<std::array<double, 8>> mlp_stat;
  bool mlp_ready = false;
  double mlp_tau = 2.5e-5;
  long mlp_gen = 0;

  bool timing = false;
  struct Ev {
    std::string name;
    hipEvent_t a, b;
  };
  std::vector<Ev> events;
  std::vector<hipEvent_t> pool;

  int fail(int code, const char *what, hipError_t e = hipSuccess) {
    err = what;
    if (e != hipSuccess) {
      err += ": ";
      err += hipGetErrorString(e);
    }
    return code;
  }

  DForce fview() const {
    DForce f{};
    f.rho = d_rho.as<double>();
    for (int i = 0; i < 3; i++) f.dmat[i] = d_dmat[i].as<double>();
    f.fatom = d_fatom.as<double>();
    f.fontop = d_fontop.as<double>();
    f.f3a = d_f3a.as<double>();
    f.f3b = d_f3b.as<double>();
    f.f3c = d_f3c.as<double>();
    f.f3 = d_f3.as<double>();
    f.trip = d_trip_orig.as<int4>();
    f.tpair = d_tpair_orig.as<int>();
    f.toff = d_toff.as<int64_t>();
    f.inv_perm = d_inv_perm.as<int>();
    f.ioff = d_ioff.as<int64_t>();
    f.inv_perm_j = d_inv_perm_j.as<int>();
    f.joff = d_joff.as<int64_t>();
    return f;
  }

  DView view() const {
    DView v{};
    v.sp = sp;
    for (auto &st : structs)
      for (double q : st.dQ)
        if (q != 0.0) v.any_dq = 1;
    v.cell2c = d_cell2c.as<DCell2c>();
    v.ncell2c = ncell2c;
    v.ngrid2c = ngrid2c;
    v.idx2c = d_idx2c.as<int>();
    v.grp3c = d_grp3c.as<DGroup3c>();
    v.idx3c = d_idx3c.as<int>();
    v.idx3r = d_idx3r.as<int>();
    v.tab2c = d_tab2c.as<double>();
    v.tab2c32 = d_tab2c32.as<float>();
    v.tab3c = d_tab3c.as<double>();
    v.tab3c32 = d_tab3c32.as<float>();
    v.meth3c = d_meth3c.as<unsigned char>();
    v.ntab3c32 = ntab3c32;
    v.tab3c32q = d_tab3c32q.as<float>();
    v.ntab3c32q = ntab3c32q;
    v.mu = d_mu.as<int>();
    v.nu = d_nu.as<int>();
    v.mv = d_mv.as<int>();
    v.grid2c = d_grid2c.as<DGrid2c>();
    v.pden = d_pden.as<double>();
    v.canon_sp = canon_sp;
    v.wg2c = d_wg2c.as<int>();
    v.nwg2c = nwg2c;
    v.pair_disp = d_pair_disp.as<double>();
    v.cell_start = d_cell_start.as<int64_t>();
    v.cell_atoms = d_cell_atoms.as<int>();
    v.cell_bits = cell_bits;
    v.nstruct = (int)structs.size();
    v.natoms = natoms;
    v.st = d_st.as<DStruct>();
    v.xl = d_xl.as<double>();
    v.ratom = d_ratom.as<double>();
    v.imass = d_imass.as<int>();
    v.atom_struct = d_atom_struct.as<int>();
    v.dQ = d_dQ.as<double>();
    v.qin = d_qin.as<double>();
    v.npairs = npairs;
    v.neighn = d_neighn.as<int>();
    v.noff = d_noff.as<int64_t>();
    v.neigh_j = d_neigh_j.as<int>();
    v.neigh_b = d_neigh_b.as<int>();
    v.neigh_back = d_neigh_back.as<int>();
    v.neigh_self = d_neigh_self.as<int>();
    v.pair_atom = d_pair_atom.as<int>();
    v.boff = d_boff.as<int64_t>();
    v.pair_rec = d_pair_rec.as<int4>();
    v.pair_perm = lists_ok ? d_pair_perm.as<unsigned char>() : nullptr;
    v.wboff = d_wboff.as<int64_t>();
    v.ntrip = ntrip;
    v.nseg = nseg;
    v.trip = d_trip.as<int4>();
    v.trip_pair = d_tpair.as<int>();
    v.seg_off = d_seg_off.as<int64_t>();
    v.seg_pair = d_seg_pair.as<int>();
    v.pair_tgt = d_pair_tgt.as<int>();
    v.trip_aux = d_trip_aux.as<int>();
    for (int k = 0; k < 3; k++) {
      v.tiles[k] = d_tiles[k].as<int4>();
      v.ntiles[k] = ntiles[k];
    }
    v.trip_seg = d_trip_seg.as<int>();
    for (int i = 0; i < FB_NMAT; i++) v.mat[i] = d_mat[i].as<double>();
    v.mlp = d_mlp.as<DMlp>();
    v.mlp_err = d_mlp_err.as<float>();
    v.mlp_tau = (float)mlp_tau;
    v.mlp_count = d_mlp_count.as<unsigned long long>();
    v.status = d_status.as<int>();
    return v;
  }

  hipEvent_t ev() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    (void)hipEventCreate(&e);
    return e;
  }

  // run a launcher, bracketed by events when timing is on
  template <typename F> int launch(const char *name, F f) {
    hipEvent_t a = nullptr, b = nullptr;
    if (timing) {
      a = ev();
      b = ev();
      (void)hipEventRecord(a, stream);
    }
    int rc = f();
    if (timing) {
      (void)hipEventRecord(b, stream);
      events.push_back({name, a, b});
    }
    if (events.size() > 8192) {  // timing left on without readers: keep the newest
      hipStreamSynchronize(stream);
      pool.push_back(events.front().a);
      pool.push_back(events.front().b);
      events.erase(events.begin());
    }
    if (rc != 0) return fail(FB_E_HIP, name, (hipError_t)rc);
    return FB_OK;
  }

  void reset_events() {
    for (auto &e : events) {
      pool.push_back(e.a);
      pool.push_back(e.b);
    }
    events.clear();
  }

  int check_status(const char *where) {
    int st = 0;
    hipError_t e = hipMemcpyAsync(&st, d_status.p, sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e != hipSuccess) return fail(FB_E_HIP, where, e);
    e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail(FB_E_HIP, where, e);
    return decode_status(st, where);
  }

  // result of the asynchronous builds enqueued since the last check, if any:
  // their kernels OR their error bits into the status word, which stays set
  // until read here and cleared
  int settle(const char *where) {
    if (!pending) return FB_OK;
    pending = false;
    hipError_t e = hipMemcpyAsync(h_status, d_status.p, sizeof(int), hipMemcpyDeviceToHost, stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return fail(FB_E_HIP, where, e);
    const int st = *h_status;
    if (st) {
      e = hipMemsetAsync(d_status.p, 0, sizeof(int), stream);
      if (e == hipSuccess) e = hipStreamSynchronize(stream);
      if (e != hipSuccess) return fail(FB_E_HIP, where, e);
    }
    return decode_status(st, where);
  }

  int decode_status(int st, const char *where) {
    if (st == 0) return FB_OK;
    std::string w(where);
    if (st & ST_TOO_CLOSE) return fail(FB_E_ATOMS_TOO_CLOSE, (w + ": atoms dangerously close (< 0.5 A)").c_str());
    if (st & ST_BACK) return fail(FB_E_NEIGH_BACK, (w + ": neigh_back missing or not unique").c_str());
    if (st & ST_SELF) return fail(FB_E_NEIGH_SELF, (w + ": neigh_self missing").c_str());
    if (st & ST_X_NEG) return fail(FB_E_X_NEGATIVE, (w + ": negative x in 2c interpolation").c_str());
    if (st & ST_X_RANGE) return fail(FB_E_X_OUT_OF_RANGE, (w + ": x beyond the table range").c_str());
    if (st & ST_NO_TABLE) return fail(FB_E_NO_TABLE, (w + ": interaction table missing").c_str());
    if (st & ST_CAPACITY) return fail(FB_E_CAPACITY, (w + ": list exceeds kernel capacity").c_str());
    return fail(FB_E_HIP, where);
  }

  int clear_status() {
    hipError_t e = hipMemsetAsync(d_status.p, 0, sizeof(int), stream);
    if (e != hipSuccess) return fail(FB_E_HIP, "status clear", e);
    return FB_OK;
  }

  ~fb_ctx() {
    DevBuf *bufs[] = {&d_cell2c, &d_idx2c, &d_grp3c, &d_idx3c, &d_tab2c, &d_tab2c32, &d_tab3c, &d_tab3c32, &d_tab3c32q,
                      &d_mu, &d_nu, &d_mv, &d_grid2c, &d_pden, &d_wg2c, &d_cell_cnt, &d_cell_start, &d_cell_atoms, &d_atom_cell, &d_pair_disp, &d_tile_cnt, &d_tile_off, &d_gidx, &d_gout, &d_st, &d_xl, &d_ratom, &d_imass, &d_atom_struct, &d_dQ,
                      &d_neighn, &d_noff, &d_neigh_j, &d_neigh_b, &d_neigh_back, &d_neigh_self, &d_pair_atom, &d_pair_rec, &d_pair_perm,
                      &d_bsz, &d_boff, &d_ncommon, &d_toff, &d_trip_orig, &d_tpair_orig, &d_trip, &d_tpair,
                      &d_perm_in, &d_perm, &d_flag, &d_segincl, &d_seg_off, &d_seg_pair, &d_trip_seg, &d_pair_tgt, &d_trip_aux, &d_tiles[0], &d_tiles[1], &d_tiles[2], &d_tmp,
                      &d_mlp, &d_mlp_err, &d_mlp_count, &d_queue, &d_meth3c,
                      &d_status, &d_idx3r, &d_qin, &d_wbsz, &d_wboff, &d_pp_map, &d_px_map, &d_px_point,
                      &d_inv_pp, &d_inv_px, &d_cand_cnt, &d_cand_off, &d_cand_j, &d_cand_b, &d_pc_n, &d_pc_off,
                      &d_pc_trip, &d_pc_tpair, &d_pc_trip_s, &d_pc_tpair_s, &d_pc_perm_in, &d_pc_perm, &d_pc_flag,
                      &d_pc_segincl, &d_pc_seg_off, &d_pc_seg_pair, &d_pc_trip_seg, &d_pc_tgt, &d_pc_mp, &d_pc_tiles, &d_ppp_disp, &d_pk_size, &d_pk_off, &d_pk_key, &d_pk_idx, &d_sv_bsz, &d_sv_off,
                      &d_vl_bsz, &d_vl_off, &d_svnl, &d_vnl};
    for (DevBuf *b : bufs) b->release();
    for (auto &m : d_mat) m.release();
    for (int l = 0; l < 3; l++)
      for (DevBuf *b : {&d_ppn[l], &d_ppoff[l], &d_ppj[l], &d_ppb[l], &d_ppatom[l], &d_ppself[l]}) b->release();
    DevBuf *fbufs[] = {&d_rho, &d_dmat[0], &d_dmat[1], &d_dmat[2], &d_fatom, &d_fontop, &d_f3a, &d_f3b, &d_f3c,
                       &d_f3, &d_jkey, &d_jkey_s, &d_iota_t, &d_perm_j, &d_inv_perm, &d_inv_perm_j, &d_ioff, &d_joff};
    for (DevBuf *b : fbufs) b->release();
    reset_events();
    for (auto e : pool) (void)hipEventDestroy(e);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    if (h_status) (void)hipHostFree(h_status);
    if (own_stream) (void)hipStreamDestroy(own_stream);
    if (side) (void)hipStreamDestroy(side);
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (ev_join) (void)hipEventDestroy(ev_join);
  }
};

// ------------------------------------------------------------------ helpers
static int upload(fb_ctx *ctx, DevBuf &b, const void *src, size_t bytes, const char *what) {
  hipError_t e = b.ensure(bytes);
  if (e != hipSuccess) return ctx->fail(FB_E_HIP, what, e);
  if (bytes) {
    e = hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) return ctx->fail(FB_E_HIP, what, e);
  }
  return FB_OK;
}

#define RC(x)                 \
  do {                        \
    int rc_ = (x);            \
    if (rc_ != FB_OK) return rc_; \
  } while (0)

// make_cells, c.SYSTEM/make_cells.f90:115-144 (same expression order)
static int make_cells(const double *lat, int icluster, double *xl) {
  xl[0] = xl[1] = xl[2] = 0.0;
  if (icluster == 1) return 0;
  int mbeta = 1;
  for (int iz = -4; iz <= 4; iz++)
    for (int iy = -4; iy <= 4; iy++)
      for (int ix = -4; ix <= 4; ix++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        for (int k = 0; k < 3; k++)
          xl[3 * mbeta + k] = (double)ix * lat[k] + (double)iy * lat[3 + k] + (double)iz * lat[6 + k];
        mbeta++;
      }
  return mbeta - 1;
}

// ------------------------------------------------------------------- C ABI
extern "C" {

// why the last fb_create of this thread failed (fb_last_error(NULL) returns it)
// ---------------------------------------------- scratch of the allocation-free entry point
// fb_mlp_gradients_noalloc keeps the work arrays of one training batch (the
// MlpWork arrays for FB_MLP_NOALLOC_MAX_BATCH samples, laid out at fixed
// offsets) in this file-scope device array instead of allocating them.  Each
// GPU holds its own copy until the process ends.  The host never passes the
// array's name to a kernel: fb_create looks up its device address once per
// device (with that device current) and every use goes through that address.
namespace {
constexpr int64_t NA_B = FB_MLP_NOALLOC_MAX_BATCH;
constexpr int64_t NA_HID = NA_B * MLP_H;            // floats of one [B][64] array
constexpr int64_t NA_OFF_Y = 7 * NA_HID;            // Z1, A1, Z2, A2, dA, dZ2, dZ1 first
constexpr int64_t NA_OFF_DY = NA_OFF_Y + NA_B * MLP_O;
constexpr int64_t NA_OFF_ROW = NA_OFF_DY + NA_B * MLP_O;
constexpr int64_t NA_OFF_PART = NA_OFF_ROW + NA_B;
constexpr int64_t NA_OFF_ONES = NA_OFF_PART + 64LL * MLP_H * MLP_O;
constexpr int64_t NA_FLOATS = NA_OFF_ONES + 64;     // ones[1], padded to 256 bytes
constexpr int NA_MAXDEV = 64;
std::mutex g_scratch_mu;
float *g_scratch_dev[NA_MAXDEV] = {};
}  // namespace
static __device__ __align__(256) float scratch[NA_FLOATS];

// device address of this GPU's scratch array; the device must be current
static hipError_t scratch_address(int device, float **out) {
  if (device < 0 || device >= NA_MAXDEV) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_dev[device] = static_cast<float *>(ptr);
  }
  *out = g_scratch_dev[device];
  return hipSuccess;
}

static thread_local std::string g_create_err = "null context";

static int create_failed(fb_ctx *ctx, const char *what, hipError_t e) {
  delete ctx;
  g_create_err = std::string("fb_create: ") + what + ": " + hipGetErrorString(e);
  return FB_E_HIP;
}

int fb_create(fb_ctx **out, int device) {
  if (!out) return FB_E_ARG;
  *out = nullptr;
  fb_ctx *ctx = new fb_ctx();
  ctx->device = device;
  hipError_t e = hipSetDevice(device);
  if (e != hipSuccess) return create_failed(ctx, "hipSetDevice", e);
  e = hipStreamCreateWithFlags(&ctx->own_stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_fork, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_join, hipEventDisableTiming);
  if (e != hipSuccess) return create_failed(ctx, "streams", e);
  e = scratch_address(device, &ctx->scratch);
  if (e != hipSuccess) return create_failed(ctx, "scratch address", e);
  e = hipDeviceGetAttribute(&ctx->ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return create_failed(ctx, "device attributes", e);
  ctx->stream = ctx->own_stream;
  e = ctx->d_status.ensure(sizeof(int) * 4);
  if (e == hipSuccess) e = hipMemset(ctx->d_status.p, 0, sizeof(int) * 4);
  if (e == hipSuccess) e = hipHostMalloc((void **)&ctx->h_status, sizeof(int) * 4);
  if (e != hipSuccess) return create_failed(ctx, "status buffers", e);
  ctx->h_status[0] = 0;
  *out = ctx;
  return FB_OK;
}

int fb_destroy(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  delete ctx;
  return FB_OK;
}

const char *fb_last_error(const fb_ctx *ctx) { return ctx ? ctx->err.c_str() : g_create_err.c_str(); }

int fb_read_positions(const char *basisfile, int nspecies, const int *nZ, const int *norb, int max_atoms,
                      int *natoms, int *icluster, double *lattice, double *ratom, int *imass, int *norbitals,
                      int *iblock_slot, int *shifted) {
  if (!basisfile || nspecies < 1 || !nZ || !norb || max_atoms < 0 || !natoms || !icluster || !lattice || !ratom ||
      !imass || !norbitals || !shifted) {
    g_create_err = "read_positions arguments";
    return FB_E_ARG;
  }
  std::string err;
  const int rc = fb_positions_read(basisfile, nspecies, nZ, norb, max_atoms, natoms, icluster, lattice, ratom, imass,
                                   norbitals, iblock_slot, shifted, err);
  if (rc == 0) return FB_OK;
  g_create_err = "read_positions: " + err;
  return rc == -2 ? FB_E_CAPACITY : FB_E_ARG;
}

int fb_set_stream(fb_ctx *ctx, void *s) {
  if (!ctx) return FB_E_ARG;
  ctx->stream = s ? (hipStream_t)s : ctx->own_stream;
  ctx->graph_key = -1;  // a captured graph is tied to its stream's device
  return FB_OK;
}

int fb_synchronize(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  FB_HIP(hipStreamSynchronize(ctx->stream));
  return ctx->settle("assemble_all");
}

int fb_set_precision(fb_ctx *ctx, int mode) {
  if (!ctx || (mode != FB_PREC_FP64 && mode != FB_PREC_MIXED)) return FB_E_ARG;
  ctx->precision = mode;
  return FB_OK;
}

int fb_set_3c_method(fb_ctx *ctx, int method) {
  if (!ctx || (method != FB_3C_TABLE && method != FB_3C_MLP)) return FB_E_ARG;
  ctx->method3c = method;
  return FB_OK;
}

int fb_set_species(fb_ctx *ctx, int ns, const int *nZ, const int *nssh, const int *lssh, const double *rcA) {
  if (!ctx || ns < 1 || ns > FB_MAXSPECIES || !nssh || !lssh || !rcA) return ctx ? ctx->fail(FB_E_ARG, "species") : FB_E_ARG;
  DSpecies sp{};
  sp.ns = ns;
  for (int i = 0; i < ns; i++) {
    if (nssh[i] < 1 || nssh[i] > FB_MAXSH) return ctx->fail(FB_E_ARG, "species: nssh out of range");
    sp.nssh[i] = nssh[i];
    sp.norb[i] = 0;
    sp.rcmax[i] = -99.0;
    sp.rcmin[i] = 99.0;
    for (int k = 0; k < nssh[i]; k++) {
      int l = lssh[i * FB_MAXSH + k];
      if (l < 0 || l > 2) return ctx->fail(FB_E_UNSUPPORTED, "species: shell with l > 2");
      if (l == 2) sp.have_d = 1;
      sp.lssh[i][k] = l;
      sp.norb[i] += 2 * l + 1;
      // c.SYSTEM/M_neighbors.f90:299-304 (max) and M_assemble_2c_DOGS.f90:764-771 (min)
      sp.rcmax[i] = std::max(sp.rcmax[i], rcA[i * FB_MAXSH + k]);
      sp.rcmin[i] = std::min(sp.rcmin[i], rcA[i * FB_MAXSH + k]);
    }
    if (sp.norb[i] > FB_MOD) return ctx->fail(FB_E_UNSUPPORTED, "species: more than 9 orbitals");
    ctx->nZ[i] = nZ ? nZ[i] : 0;
  }
  ctx->sp = sp;
  ctx->maxsh = 1;
  for (int i = 0; i < ns; i++) ctx->maxsh = std::max(ctx->maxsh, sp.nssh[i]);
  ctx->species_set = true;
  ctx->tables_ok = false;
  ctx->lists_ok = false;
  ctx->pp_lists_ok = false;
  return FB_OK;
}

int fb_set_table_2c(fb_ctx *ctx, int is, int js, int itype, int isub, int nME, int nx, double xmax, const int *mu,
                    const int *nu, const int *mvalue, const double *fdata) {
  if (!ctx || !ctx->species_set) return ctx ? ctx->fail(FB_E_ORDER, "set_table_2c before set_species") : FB_E_ARG;
  if (is < 1 || js < 1 || is > ctx->sp.ns || js > ctx->sp.ns || isub < 0 || isub >= FB_MAXSUB2 || nx < 7 ||
      nME < 1 || nME > FB_NME3_WIDE || !mu || !nu || !mvalue || !fdata)
    return ctx->fail(FB_E_ARG, "set_table_2c: bad arguments");
  HCell2c c;
  c.is = is; c.js = js; c.itype = itype; c.isub = isub; c.nME = nME; c.nx = nx; c.xmax = xmax;
  c.mu.assign(mu, mu + nME);
  c.nu.assign(nu, nu + nME);
  c.mv.assign(mvalue, mvalue + nME);
  c.data.resize((size_t)nx * nME);
  for (int m = 0; m < nME; m++)
    for (int ix = 0; ix < nx; ix++) c.data[(size_t)ix * nME + m] = fdata[(size_t)m * nx + ix];
  for (auto &o : ctx->c2)
    if (o.is == is && o.js == js && o.itype == itype && o.isub == isub) { o = std::move(c); ctx->tables_ok = false; return FB_OK; }
  ctx->c2.push_back(std::move(c));
  ctx->tables_ok = false;
  return FB_OK;
}

int fb_set_table_3c(fb_ctx *ctx, int is, int js, int ks, int itype, int isub, int itheta, int nME, int nx,
                    double xmax, int ny, double ymax, const int *mu, const int *nu, const int *mvalue,
                    const double *fdata) {
  if (!ctx || !ctx->species_set) return ctx ? ctx->fail(FB_E_ORDER, "set_table_3c before set_species") : FB_E_ARG;
  if (is < 1 || js < 1 || ks < 1 || is > ctx->sp.ns || js > ctx->sp.ns || ks > ctx->sp.ns || isub < 0 ||
      isub >= FB_MAXSUB3 || itheta < 1 || itheta > FB_NTHETA || nx < 4 || ny < 4 || nME < 1 || nME > FB_NME3_WIDE || !mu ||
      !nu || !mvalue || !fdata)
    return ctx->fail(FB_E_ARG, "set_table_3c: bad arguments");
  HCell3c c;
  c.is = is; c.js = js; c.ks = ks; c.itype = itype; c.isub = isub; c.itheta = itheta;
  c.nME = nME; c.nx = nx; c.ny = ny; c.xmax = xmax; c.ymax = ymax;
  c.mu.assign(mu, mu + nME);
  c.nu.assign(nu, nu + nME);
  c.mv.assign(mvalue, mvalue + nME);
  c.data.resize((size_t)nx * ny * nME);
  for (int m = 0; m < nME; m++)
    for (int iy = 0; iy < ny; iy++)
      for (int ix = 0; ix < nx; ix++)
        c.data[((size_t)iy * nx + ix) * nME + m] = fdata[((size_t)m * ny + iy) * nx + ix];
  for (auto &o : ctx->c3)
    if (o.is == is && o.js == js && o.ks == ks && o.itype == itype && o.isub == isub && o.itheta == itheta) {
      o = std::move(c);
      ctx->tables_ok = false;
      return FB_OK;
    }
  ctx->c3.push_back(std::move(c));
  ctx->tables_ok = false;
  return FB_OK;
}

namespace {
bool fdata_type2c(int itype) { return slot_of(itype) >= 0; }
// rho_3c 1, rhoS_3c 2, bcna 3 (include/interactions_3c.h:6-8)
bool fdata_type3c(int itype) { return itype >= 1 && itype <= 3; }

struct CtxSink : FdataSink {
  fb_ctx *ctx;
  explicit CtxSink(fb_ctx *c) : ctx(c) {}
  int cell2c(int is, int js, int itype, int isub, int nME, int nx, double xmax, const int *mu, const int *nu,
             const int *mv, const double *fdata) override {
    return fb_set_table_2c(ctx, is, js, itype, isub, nME, nx, xmax, mu, nu, mv, fdata);
  }
  int cell3c(int is, int js, int ks, int itype, int isub, int itheta, int nME, int nx, double xmax, int ny,
             double ymax, const int *mu, const int *nu, const int *mv, const double *fdata) override {
    return fb_set_table_3c(ctx, is, js, ks, itype, isub, itheta, nME, nx, xmax, ny, ymax, mu, nu, mv, fdata);
  }
};
}  // namespace

int fb_read_fdata(fb_ctx *ctx, const char *fdata_dir) {
  if (!ctx || !fdata_dir) return FB_E_ARG;
  if (!ctx->species_set) return ctx->fail(FB_E_ORDER, "read_fdata before set_species");
  for (int i = 0; i < ctx->sp.ns; i++)
    if (ctx->nZ[i] < 1) return ctx->fail(FB_E_ARG, "read_fdata: species need their atomic numbers");
  CtxSink sink(ctx);
  std::string err;
  FdataWanted want{fdata_type2c, fdata_type3c};
  int rc = fb_fdata_read(fdata_dir, ctx->sp.ns, ctx->nZ, want, sink, err);
  if (rc == -1) return ctx->fail(FB_E_NO_TABLE, ("read_fdata: " + err).c_str());
  if (rc == 0 && ctx->pp_species) {
    double cl[FB_MAXSPECIES * FB_MAXSH] = {0.0};
    int nsh[FB_MAXSPECIES] = {0};
    for (int i = 0; i < ctx->sp.ns; i++) nsh[i] = ctx->hpp.nssh_pp[i];
    if (fb_fdata_read_cl(fdata_dir, ctx->sp.ns, ctx->nZ, nsh, cl, err))
      return ctx->fail(FB_E_NO_TABLE, ("read_fdata: " + err).c_str());
    for (int i = 0; i < ctx->sp.ns; i++)
      for (int k = 0; k < nsh[i]; k++) ctx->hpp.cl[i][k] = cl[i * FB_MAXSH + k];
  }
  return rc;
}

int fb_commit_tables(fb_ctx *ctx) {
  if (!ctx || !ctx->species_set) return ctx ? ctx->fail(FB_E_ORDER, "commit_tables before set_species") : FB_E_ARG;
  FB_HIP(hipSetDevice(ctx->device));
  const int ns = ctx->sp.ns;
  const bool wide = fb_wide_species(ctx->sp);
  std::vector<int> idx2((size_t)ns * ns * FB_NSLOT2 * FB_MAXSUB2, -1);
  std::vector<DCell2c> cells;
  std::vector<double> tab2;
  std::vector<int> mu, nu, mv;
  for (auto &c : ctx->c2) {
    int slot = slot_of(c.itype);
    if (slot < 0) continue;  // interactions outside the assembled path
    if (slot == SLOT_VNL && !ctx->pp_species) continue;   // projector tables need the PP shells (fb_set_species_pp)
    {
      // the block of an atom term lives on the first atom's orbitals (make_munu_atom);
      // the shell-averaged density tables index shells (make_munuS, make_munuS_atom)
      const bool shell = (slot == SLOT_RHOS_ONL || slot == SLOT_RHOS_ONR || slot == SLOT_RHOS_ATOM);
      const bool atom = (slot == SLOT_ATOM || slot == SLOT_RHO_ATOM || slot == SLOT_RHOS_ATOM);
      const int *dim = shell ? ctx->sp.nssh : ctx->sp.norb;
      int rows = dim[c.is - 1], cols = atom ? rows : dim[c.js - 1];
      if (slot == SLOT_VNL) cols = ctx->hpp.norb_pp[c.js - 1];   // PP projectors of the second species
      if (c.nME > (wide ? FB_NME2_WIDE : 16))
        return ctx->fail(FB_E_UNSUPPORTED, "commit_tables: 2c cell with more matrix elements than the kernels hold");
      for (int k = 0; k < c.nME; k++)
        if (c.mu[k] < 1 || c.mu[k] > rows || c.nu[k] < 1 || c.nu[k] > cols)
          return ctx->fail(FB_E_ARG, "commit_tables: 2c mu/nu outside the orbital block");
    }
    DCell2c d{};
    d.nME = c.nME;
    d.nx = c.nx;
    d.xmax = c.xmax;
    d.dx = c.xmax / (double)(c.nx - 1);  // e.FDATA/M_Fdata_2c.f90:222
    if (tab2.size() & 1) tab2.push_back(0.0);   // cells start on 16-byte boundaries (paired loads)
    d.off = (int64_t)tab2.size();
    d.off32 = d.off;
    d.munu = (int)mu.size();
    tab2.insert(tab2.end(), c.data.begin(), c.data.end());
    mu.insert(mu.end(), c.mu.begin(), c.mu.end());
    nu.insert(nu.end(), c.nu.begin(), c.nu.end());
    mv.insert(mv.end(), c.mv.begin(), c.mv.end());
    idx2[(((size_t)(c.is - 1) * ns + (c.js - 1)) * FB_NSLOT2 + slot) * FB_MAXSUB2 + c.isub] = (int)cells.size();
    cells.push_back(d);
  }
  // 3c: group the five theta cells of each (is, js, ks, isub) of bcna (itype 3)
  std::map<std::tuple<int, int, int, int>, std::vector<const HCell3c *>> groups;
  for (auto &c : ctx->c3)
    if (c.itype == 3) groups[std::make_tuple(c.is, c.js, c.ks, c.isub)].push_back(&c);
  std::vector<int> idx3((size_t)ns * ns * ns * FB_MAXSUB3, -1);
  std::vector<DGroup3c> grps;
  std::vector<std::array<int, 4>> grpkey;
  std::vector<unsigned char> meth;
  // method of every interior cell of a group: gradient of ME 1 of theta cell
  // 1 (data t0, layout [ny][nx][nME]) against 1e-6 (e.FDATA/M_Fdata_3c.f90:
  // 383-385), the same binary64 expression the kernels would evaluate per
  // triplet; 1 = the 3-point rule
  auto method_table = [&](DGroup3c &g, const double *t0) {
    g.moff = (int64_t)meth.size();
    meth.resize(meth.size() + (size_t)g.nx * g.ny, 0);
    for (int iy = 2; iy <= g.ny - 2; iy++)
      for (int ix = 2; ix <= g.nx - 2; ix++) {
        const size_t i00 = ((size_t)(iy - 1) * g.nx + (ix - 1)) * g.nME;
        const double f00 = t0[i00];
        const double gx = (t0[i00 + g.nME] - f00) / g.dx;
        const double gy = (t0[i00 + (size_t)g.nx * g.nME] - f00) / g.dy;
        meth[g.moff + (size_t)(iy - 1) * g.nx + (ix - 1)] = (gx * gx + gy * gy < 1.0e-6) ? 1 : 0;
      }
  };
  std::vector<double> tab3;
  std::vector<float> tab3f, tab3q;
  for (auto &kv : groups) {
    const HCell3c *th[FB_NTHETA] = {nullptr};
    for (auto *c : kv.second) th[c->itheta - 1] = c;
    for (int t = 0; t < FB_NTHETA; t++)
      if (!th[t]) return ctx->fail(FB_E_NO_TABLE, "commit_tables: a 3c group lacks a theta cell");
    const HCell3c *c1 = th[0];
    if (c1->nME > (wide ? FB_NME3_WIDE : 10))
      return ctx->fail(FB_E_UNSUPPORTED, "commit_tables: bcna cell with more matrix elements than the kernels hold");
    for (int t = 0; t < FB_NTHETA; t++)
      for (int k = 0; k < th[t]->nME; k++)
        if (th[t]->mu[k] < 1 || th[t]->mu[k] > ctx->sp.norb[th[t]->is - 1] || th[t]->nu[k] < 1 ||
            th[t]->nu[k] > ctx->sp.norb[th[t]->js - 1])
          return ctx->fail(FB_E_ARG, "commit_tables: 3c mu/nu outside the orbital block");
    for (int t = 1; t < FB_NTHETA; t++)
      if (th[t]->nME != c1->nME || th[t]->nx != c1->nx || th[t]->ny != c1->ny || th[t]->xmax != c1->xmax ||
          th[t]->ymax != c1->ymax)
        return ctx->fail(FB_E_UNSUPPORTED, "commit_tables: theta cells of a 3c group differ in grid");
    DGroup3c g{};
    g.nME = c1->nME;
    g.nx = c1->nx;
    g.ny = c1->ny;
    g.xmax = c1->xmax;
    g.ymax = c1->ymax;
    g.dx = c1->xmax / (c1->nx - 1.0);  // e.FDATA/M_Fdata_3c.f90:367-368
    g.dy = c1->ymax / (c1->ny - 1.0);
    g.rdx = 1.0 / g.dx;
    g.rdy = 1.0 / g.dy;
    g.cls = (ctx->sp.norb[c1->is - 1] == 4 ? 2 : 0) + (ctx->sp.norb[c1->js - 1] == 4 ? 1 : 0);
    g.nq = (c1->nME + 3) / 4;
    g.off = (int64_t)tab3.size();
    g.off32 = (int64_t)tab3f.size();
    g.off32q = (int64_t)tab3q.size();
    // the reference recovers the block with the mu/nu/mvalue of the last theta cell
    const HCell3c *cl = th[FB_NTHETA - 1];
    g.munu = (int)mu.size();
    mu.insert(mu.end(), cl->mu.begin(), cl->mu.end());
    nu.insert(nu.end(), cl->nu.begin(), cl->nu.end());
    mv.insert(mv.end(), cl->mv.begin(), cl->mv.end());
    for (int t = 0; t < FB_NTHETA; t++) tab3.insert(tab3.end(), th[t]->data.begin(), th[t]->data.end());
    method_table(g, th[0]->data.data());
    // FP32 slab [iy][ix][ME][theta]: the five values of one (point, ME) are
    // contiguous and load as one float4 + one float
    size_t plane = (size_t)g.nx * g.ny;
    tab3f.resize(tab3f.size() + plane * g.nME * FB_TH32, 0.0f);
    float *dst = tab3f.data() + g.off32;
    for (size_t q = 0; q < plane; q++)
      for (int m = 0; m < g.nME; m++)
        for (int t = 0; t < FB_NTHETA; t++) dst[(q * g.nME + m) * FB_TH32 + t] = (float)th[t]->data[q * g.nME + m];
    // FP32 quad slab [iy][ix][theta][4 nq]: one float4 holds four MEs of one
    // theta cell at one point
    const size_t qrow = (size_t)4 * g.nq;
    tab3q.resize(tab3q.size() + plane * FB_NTHETA * qrow, 0.0f);
    float *dq = tab3q.data() + g.off32q;
    for (size_t q = 0; q < plane; q++)
      for (int t = 0; t < FB_NTHETA; t++)
        for (int m = 0; m < g.nME; m++) dq[(q * FB_NTHETA + t) * qrow + m] = (float)th[t]->data[q * g.nME + m];
    int is, js, ks, isub;
    std::tie(is, js, ks, isub) = kv.first;
    // the mixed kernel addresses every charge subtype with the neutral group's stencil
    if (isub > 0) {
      auto k0 = groups.find(std::make_tuple(is, js, ks, 0));
      if (k0 == groups.end() || k0->second[0]->nx != g.nx || k0->second[0]->ny != g.ny ||
          k0->second[0]->nME != g.nME || k0->second[0]->xmax != g.xmax || k0->second[0]->ymax != g.ymax)
        return ctx->fail(FB_E_UNSUPPORTED, "commit_tables: 3c charge subtypes must share the neutral grid");
    }
    idx3[(((size_t)(is - 1) * ns + (js - 1)) * ns + (ks - 1)) * FB_MAXSUB3 + isub] = (int)grps.size();
    grps.push_back(g);
    grpkey.push_back({is, js, ks, isub});
  }
  const size_t nbcna = grps.size();
  // McWEDA density groups: rho_3c (itype 1, orbital blocks, make_munu_3c) and
  // rhoS_3c (itype 2, shell blocks, make_munuS_3c), subtypes isorp =
  // 1..nssh(third atom); the binary64 density kernels read the FP64 slab only
  std::map<std::tuple<int, int, int, int, int>, std::vector<const HCell3c *>> rgroups;
  for (auto &c : ctx->c3)
    if (c.itype == 1 || c.itype == 2) rgroups[std::make_tuple(c.itype, c.is, c.js, c.ks, c.isub)].push_back(&c);
  std::vector<int> idx3r((size_t)ns * ns * ns * 2 * FB_MAXSUB3, -1);
  for (auto &kv : rgroups) {
    int itype, is, js, ks, isub;
    std::tie(itype, is, js, ks, isub) = kv.first;
    const HCell3c *th[FB_NTHETA] = {nullptr};
    for (auto *c : kv.second) th[c->itheta - 1] = c;
    for (int t = 0; t < FB_NTHETA; t++)
      if (!th[t]) return ctx->fail(FB_E_NO_TABLE, "commit_tables: a 3c density group lacks a theta cell");
    const HCell3c *c1 = th[0];
    const int *dim = (itype == 2) ? ctx->sp.nssh : ctx->sp.norb;
    for (int t = 0; t < FB_NTHETA; t++) {
      if (th[t]->nME != c1->nME || th[t]->nx != c1->nx || th[t]->ny != c1->ny || th[t]->xmax != c1->xmax ||
          th[t]->ymax != c1->ymax)
        return ctx->fail(FB_E_UNSUPPORTED, "commit_tables: theta cells of a 3c group differ in grid");
      if (th[t]->nME > (wide ? FB_NME3_WIDE : 16))
        return ctx->fail(FB_E_UNSUPPORTED, "commit_tables: 3c cell with more matrix elements than the kernels hold");
      for (int k = 0; k < th[t]->nME; k++)
        if (th[t]->mu[k] < 1 || th[t]->mu[k] > dim[is - 1] || th[t]->nu[k] < 1 || th[t]->nu[k] > dim[js - 1])
          return ctx->fail(FB_E_ARG, "commit_tables: 3c mu/nu outside the block");
    }
    if (isub < 1 || isub >= FB_MAXSUB3) return ctx->fail(FB_E_ARG, "commit_tables: 3c density subtype out of range");
    DGroup3c g{};
    g.nME = c1->nME;
    g.nx = c1->nx;
    g.ny = c1->ny;
    g.xmax = c1->xmax;
    g.ymax = c1->ymax;
    g.dx = c1->xmax / (c1->nx - 1.0);
    g.dy = c1->ymax / (c1->ny - 1.0);
    g.off = (int64_t)tab3.size();
    g.off32 = -1;
    method_table(g, th[0]->data.data());
    const HCell3c *cl = th[FB_NTHETA - 1];
    g.munu = (int)mu.size();
    mu.insert(mu.end(), cl->mu.begin(), cl->mu.end());
    nu.insert(nu.end(), cl->nu.begin(), cl->nu.end());
    mv.insert(mv.end(), cl->mv.begin(), cl->mv.end());
    for (int t = 0; t < FB_NTHETA; t++) tab3.insert(tab3.end(), th[t]->data.begin(), th[t]->data.end());
    idx3r[((((size_t)(is - 1) * ns + (js - 1)) * ns + (ks - 1)) * 2 + (itype - 1)) * FB_MAXSUB3 + isub] =
        (int)grps.size();
    grps.push_back(g);
  }
  // distinct 2c grids and their Lagrange denominators (host binary64, same
  // expression as e.FDATA/M_Fdata_2c.f90:379-385)
  std::vector<DGrid2c> grids;
  std::vector<double> pden;
  for (auto &d : cells) {
    int gidx = -1;
    for (size_t k = 0; k < grids.size(); k++)
      if (grids[k].nx == d.nx && grids[k].dx == d.dx) gidx = (int)k;
    if (gidx < 0) {
      DGrid2c gr{};
      gr.nx = d.nx;
      gr.dx = d.dx;
      gr.pden_off = (int64_t)pden.size();
      for (int ileft = 1; ileft <= d.nx - 6; ileft++) {
        double xx[7];
        for (int k = 0; k < 7; k++) xx[k] = (double)(ileft + k - 1) * d.dx;
        for (int is = 0; is < 7; is++)
          for (int ip = 0; ip < 7; ip++) pden.push_back(ip == is ? 0.0 : 1.0 / (xx[is] - xx[ip]));
      }
      gidx = (int)grids.size();
      grids.push_back(gr);
    }
    d.grid = gidx;
  }
  // s / p species with canonical matrix-element order -> register-resident kernels
  int canon = 1;
  std::vector<std::vector<int>> shells(ns);
  for (int i = 0; i < ns; i++) {
    for (int k = 0; k < ctx->sp.nssh[i]; k++) shells[i].push_back(ctx->sp.lssh[i][k]);
    bool ok = (shells[i] == std::vector<int>{0}) || (shells[i] == std::vector<int>{0, 1});
    if (!ok) canon = 0;
  }
  if (canon) {
    for (auto &c : ctx->c2) {
      int slot = slot_of(c.itype);
      if (slot < 0 || slot > SLOT_DIP) continue;   // the s/p kernels serve the H+S interactions
      MeList e = (slot == SLOT_ATOM) ? munu2c(shells[c.is - 1], shells[c.is - 1])
                                     : munu2c(shells[c.is - 1], shells[c.js - 1]);
      if (!same_list(e, c.mu, c.nu, c.mv)) canon = 0;
    }
    for (auto &c : ctx->c3) {
      if (c.itype != 3) continue;
      MeList e = munu3c_sp(shells[c.is - 1], shells[c.js - 1]);
      if (!same_list(e, c.mu, c.nu, c.mv)) canon = 0;
    }
  }
  ctx->canon_sp = canon;
  // density tables: the s/p density kernels need every species s or s+p and the
  // canonical orders of make_munu / make_munu_atom / make_munuS / make_munu_3c /
  // make_munuS_3c
  {
    int cr = 1;
    for (int i = 0; i < ns; i++)
      if (!((shells[i] == std::vector<int>{0}) || (shells[i] == std::vector<int>{0, 1}))) cr = 0;
    auto munuS = [](int a, int b) {
      MeList r;
      for (int i = 1; i <= a; i++)
        for (int j = 1; j <= b; j++) { r.mu.push_back(i); r.nu.push_back(j); r.mv.push_back(0); }
      return r;
    };
    if (cr) {
      for (auto &c : ctx->c2) {
        int slot = slot_of(c.itype);
        if (slot < SLOT_RHO_ONL || slot > SLOT_RHOS_ATOM) continue;
        const int si = (int)shells[c.is - 1].size(), sj = (int)shells[c.js - 1].size();
        MeList e = slot == SLOT_RHO_ATOM ? munu2c(shells[c.is - 1], shells[c.is - 1])
                 : slot == SLOT_RHOS_ATOM ? munuS(si, si)
                 : (slot == SLOT_RHOS_ONL || slot == SLOT_RHOS_ONR) ? munuS(si, sj)
                 : munu2c(shells[c.is - 1], shells[c.js - 1]);
        if (!same_list(e, c.mu, c.nu, c.mv)) cr = 0;
      }
      for (auto &c : ctx->c3) {
        if (c.itype != 1 && c.itype != 2) continue;
        MeList e = c.itype == 1 ? munu3c_sp(shells[c.is - 1], shells[c.js - 1])
                                : munuS((int)shells[c.is - 1].size(), (int)shells[c.js - 1].size());
        if (!same_list(e, c.mu, c.nu, c.mv)) cr = 0;
      }
    }
    ctx->canon_rho = cr;
    // projector tables: s/p species and s/p PP shells, make_munu order with the PP shells on the right
    int cp = ctx->pp_species ? 1 : 0;
    std::vector<std::vector<int>> ppsh(ns);
    for (int i = 0; i < ns && cp; i++) {
      for (int k = 0; k < ctx->hpp.nssh_pp[i]; k++) ppsh[i].push_back(ctx->hpp.lssh_pp[i][k]);
      if (!((shells[i] == std::vector<int>{0}) || (shells[i] == std::vector<int>{0, 1}))) cp = 0;
      if (!(ppsh[i].empty() || ppsh[i] == std::vector<int>{0} || ppsh[i] == std::vector<int>{0, 1})) cp = 0;
    }
    for (auto &c : ctx->c2) {
      if (!cp) break;
      if (slot_of(c.itype) != SLOT_VNL) continue;
      if (!same_list(munu2c(shells[c.is - 1], ppsh[c.js - 1]), c.mu, c.nu, c.mv)) cp = 0;
    }
    ctx->hpp.canon = cp;
  }
  std::vector<float> tab2f(tab2.begin(), tab2.end());
  RC(upload(ctx, ctx->d_idx2c, idx2.data(), idx2.size() * sizeof(int), "upload idx2c"));
  RC(upload(ctx, ctx->d_grp3c, grps.data(), grps.size() * sizeof(DGroup3c), "upload groups3c"));
  RC(upload(ctx, ctx->d_idx3c, idx3.data(), idx3.size() * sizeof(int), "upload idx3c"));
  RC(upload(ctx, ctx->d_idx3r, idx3r.data(), idx3r.size() * sizeof(int), "upload idx3r"));
  RC(upload(ctx, ctx->d_tab2c, tab2.data(), tab2.size() * sizeof(double), "upload tab2c"));
  RC(upload(ctx, ctx->d_tab2c32, tab2f.data(), tab2f.size() * sizeof(float), "upload tab2c32"));
  RC(upload(ctx, ctx->d_tab3c, tab3.data(), tab3.size() * sizeof(double), "upload tab3c"));
  RC(upload(ctx, ctx->d_meth3c, meth.data(), meth.size(), "upload 3c methods"));
  if (tab3f.size() * sizeof(float) >= (size_t)INT32_MAX)
    return ctx->fail(FB_E_CAPACITY, "commit_tables: FP32 3c slab above 2 GiB");
  RC(upload(ctx, ctx->d_tab3c32, tab3f.data(), tab3f.size() * sizeof(float), "upload tab3c32"));
  ctx->ntab3c32 = (int64_t)tab3f.size();
  if (tab3q.size() * sizeof(float) >= (size_t)INT32_MAX)
    return ctx->fail(FB_E_CAPACITY, "commit_tables: FP32 3c quad slab above 2 GiB");
  RC(upload(ctx, ctx->d_tab3c32q, tab3q.data(), tab3q.size() * sizeof(float), "upload tab3c32q"));
  ctx->ntab3c32q = (int64_t)tab3q.size();
  RC(upload(ctx, ctx->d_cell2c, cells.data(), cells.size() * sizeof(DCell2c), "upload cells2c"));
  ctx->ncell2c = (int)cells.size();
  ctx->ngrid2c = (int)grids.size();
  RC(upload(ctx, ctx->d_grid2c, grids.data(), grids.size() * sizeof(DGrid2c), "upload grids2c"));
  RC(upload(ctx, ctx->d_pden, pden.data(), pden.size() * sizeof(double), "upload pden"));
  RC(upload(ctx, ctx->d_mu, mu.data(), mu.size() * sizeof(int), "upload mu"));
  RC(upload(ctx, ctx->d_nu, nu.data(), nu.size() * sizeof(int), "upload nu"));
  RC(upload(ctx, ctx->d_mv, mv.data(), mv.size() * sizeof(int), "upload mvalue"));
  FB_HIP(hipStreamSynchronize(ctx->stream));
  ctx->h_grp3c.assign(grps.begin(), grps.begin() + nbcna);   // the networks serve bcna groups only
  ctx->h_grpkey = grpkey;
  ctx->mlp_ready = false;   // networks belong to the previous tables
  ctx->mlp_gen++;
  ctx->tables_ok = true;
  // the per-triplet group index of existing lists names the new tables' groups
  if (ctx->lists_ok && ctx->ntrip > 0) {
    DView v = ctx->view();
    RC(ctx->launch("trip_aux", [&] { return fbk_trip_aux(v, ctx->d_trip_aux.as<int>(), ctx->stream); }));
    FB_HIP(hipStreamSynchronize(ctx->stream));
  }
  return FB_OK;
}

int fb_clear_structures(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  ctx->structs.clear();
  ctx->structs_dirty = true;
  ctx->lists_ok = false;
  ctx->pp_lists_ok = false;
  return FB_OK;
}

int fb_add_structure(fb_ctx *ctx, int natoms, const double *ratom, const int *imass, const double *lattice,
                     int icluster, const double *dQ, int *istruct) {
  if (!ctx || !ctx->species_set) return ctx ? ctx->fail(FB_E_ORDER, "add_structure before set_species") : FB_E_ARG;
  if (natoms < 1 || !ratom || !imass || !lattice) return ctx->fail(FB_E_ARG, "add_structure: bad arguments");
  HStruct s;
  s.natoms = natoms;
  s.icluster = icluster;
  s.ratom.assign(ratom, ratom + 3 * (size_t)natoms);
  s.imass.assign(imass, imass + natoms);
  for (int i = 0; i < natoms; i++)
    if (imass[i] < 1 || imass[i] > ctx->sp.ns) return ctx->fail(FB_E_ARG, "add_structure: species out of range");
  std::memcpy(s.lattice, lattice, sizeof(s.lattice));
  s.dQ.assign((size_t)natoms * FB_MAXSH, 0.0);
  if (dQ) std::memcpy(s.dQ.data(), dQ, sizeof(double) * natoms * FB_MAXSH);
  ctx->structs.push_back(std::move(s));
  if (istruct) *istruct = (int)ctx->structs.size() - 1;
  ctx->structs_dirty = true;
  ctx->lists_ok = false;
  ctx->pp_lists_ok = false;
  return FB_OK;
}

int fb_set_positions(fb_ctx *ctx, int is, const double *ratom) {
  if (!ctx || is < 0 || is >= (int)ctx->structs.size() || !ratom) return FB_E_ARG;
  auto &s = ctx->structs[is];
  std::memcpy(s.ratom.data(), ratom, sizeof(double) * 3 * s.natoms);
  ctx->structs_dirty = true;
  ctx->lists_ok = false;
  ctx->pp_lists_ok = false;
  return FB_OK;
}

int fb_set_charges(fb_ctx *ctx, int is, const double *dQ) {
  if (!ctx || is < 0 || is >= (int)ctx->structs.size()) return FB_E_ARG;
  auto &s = ctx->structs[is];
  std::fill(s.dQ.begin(), s.dQ.end(), 0.0);
  if (dQ) std::memcpy(s.dQ.data(), dQ, sizeof(double) * s.natoms * FB_MAXSH);
  ctx->structs_dirty = true;  // charges change no list, but must be re-uploaded
  // only the Vna blocks depend on the shell charges: S, T and dipole_z stay valid,
  // as in a self-consistent loop that re-assembles Vna per charge update
  ctx->done[FB_MAT_VNA] = ctx->done[FB_MAT_VNAO] = 0;
  ctx->ddone[3] = ctx->ddone[4] = 0;   // the Vna force pieces too
  return FB_OK;
}

int fb_set_qin(fb_ctx *ctx, int is, const double *qin) {
  if (!ctx || is < 0 || is >= (int)ctx->structs.size() || !qin) return ctx ? ctx->fail(FB_E_ARG, "set_qin arguments") : FB_E_ARG;
  auto &s = ctx->structs[is];
  s.qin.assign(qin, qin + (size_t)s.natoms * FB_MAXSH);
  s.has_qin = true;
  ctx->structs_dirty = true;   // the next density assembler reads them, as the reference reads Qin per call
  return FB_OK;
}

static int upload_structures(fb_ctx *ctx) {
  const int ns = (int)ctx->structs.size();
  int natoms = 0;
  for (auto &s : ctx->structs) natoms += s.natoms;
  ctx->natoms = natoms;
  std::vector<double> ratom((size_t)natoms * 3), dq((size_t)natoms * FB_MAXSH), xl((size_t)ns * FB_NIMG * 3, 0.0);
  std::vector<double> qin((size_t)natoms * FB_MAXSH, 0.0);
  std::vector<int> imass(natoms), astr(natoms);
  ctx->h_st.assign(ns, DStruct{});
  int off = 0, max_atoms = 0;
  int64_t ncells = 0;
  for (int k = 0; k < ns; k++) {
    const HStruct &s = ctx->structs[k];
    DStruct &d = ctx->h_st[k];
    d.atom_off = off;
    d.natoms = s.natoms;
    d.icluster = s.icluster;
    d.mbeta_max = make_cells(s.lattice, s.icluster, &xl[(size_t)k * FB_NIMG * 3]);
    d.rcmax_all = -99.0;
    for (int q = 0; q < 3; q++) { d.bmin[q] = 1e300; d.bmax[q] = -1e300; }
    for (int i = 0; i < s.natoms; i++) {
      for (int q = 0; q < 3; q++) {
        double x = s.ratom[3 * i + q];
        ratom[3 * (size_t)(off + i) + q] = x;
        d.bmin[q] = std::min(d.bmin[q], x);
        d.bmax[q] = std::max(d.bmax[q], x);
      }
      imass[off + i] = s.imass[i];
      astr[off + i] = k;
      d.rcmax_all = std::max(d.rcmax_all, ctx->sp.rcmax[s.imass[i] - 1]);
      for (int q = 0; q < FB_MAXSH; q++) dq[(size_t)(off + i) * FB_MAXSH + q] = s.dQ[(size_t)i * FB_MAXSH + q];
      if (s.has_qin)
        for (int q = 0; q < FB_MAXSH; q++) qin[(size_t)(off + i) * FB_MAXSH + q] = s.qin[(size_t)i * FB_MAXSH + q];
    }
    // cell grid: width at least the largest cutoff of any list (H+S: rc_i + rc_j,
    // PP and PPx: rc + rcutoff_PP), at most 2^22 cells per structure
    double rcpp_max = 0.0;
    if (ctx->pp_species)
      for (int i = 0; i < ctx->sp.ns; i++) rcpp_max = std::max(rcpp_max, ctx->hpp.rcpp[i]);
    d.crange = 2.0 * std::max(d.rcmax_all, rcpp_max) + 1.0e-6;
    double ext[3];
    for (int q = 0; q < 3; q++) {
      ext[q] = s.natoms > 0 ? d.bmax[q] - d.bmin[q] : 0.0;
      d.clo[q] = s.natoms > 0 ? d.bmin[q] : 0.0;
      d.cn[q] = (int)std::min(1024.0, std::floor(ext[q] / d.crange) + 1.0);
    }
    while ((int64_t)d.cn[0] * d.cn[1] * d.cn[2] > (1ll << 22)) {
      const int q = (d.cn[0] >= d.cn[1] && d.cn[0] >= d.cn[2]) ? 0 : (d.cn[1] >= d.cn[2] ? 1 : 2);
      d.cn[q] = (d.cn[q] + 1) / 2;
    }
    // widths strictly above crange: the box point +- crange then spans at most
    // three cells per axis after rounding
    const double wmin = d.crange * (1.0 + 1.0e-9);
    for (int q = 0; q < 3; q++)
      d.cw[q] = d.cn[q] > 1 ? std::max(wmin, ext[q] / (d.cn[q] - 0.5)) : std::max(wmin, ext[q] + 1.0);
    d.cell_off = (int)ncells;
    ncells += (int64_t)d.cn[0] * d.cn[1] * d.cn[2];
    max_atoms = std::max(max_atoms, s.natoms);
    off += s.natoms;
  }
  ctx->ncells = ncells;
  // per-image bitmap of the neighbour search in LDS (32 KB at most), else the full scan
  ctx->cell_bits = (max_atoms <= 262144 && ncells < (1ll << 31)) ? (max_atoms + 31) / 32 : 0;
  ctx->h_imass = imass;
  RC(upload(ctx, ctx->d_ratom, ratom.data(), ratom.size() * sizeof(double), "upload ratom"));
  RC(upload(ctx, ctx->d_imass, imass.data(), imass.size() * sizeof(int), "upload imass"));
  RC(upload(ctx, ctx->d_atom_struct, astr.data(), astr.size() * sizeof(int), "upload atom_struct"));
  RC(upload(ctx, ctx->d_dQ, dq.data(), dq.size() * sizeof(double), "upload dQ"));
  RC(upload(ctx, ctx->d_qin, qin.data(), qin.size() * sizeof(double), "upload Qin"));
  RC(upload(ctx, ctx->d_xl, xl.data(), xl.size() * sizeof(double), "upload xl"));
  RC(upload(ctx, ctx->d_st, ctx->h_st.data(), ctx->h_st.size() * sizeof(DStruct), "upload structs"));
  ctx->structs_dirty = false;
  return FB_OK;
}

static int ensure_tmp(fb_ctx *ctx, size_t bytes) {
  hipError_t e = ctx->d_tmp.ensure(bytes);
  if (e != hipSuccess) return ctx->fail(FB_E_HIP, "temporary storage", e);
  return FB_OK;
}

static int scan_i32(fb_ctx *ctx, const int *in, int64_t *out, int64_t n) {
  size_t bytes = 0;
  int rc = fbk_scan_i32(in, out, n, nullptr, &bytes, ctx->stream);
  if (rc) return ctx->fail(FB_E_HIP, "scan sizing", (hipError_t)rc);
  RC(ensure_tmp(ctx, bytes));
  bytes = ctx->d_tmp.cap;
  return ctx->launch("scan", [&] { return fbk_scan_i32(in, out, n, ctx->d_tmp.p, &bytes, ctx->stream); });
}

static int scan_i64(fb_ctx *ctx, const int64_t *in, int64_t *out, int64_t n) {
  size_t bytes = 0;
  int rc = fbk_scan_i64(in, out, n, nullptr, &bytes, ctx->stream);
  if (rc) return ctx->fail(FB_E_HIP, "scan sizing", (hipError_t)rc);
  RC(ensure_tmp(ctx, bytes));
  bytes = ctx->d_tmp.cap;
  return ctx->launch("scan", [&] { return fbk_scan_i64(in, out, n, ctx->d_tmp.p, &bytes, ctx->stream); });
}

// Tiles of whole target segments, at most ch triplets each (a longer segment
// alone), packed per atom on the device: count, scan, fill.  Returns the
// number of tiles in *ntiles.
static int build_tiles(fb_ctx *ctx, const int64_t *pair_off, const int64_t *seg_off, const int *seg_pair, int64_t nseg,
                       int ch, DevBuf &tiles, int *ntiles) {
  const int N = ctx->natoms;
  hipStream_t st = ctx->stream;
  FB_HIP(ctx->d_tile_cnt.ensure(sizeof(int) * std::max(N, 1)));
  FB_HIP(ctx->d_tile_off.ensure(sizeof(int64_t) * (N + 1)));
  RC(ctx->launch("tiles_count", [&] {
    return fbk_tiles_atom(N, pair_off, seg_off, seg_pair, nseg, ch, 0, ctx->d_tile_cnt.as<int>(), nullptr, nullptr, st);
  }));
  RC(scan_i32(ctx, ctx->d_tile_cnt.as<int>(), ctx->d_tile_off.as<int64_t>(), N));
  int64_t n = 0;
  FB_HIP(hipMemcpyAsync(&n, ctx->d_tile_off.as<int64_t>() + N, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  FB_HIP(tiles.ensure(sizeof(int4) * std::max<int64_t>(n, 1)));
  RC(ctx->launch("tiles_fill", [&] {
    return fbk_tiles_atom(N, pair_off, seg_off, seg_pair, nseg, ch, 1, nullptr, ctx->d_tile_off.as<int64_t>(),
                          tiles.as<int4>(), st);
  }));
  *ntiles = (int)n;
  return FB_OK;
}

// out[k] = src[idx[k]] from a device array, one gather and one copy
static int gather_to_host(fb_ctx *ctx, const int64_t *src, const std::vector<int64_t> &idx, int64_t *out) {
  const int n = (int)idx.size();
  if (n == 0) return FB_OK;
  hipStream_t st = ctx->stream;
  FB_HIP(ctx->d_gidx.ensure(sizeof(int64_t) * n));
  FB_HIP(ctx->d_gout.ensure(sizeof(int64_t) * n));
  FB_HIP(hipMemcpyAsync(ctx->d_gidx.p, idx.data(), sizeof(int64_t) * n, hipMemcpyHostToDevice, st));
  RC(ctx->launch("gather", [&] {
    return fbk_gather_i64(n, src, ctx->d_gidx.as<int64_t>(), ctx->d_gout.as<int64_t>(), st);
  }));
  FB_HIP(hipMemcpyAsync(out, ctx->d_gout.p, sizeof(int64_t) * n, hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  return FB_OK;
}

int fb_build_lists(fb_ctx *ctx, int64_t *npairs_total, int64_t *ntrip_total) {
  if (!ctx) return FB_E_ARG;
  if (!ctx->tables_ok) return ctx->fail(FB_E_ORDER, "build_lists before commit_tables");
  if (ctx->structs.empty()) return ctx->fail(FB_E_ORDER, "build_lists without structures");
  FB_HIP(hipSetDevice(ctx->device));
  RC(upload_structures(ctx));
  const int N = ctx->natoms;
  RC(ctx->clear_status());
  hipStream_t st = ctx->stream;
  FB_HIP(ctx->d_neighn.ensure(sizeof(int) * N));
  FB_HIP(ctx->d_noff.ensure(sizeof(int64_t) * (N + 1)));
  FB_HIP(ctx->d_neigh_self.ensure(sizeof(int) * N));
  if (ctx->cell_bits > 0) {
    // atoms binned into their structure's cells (counting sort; order inside a
    // cell is free: the search orders its results through a bitmap)
    const int64_t C = ctx->ncells;
    FB_HIP(ctx->d_cell_cnt.ensure(sizeof(int) * C));
    FB_HIP(ctx->d_cell_start.ensure(sizeof(int64_t) * (C + 1)));
    FB_HIP(ctx->d_cell_atoms.ensure(sizeof(int) * std::max(N, 1)));
    FB_HIP(ctx->d_atom_cell.ensure(sizeof(int) * std::max(N, 1)));
    DView vb = ctx->view();
    RC(ctx->launch("cell_clear", [&] { return fbk_fill_i32(ctx->d_cell_cnt.as<int>(), C, 0, st); }));
    RC(ctx->launch("cell_count", [&] {
      return fbk_cell_count(vb, ctx->d_cell_cnt.as<int>(), ctx->d_atom_cell.as<int>(), st);
    }));
    RC(scan_i32(ctx, ctx->d_cell_cnt.as<int>(), ctx->d_cell_start.as<int64_t>(), C));
    RC(ctx->launch("cell_clear", [&] { return fbk_fill_i32(ctx->d_cell_cnt.as<int>(), C, 0, st); }));
    vb = ctx->view();
    RC(ctx->launch("cell_fill", [&] {
      return fbk_cell_fill(vb, ctx->d_atom_cell.as<int>(), ctx->d_cell_cnt.as<int>(), ctx->d_cell_atoms.as<int>(), st);
    }));
  }
  DView v = ctx->view();
  RC(ctx->launch("neighbors_count", [&] { return fbk_neighbors_count(v, ctx->d_neighn.as<int>(), v.status, st); }));
  RC(scan_i32(ctx, ctx->d_neighn.as<int>(), ctx->d_noff.as<int64_t>(), N));
  ctx->h_noff.resize(N + 1);
  FB_HIP(hipMemcpyAsync(ctx->h_noff.data(), ctx->d_noff.p, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost, st));
  RC(ctx->check_status("find_neighbors"));
  ctx->npairs = ctx->h_noff[N];
  const int64_t P = ctx->npairs;
  FB_HIP(ctx->d_neigh_j.ensure(sizeof(int) * P));
  FB_HIP(ctx->d_neigh_b.ensure(sizeof(int) * P));
  FB_HIP(ctx->d_neigh_back.ensure(sizeof(int) * P));
  FB_HIP(ctx->d_pair_atom.ensure(sizeof(int) * P));
  FB_HIP(ctx->d_pair_rec.ensure(sizeof(int4) * std::max<int64_t>(P, 1)));
  FB_HIP(ctx->d_pair_perm.ensure(std::max<int64_t>(P, 1)));
  FB_HIP(ctx->d_bsz.ensure(sizeof(int64_t) * P));
  FB_HIP(ctx->d_boff.ensure(sizeof(int64_t) * (P + 1)));
  v = ctx->view();
  RC(ctx->launch("fill_self", [&] { return fbk_fill_i32(ctx->d_neigh_self.as<int>(), N, -999, st); }));
  RC(ctx->launch("neighbors_fill", [&] {
    return fbk_neighbors_fill(v, ctx->d_noff.as<int64_t>(), ctx->d_neigh_j.as<int>(), ctx->d_neigh_b.as<int>(),
                              ctx->d_neigh_self.as<int>(), ctx->d_pair_atom.as<int>(), v.status, st);
  }));
  RC(ctx->launch("neigh_back", [&] { return fbk_neigh_back(v, ctx->d_neigh_back.as<int>(), v.status, st); }));
  FB_HIP(ctx->d_pair_disp.ensure(sizeof(double) * 3 * std::max<int64_t>(P, 1)));
  v = ctx->view();
  RC(ctx->launch("pair_disp", [&] { return fbk_pair_disp(v, ctx->d_pair_disp.as<double>(), st); }));
  {
    std::vector<int> self(N);
    FB_HIP(hipMemcpyAsync(self.data(), ctx->d_neigh_self.p, sizeof(int) * N, hipMemcpyDeviceToHost, st));
    RC(ctx->check_status("find_neighbors"));
    for (int i = 0; i < N; i++)
      if (self[i] == -999) return ctx->fail(FB_E_NEIGH_SELF, "find_neighbors: neigh_self missing");
  }
  // common neighbours
  FB_HIP(ctx->d_ncommon.ensure(sizeof(int) * N));
  FB_HIP(ctx->d_toff.ensure(sizeof(int64_t) * (N + 1)));
  RC(ctx->launch("common_count", [&] { return fbk_common_count(v, ctx->d_ncommon.as<int>(), v.status, st); }));
  RC(scan_i32(ctx, ctx->d_ncommon.as<int>(), ctx->d_toff.as<int64_t>(), N));
  ctx->h_toff.resize(N + 1);
  FB_HIP(hipMemcpyAsync(ctx->h_toff.data(), ctx->d_toff.p, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost, st));
  RC(ctx->check_status("find_common_neighbors"));
  ctx->ntrip = ctx->h_toff[N];
  const int64_t T = ctx->ntrip;
  if (P >= (1ll << 31) || T >= (1ll << 31)) return ctx->fail(FB_E_CAPACITY, "more than 2^31 pairs or triplets");
  FB_HIP(ctx->d_trip_orig.ensure(sizeof(int4) * T));
  FB_HIP(ctx->d_tpair_orig.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_trip.ensure(sizeof(int4) * T));
  FB_HIP(ctx->d_tpair.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_perm_in.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_perm.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_flag.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_segincl.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_seg_off.ensure(sizeof(int64_t) * (T + 1)));
  FB_HIP(ctx->d_seg_pair.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_trip_seg.ensure(sizeof(int) * T));
  FB_HIP(ctx->d_pair_tgt.ensure(sizeof(int) * std::max<int64_t>(P, 1)));
  RC(ctx->launch("clear_targets", [&] { return fbk_fill_i32(ctx->d_pair_tgt.as<int>(), P, 0, st); }));
  RC(ctx->launch("common_fill", [&] {
    return fbk_common_fill(v, ctx->d_toff.as<int64_t>(), ctx->d_trip_orig.as<int4>(), ctx->d_tpair_orig.as<int>(),
                           v.status, st);
  }));
  ctx->nseg = 0;
  for (int k = 0; k < 3; k++) ctx->ntiles[k] = 0;
  if (T > 0) {
    // stable radix sort of triplet ids by target pair
    int end_bit = 1;
    while ((1ll << end_bit) <= P) end_bit++;
    RC(ctx->launch("iota", [&] { return fbk_iota(ctx->d_perm_in.as<int>(), T, st); }));
    size_t bytes = 0;
    int rc = fbk_sort_pairs(ctx->d_tpair_orig.as<int>(), ctx->d_tpair.as<int>(), ctx->d_perm_in.as<int>(),
                            ctx->d_perm.as<int>(), T, end_bit, nullptr, &bytes, st);
    if (rc) return ctx->fail(FB_E_HIP, "sort sizing", (hipError_t)rc);
    size_t b2 = 0;
    rc = fbk_scan_incl_i32(ctx->d_flag.as<int>(), ctx->d_segincl.as<int>(), T, nullptr, &b2, st);
    if (rc) return ctx->fail(FB_E_HIP, "scan sizing", (hipError_t)rc);
    RC(ensure_tmp(ctx, std::max(bytes, b2)));
    bytes = ctx->d_tmp.cap;
    RC(ctx->launch("sort_triplets", [&] {
      return fbk_sort_pairs(ctx->d_tpair_orig.as<int>(), ctx->d_tpair.as<int>(), ctx->d_perm_in.as<int>(),
                            ctx->d_perm.as<int>(), T, end_bit, ctx->d_tmp.p, &bytes, st);
    }));
    RC(ctx->launch("seg_mark", [&] { return fbk_seg_mark(T, ctx->d_tpair.as<int>(), ctx->d_flag.as<int>(), st); }));
    b2 = ctx->d_tmp.cap;
    RC(ctx->launch("seg_scan", [&] {
      return fbk_scan_incl_i32(ctx->d_flag.as<int>(), ctx->d_segincl.as<int>(), T, ctx->d_tmp.p, &b2, st);
    }));
    int nseg = 0;
    FB_HIP(hipMemcpyAsync(&nseg, ctx->d_segincl.as<int>() + (T - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    RC(ctx->launch("seg_fill", [&] {
      return fbk_seg_fill(T, ctx->d_tpair.as<int>(), ctx->d_segincl.as<int>(), ctx->d_seg_off.as<int64_t>(),
                          ctx->d_seg_pair.as<int>(), ctx->d_trip_seg.as<int>(), ctx->d_pair_tgt.as<int>(), st);
    }));
    RC(ctx->launch("gather_triplets", [&] {
      return fbk_gather_trip(T, ctx->d_perm.as<int>(), ctx->d_trip_orig.as<int4>(), ctx->d_trip.as<int4>(), st);
    }));
    FB_HIP(hipStreamSynchronize(st));
    ctx->nseg = nseg;
    // per-triplet group and structure, per-tile segment ownership (list
    // properties: species and triplet order only)
    FB_HIP(ctx->d_trip_aux.ensure(sizeof(int) * T));
    v = ctx->view();
    v.ntrip = T;
    v.nseg = nseg;
    RC(ctx->launch("trip_aux", [&] { return fbk_trip_aux(v, ctx->d_trip_aux.as<int>(), st); }));
    // mixed-kernel tiles: whole segments, at most one chunk of triplets per
    // tile unless a single segment is longer, so that a workgroup makes one
    // pass over its triplets
    const int chs[3] = {64, 128, 256};
    for (int k = 0; k < 3; k++)
      RC(build_tiles(ctx, ctx->d_noff.as<int64_t>(), ctx->d_seg_off.as<int64_t>(), ctx->d_seg_pair.as<int>(), nseg,
                     chs[k], ctx->d_tiles[k], &ctx->ntiles[k]));
  }
  // block offsets
  v = ctx->view();
  RC(ctx->launch("block_sizes", [&] {
    return fbk_block_sizes(v, ctx->d_bsz.as<int64_t>(), ctx->d_pair_rec.as<int4>(), st);
  }));
  RC(scan_i64(ctx, ctx->d_bsz.as<int64_t>(), ctx->d_boff.as<int64_t>(), P));
  FB_HIP(hipMemcpyAsync(&ctx->blen, ctx->d_boff.as<int64_t>() + P, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  RC(ctx->check_status("find_common_neighbors"));
  ctx->h_boff_struct.resize(ctx->structs.size() + 1);
  {
    std::vector<int64_t> idx(ctx->structs.size());
    for (size_t k = 0; k < idx.size(); k++) idx[k] = ctx->h_noff[ctx->h_st[k].atom_off];
    RC(gather_to_host(ctx, ctx->d_boff.as<int64_t>(), idx, ctx->h_boff_struct.data()));
  }
  ctx->h_boff_struct[ctx->structs.size()] = ctx->blen;
  for (int i = 0; i < 5; i++) {
    FB_HIP(ctx->d_mat[i].ensure(sizeof(double) * std::max<int64_t>(ctx->blen, 1)));
    ctx->done[i] = 0;
  }
  // 2c workgroups: whole atoms, at most 16 atoms and FB_2C_PCAP pairs each (an
  // atom with more neighbours gets a workgroup of its own, run in passes)
  {
    std::vector<int> wg;
    wg.push_back(0);
    int na = 0;
    int64_t np = 0;
    for (int a = 0; a < N; a++) {
      int64_t n = ctx->h_noff[a + 1] - ctx->h_noff[a];
      if (na > 0 && (na == 16 || np + n > FB_2C_PCAP)) {
        wg.push_back(a);
        na = 0;
        np = 0;
      }
      na++;
      np += n;
    }
    wg.push_back(N);
    ctx->nwg2c = (int)wg.size() - 1;
    RC(upload(ctx, ctx->d_wg2c, wg.data(), wg.size() * sizeof(int), "upload wg2c"));
    v = ctx->view();
    RC(ctx->launch("pair_perm", [&] { return fbk_pair_perm(v, ctx->d_pair_perm.as<unsigned char>(), st); }));
    FB_HIP(hipStreamSynchronize(st));
  }
  ctx->lists_ok = true;
  ctx->lists_gen++;
  if (npairs_total) *npairs_total = P;
  if (ntrip_total) *ntrip_total = T;
  return FB_OK;
}

int fb_get_list_sizes(fb_ctx *ctx, int is, int64_t *npairs, int64_t *ntrip, int *mbeta_max) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size()) return ctx->fail(FB_E_ARG, "structure index");
  const DStruct &d = ctx->h_st[is];
  if (npairs) *npairs = ctx->h_noff[d.atom_off + d.natoms] - ctx->h_noff[d.atom_off];
  if (ntrip) *ntrip = ctx->h_toff[d.atom_off + d.natoms] - ctx->h_toff[d.atom_off];
  if (mbeta_max) *mbeta_max = d.mbeta_max;
  return FB_OK;
}

static int d2h(fb_ctx *ctx, void *dst, const void *src, size_t bytes) {
  if (!dst || bytes == 0) return FB_OK;
  hipError_t e = hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return ctx->fail(FB_E_HIP, "device to host copy", e);
  return FB_OK;
}

int fb_get_neighbors(fb_ctx *ctx, int is, int *neighn, int *neigh_j, int *neigh_b, int *neigh_back, int *neigh_self) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size()) return ctx->fail(FB_E_ARG, "structure index");
  FB_HIP(hipStreamSynchronize(ctx->stream));
  const DStruct &d = ctx->h_st[is];
  int64_t p0 = ctx->h_noff[d.atom_off], p1 = ctx->h_noff[d.atom_off + d.natoms];
  RC(d2h(ctx, neighn, ctx->d_neighn.as<int>() + d.atom_off, sizeof(int) * d.natoms));
  RC(d2h(ctx, neigh_j, ctx->d_neigh_j.as<int>() + p0, sizeof(int) * (p1 - p0)));
  RC(d2h(ctx, neigh_b, ctx->d_neigh_b.as<int>() + p0, sizeof(int) * (p1 - p0)));
  RC(d2h(ctx, neigh_back, ctx->d_neigh_back.as<int>() + p0, sizeof(int) * (p1 - p0)));
  RC(d2h(ctx, neigh_self, ctx->d_neigh_self.as<int>() + d.atom_off, sizeof(int) * d.natoms));
  return FB_OK;
}

int fb_get_common(fb_ctx *ctx, int is, int *ncommon, int *iatom_j, int *iatom_b, int *jatom_j, int *jatom_b,
                  int *neigh_common) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size()) return ctx->fail(FB_E_ARG, "structure index");
  FB_HIP(hipStreamSynchronize(ctx->stream));
  const DStruct &d = ctx->h_st[is];
  int64_t t0 = ctx->h_toff[d.atom_off], t1 = ctx->h_toff[d.atom_off + d.natoms];
  int64_t n = t1 - t0;
  RC(d2h(ctx, ncommon, ctx->d_ncommon.as<int>() + d.atom_off, sizeof(int) * d.natoms));
  std::vector<int4> tr(n);
  std::vector<int> tp(n);
  RC(d2h(ctx, tr.data(), ctx->d_trip_orig.as<int4>() + t0, sizeof(int4) * n));
  RC(d2h(ctx, tp.data(), ctx->d_tpair_orig.as<int>() + t0, sizeof(int) * n));
  for (int64_t t = 0; t < n; t++) {
    if (iatom_j) iatom_j[t] = tr[t].y - d.atom_off + 1;
    if (jatom_j) jatom_j[t] = tr[t].z - d.atom_off + 1;
    if (iatom_b) iatom_b[t] = tr[t].w & 0xffff;
    if (jatom_b) jatom_b[t] = (tr[t].w >> 16) & 0xffff;
    if (neigh_common) neigh_common[t] = (int)(tp[t] - ctx->h_noff[tr[t].y] + 1);
  }
  return FB_OK;
}

// ---------------------------------------------------------------- assembly
// The density, pseudopotential and force kernels hold blocks of s and p
// species (at most FB_MO orbitals); species with d shells (or more than four
// orbitals) are assembled by S, T, dipole_z, vna_2c and vna_3c only.
static int pre_assemble(fb_ctx *ctx, const char *what) {
  if (!ctx) return FB_E_ARG;
  if (!ctx->lists_ok) return ctx->fail(FB_E_ORDER, (std::string(what) + " before build_lists").c_str());
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("previous assemble_all"));
  if (ctx->structs_dirty) RC(upload_structures(ctx));
  RC(ctx->clear_status());
  return FB_OK;
}

static int launch_2c(fb_ctx *ctx, const DView &v, int mask, int variant, hipStream_t st) {
  return variant == 1 ? fbk_2c_sp(v, mask, st) : fbk_2c_lds(v, mask, st);
}

static int run_2c(fb_ctx *ctx, int mask, const char *name, hipStream_t st = nullptr) {
  if (!st) st = ctx->stream;
  DView v = ctx->view();
  if (ctx->canon_sp) return ctx->launch(name, [&] { return launch_2c(ctx, v, mask, ctx->k2c, st); });
  return ctx->launch(name, [&] { return fbk_2c(v, mask, st); });
}

static int any_charge(fb_ctx *ctx) {
  for (auto &s : ctx->structs)
    for (double q : s.dQ)
      if (q != 0.0) return 1;
  return 0;
}

static int run_3c(fb_ctx *ctx) {
  DView v = ctx->view();
  if (ctx->canon_sp) {
    // the learned surrogate serves the neutral bcna term of the mixed path;
    // charged builds keep the table stencil for every subtype
    if (ctx->precision == FB_PREC_MIXED && ctx->method3c == FB_3C_MLP && ctx->mlp_ready && !any_charge(ctx))
      return ctx->launch("assemble_3c_mlp", [&] {
        return fbk_3c_mlp(v, ctx->d_queue.as<int>(), 4 * ctx->ncu, ctx->stream);
      });
    if (ctx->precision == FB_PREC_MIXED)
      return ctx->launch("assemble_3c_mixed", [&] { return fbk_3c_mixed(v, ctx->chunk3c, any_charge(ctx), ctx->stream); });
    return ctx->launch("assemble_3c_fp64", [&] { return fbk_3c_fp64_ph(v, any_charge(ctx), ctx->stream); });
  }
  // generic shells: binary64 path only
  return ctx->launch("assemble_3c_fp64_generic", [&] { return fbk_3c_fp64(v, ctx->stream); });
}

int fb_assemble_S(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_assemble(ctx, "assemble_S"));
  RC(run_2c(ctx, 1, "assemble_2c_S"));
  RC(ctx->check_status("assemble_S"));
  ctx->done[FB_MAT_S] = 1;
  return FB_OK;
}

int fb_assemble_T(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_assemble(ctx, "assemble_T"));
  RC(run_2c(ctx, 2, "assemble_2c_T"));
  RC(ctx->check_status("assemble_T"));
  ctx->done[FB_MAT_T] = 1;
  return FB_OK;
}

int fb_assemble_dipole_z(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_assemble(ctx, "assemble_dipole_z"));
  RC(run_2c(ctx, 4, "assemble_2c_dipole_z"));
  RC(ctx->check_status("assemble_dipole_z"));
  ctx->done[FB_MAT_DIP] = 1;
  return FB_OK;
}

int fb_assemble_vna_2c(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_assemble(ctx, "assemble_vna_2c"));
  if (!ctx->done[FB_MAT_S]) return ctx->fail(FB_E_ORDER, "assemble_vna_2c needs assemble_S first");
  RC(run_2c(ctx, 8, "assemble_2c_vna"));
  RC(ctx->check_status("assemble_vna_2c"));
  ctx->done[FB_MAT_VNA] = ctx->done[FB_MAT_VNAO] = 1;
  return FB_OK;
}

int fb_assemble_vna_3c(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_assemble(ctx, "assemble_vna_3c"));
  if (!ctx->done[FB_MAT_VNA] || !ctx->done[FB_MAT_S] || !ctx->done[FB_MAT_DIP])
    return ctx->fail(FB_E_ORDER, "assemble_vna_3c needs assemble_S, assemble_dipole_z and assemble_vna_2c first");
  RC(run_3c(ctx));
  RC(ctx->check_status("assemble_vna_3c"));
  return FB_OK;
}

// The launches of one H+S build (status clear, 2c, 3c), enqueued on the
// stream.  Without per-kernel timing the 2c launch that the 3c launch does not
// read (all of 2c when neutral; the Vna part when charged, whose terms read
// S and dipole_z) runs on a second stream beside the 3c launch; it then leaves
// the Vna blocks of 3c targets to the 3c launch, which overwrites them.
// the one-launch H+S build serves neutral mixed-precision table builds of s/p
// species
static bool fused_build(const fb_ctx *ctx, int charged) {
  return ctx->fused && !charged && ctx->canon_sp && ctx->precision == FB_PREC_MIXED &&
         !(ctx->method3c == FB_3C_MLP && ctx->mlp_ready);
}

// The status word is not cleared here: a build only ORs error bits into it,
// and settle() reads and clears it (so back-to-back builds need no host
// round trip between them).
static int enqueue_build(fb_ctx *ctx, int charged) {
  if (fused_build(ctx, charged)) {
    DView v = ctx->view();
    RC(ctx->launch("assemble_hs_fused", [&] {
      return fbk_hs_fused(v, ctx->fused, ctx->persist ? ctx->d_queue.as<int>() : nullptr, ctx->ncu, ctx->stream);
    }));
    return FB_OK;
  }
  const bool fork = !ctx->timing && ctx->side != nullptr;
  if (charged) RC(run_2c(ctx, 7, "assemble_2c_STD"));
  if (fork) {
    FB_HIP(hipEventRecord(ctx->ev_fork, ctx->stream));
    FB_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_fork, 0));
    RC(run_2c(ctx, (charged ? 8 : 15) | FB_2C_SKIP_TGT, charged ? "assemble_2c_vna" : "assemble_2c", ctx->side));
    RC(run_3c(ctx));
    FB_HIP(hipEventRecord(ctx->ev_join, ctx->side));
    FB_HIP(hipStreamWaitEvent(ctx->stream, ctx->ev_join, 0));
  } else {
    RC(run_2c(ctx, charged ? 8 : 15, charged ? "assemble_2c_vna" : "assemble_2c"));
    RC(run_3c(ctx));
  }
  return FB_OK;
}

// One H+S build as fb_assemble_all runs it: replayed from the captured graph
// (captured again when its key changes), or enqueued directly with per-kernel
// timing, with graphs off or on the null stream.
static int launch_build(fb_ctx *ctx, int charged) {
  if (ctx->timing || !ctx->use_graph || ctx->stream == nullptr) return enqueue_build(ctx, charged);
  // the graph holds kernel arguments by value: any change of lists, precision,
  // charge state, chunk, 3c method, launch shape or networks needs a new capture
  const long key = ((((((((ctx->lists_gen * 4 + ctx->precision) * 2 + charged) * 512 + ctx->chunk3c) * 2 +
                         ctx->method3c) * 8 + ctx->fused) * 2 + ctx->persist) * 2 + ctx->k2c) * 4096) +
                   (ctx->mlp_gen & 4095);
  if (ctx->graph_key != key || !ctx->gexec) {
    if (ctx->gexec) { (void)hipGraphExecDestroy(ctx->gexec); ctx->gexec = nullptr; }
    if (ctx->graph) { (void)hipGraphDestroy(ctx->graph); ctx->graph = nullptr; }
    FB_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_build(ctx, charged);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc != FB_OK) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return ctx->fail(FB_E_HIP, "graph capture", e);
    ctx->graph = g;
    FB_HIP(hipGraphInstantiate(&ctx->gexec, g, nullptr, nullptr, 0));
    ctx->graph_key = key;
  }
  FB_HIP(hipGraphLaunch(ctx->gexec, ctx->stream));
  return FB_OK;
}

// Asynchronous: returns once the build is enqueued, without waiting for an
// earlier build.  A status raised by the kernels of the builds enqueued since
// the last check (x beyond a table, missing table) is returned by the next
// fb_synchronize, fb_get_blocks or other call that reads results.  Without
// per-kernel timing the build is replayed from a captured HIP graph.
int fb_assemble_all(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  if (!ctx->lists_ok) return ctx->fail(FB_E_ORDER, "assemble_all before build_lists");
  FB_HIP(hipSetDevice(ctx->device));
  if (ctx->structs_dirty) {
    RC(ctx->settle("previous assemble_all"));
    RC(upload_structures(ctx));
  }
  RC(launch_build(ctx, any_charge(ctx)));
  ctx->pending = true;
  for (int i = 0; i < 5; i++) ctx->done[i] = 1;
  return FB_OK;
}

// rocSOLVER is loaded on first use, so the library itself does not depend on it
namespace {
struct Solver {
  bool tried = false, ok = false;
  decltype(&rocblas_create_handle) create = nullptr;
  decltype(&rocblas_destroy_handle) destroy = nullptr;
  decltype(&rocblas_set_stream) set_stream = nullptr;
  decltype(&rocsolver_dsyevd) dsyevd = nullptr;
  std::string why;
};
Solver g_solver;

bool load_solver() {
  Solver &S = g_solver;
  if (S.tried) return S.ok;
  S.tried = true;
  void *hb = dlopen("librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
  if (!hb) hb = dlopen("/opt/rocm/lib/librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
  void *hs = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
  if (!hs) hs = dlopen("/opt/rocm/lib/librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
  if (!hb || !hs) {
    S.why = "rocBLAS/rocSOLVER not found";
    return false;
  }
  S.create = (decltype(S.create))dlsym(hb, "rocblas_create_handle");
  S.destroy = (decltype(S.destroy))dlsym(hb, "rocblas_destroy_handle");
  S.set_stream = (decltype(S.set_stream))dlsym(hb, "rocblas_set_stream");
  S.dsyevd = (decltype(S.dsyevd))dlsym(hs, "rocsolver_dsyevd");
  S.ok = S.create && S.destroy && S.set_stream && S.dsyevd;
  if (!S.ok) S.why = "rocBLAS/rocSOLVER symbols missing";
  return S.ok;
}
}  // namespace

int fb_gamma_eigenvalues(fb_ctx *ctx, int is, int which, double *w, int64_t nw, int64_t *n_out) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || (which != 0 && which != 1))
    return ctx->fail(FB_E_ARG, "gamma_eigenvalues arguments");
  if (which == 0 && !ctx->done[FB_MAT_S]) return ctx->fail(FB_E_ORDER, "gamma_eigenvalues: S not assembled");
  if (which == 1 && !(ctx->done[FB_MAT_T] && ctx->done[FB_MAT_VNA]))
    return ctx->fail(FB_E_ORDER, "gamma_eigenvalues: T and Vna not assembled");
  FB_HIP(hipSetDevice(ctx->device));
  FB_HIP(hipStreamSynchronize(ctx->stream));
  RC(ctx->settle("assemble_all"));
  const DStruct &st = ctx->h_st[is];
  std::vector<int> orb(st.natoms + 1, 0);
  for (int i = 0; i < st.natoms; i++) orb[i + 1] = orb[i] + ctx->sp.norb[ctx->h_imass[st.atom_off + i] - 1];
  const int n = orb[st.natoms];
  if (n_out) *n_out = n;
  if (!w || nw < n) return ctx->fail(FB_E_ARG, "gamma_eigenvalues: output shorter than the orbital count");
  if (!load_solver()) return ctx->fail(FB_E_UNSUPPORTED, g_solver.why.c_str());
  FB_HIP(hipStreamSynchronize(ctx->stream));  // buffers freed below must not be in use by queued work
  DevBuf dA, dorb, dD, dE, dinfo;
  struct Release {  // the dense matrix is transient: free every buffer on return
    DevBuf *b[5];
    ~Release() {
      for (DevBuf *x : b) x->release();
    }
  } release_all{{&dA, &dorb, &dD, &dE, &dinfo}};
  hipError_t e = dA.ensure(sizeof(double) * (size_t)n * n);
  if (e == hipSuccess) e = dorb.ensure(sizeof(int) * orb.size());
  if (e == hipSuccess) e = dD.ensure(sizeof(double) * n);
  if (e == hipSuccess) e = dE.ensure(sizeof(double) * n);
  if (e == hipSuccess) e = dinfo.ensure(sizeof(int));
  if (e != hipSuccess) return ctx->fail(FB_E_HIP, "gamma_eigenvalues buffers", e);
  FB_HIP(hipMemcpyAsync(dorb.p, orb.data(), sizeof(int) * orb.size(), hipMemcpyHostToDevice, ctx->stream));
  DView v = ctx->view();
  int rc = fbk_gamma_fold(v, st.atom_off, st.natoms, which, dorb.as<int>(), n, dA.as<double>(), ctx->stream);
  if (rc) return ctx->fail(FB_E_HIP, "gamma fold launch", (hipError_t)rc);
  rocblas_handle h = nullptr;
  if (g_solver.create(&h) != rocblas_status_success) return ctx->fail(FB_E_HIP, "rocblas_create_handle");
  rocblas_status rs = g_solver.set_stream(h, ctx->stream);
  if (rs == rocblas_status_success)
    rs = g_solver.dsyevd(h, rocblas_evect_none, rocblas_fill_upper, n, dA.as<double>(), n, dD.as<double>(),
                         dE.as<double>(), dinfo.as<int>());
  int info = 0;
  if (rs == rocblas_status_success) {
    e = hipMemcpyAsync(w, dD.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  g_solver.destroy(h);
  if (rs != rocblas_status_success) return ctx->fail(FB_E_HIP, "rocsolver_dsyevd");
  if (e != hipSuccess) return ctx->fail(FB_E_HIP, "gamma_eigenvalues copy", e);
  if (info != 0) return ctx->fail(FB_E_HIP, "rocsolver_dsyevd did not converge");
  return FB_OK;
}

// ------------------------------------------------------- learned 3c surrogate
namespace {

// work arrays of one forward + backward pass over a batch of B samples:
// Z1, A1, Z2, A2, dA, dZ2, dZ1 [B][64], Y, dY [B][64], rowloss [B], part
// [64][64][64] (split-K partial products), ones [1] = 1.0f
struct MlpWork {
  float *Z1, *A1, *Z2, *A2, *Y, *dY, *dA, *dZ2, *dZ1, *rowloss, *part;
  const float *ones;
};

struct MlpBufs {
  DevBuf X, T, osc, inv, Z1, A1, Z2, A2, Y, dY, dA, dZ2, dZ1, rowloss, P, G, M, V, part, ones;
  MlpWork work() {
    return MlpWork{Z1.as<float>(), A1.as<float>(), Z2.as<float>(), A2.as<float>(), Y.as<float>(), dY.as<float>(),
                   dA.as<float>(), dZ2.as<float>(), dZ1.as<float>(), rowloss.as<float>(), part.as<float>(),
                   ones.as<float>()};
  }
  ~MlpBufs() {
    DevBuf *b[] = {&X, &T, &osc, &inv, &Z1, &A1, &Z2, &A2, &Y, &dY, &dA, &dZ2, &dZ1, &rowloss, &P, &G, &M, &V, &part,
                   &ones};
    for (DevBuf *x : b) x->release();
  }
};

#define MLP_RC(call)                                                      \
  do {                                                                    \
    int rc_ = (call);                                                     \
    if (rc_) return ctx->fail(FB_E_HIP, #call, (hipError_t)rc_);          \
  } while (0)

int mlp_alloc_batch(fb_ctx *ctx, MlpBufs &b, int64_t B) {
  const size_t h = sizeof(float) * (size_t)B * MLP_H, o = sizeof(float) * (size_t)B * MLP_O;
  hipError_t e = hipSuccess;
  DevBuf *hb[] = {&b.Z1, &b.A1, &b.Z2, &b.A2, &b.dA, &b.dZ2, &b.dZ1};
  for (DevBuf *x : hb)
    if (e == hipSuccess) e = x->ensure(h);
  if (e == hipSuccess) e = b.Y.ensure(o);
  if (e == hipSuccess) e = b.dY.ensure(o);
  if (e == hipSuccess) e = b.rowloss.ensure(sizeof(float) * (size_t)B);
  if (e == hipSuccess) e = b.P.ensure(sizeof(float) * MLP_NP);
  if (e == hipSuccess) e = b.G.ensure(sizeof(float) * MLP_NP);
  if (e == hipSuccess) e = b.part.ensure(sizeof(float) * 64 * MLP_H * MLP_O);
  if (e == hipSuccess) e = b.ones.ensure(sizeof(float));
  if (e != hipSuccess) return ctx->fail(FB_E_HIP, "network buffers", e);
  const float one = 1.0f;
  FB_HIP(hipMemcpyAsync(b.ones.p, &one, sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  return FB_OK;
}

// forward + backward of one batch (X [B][3], T [B][64]) with parameters P;
// gradients into G, per-row losses into w.rowloss.  Every array of w is
// written before it is read, so the work area needs no clearing.
int mlp_fwd_bwd(fb_ctx *ctx, const MlpWork &w, const float *P, float *G, int nme, int64_t B, const float *X,
                const float *T) {
  hipStream_t st = ctx->stream;
  float *Z1 = w.Z1, *A1 = w.A1, *Z2 = w.Z2, *A2 = w.A2, *Y = w.Y, *dY = w.dY, *dA = w.dA;
  float *dZ2 = w.dZ2, *dZ1 = w.dZ1, *part = w.part;
  const float *ones = w.ones;
  const int Bi = (int)B;
  const int nz = (int)std::min<int64_t>(64, std::max<int64_t>(1, B / 256));
  int nvalid = 0;
  for (int o = 0; o < MLP_O; o++) {
    int th, m;
    nvalid += mlp_slot(o, nme, &th, &m) ? 1 : 0;
  }
  // forward
  MLP_RC(fbk_mlp_gemm(Bi, MLP_H, MLP_IN, X, MLP_IN, 1, P + MLP_P_W1, 1, MLP_IN, Z1, MLP_H, 1, 1, nullptr, st));
  MLP_RC(fbk_mlp_bias_act(B, Z1, A1, P + MLP_P_B1, st));
  MLP_RC(fbk_mlp_gemm(Bi, MLP_H, MLP_H, A1, MLP_H, 1, P + MLP_P_W2, 1, MLP_H, Z2, MLP_H, 1, 1, nullptr, st));
  MLP_RC(fbk_mlp_bias_act(B, Z2, A2, P + MLP_P_B2, st));
  MLP_RC(fbk_mlp_gemm(Bi, MLP_O, MLP_H, A2, MLP_H, 1, P + MLP_P_W3, 1, MLP_H, Y, MLP_O, 1, 1, nullptr, st));
  MLP_RC(fbk_mlp_out(B, nme, Y, P + MLP_P_B3, T, dY, w.rowloss, 1.0f / ((float)B * (float)nvalid), st));
  // backward
  MLP_RC(fbk_mlp_gemm(MLP_O, MLP_H, Bi, dY, 1, MLP_O, A2, MLP_H, 1, G + MLP_P_W3, MLP_H, 1, nz, part, st));
  MLP_RC(fbk_mlp_gemm(1, MLP_O, Bi, ones, 0, 0, dY, MLP_O, 1, G + MLP_P_B3, MLP_O, 1, nz, part, st));
  MLP_RC(fbk_mlp_gemm(Bi, MLP_H, MLP_O, dY, MLP_O, 1, P + MLP_P_W3, MLP_H, 1, dA, MLP_H, 1, 1, nullptr, st));
  MLP_RC(fbk_mlp_act_back(B, dA, Z2, dZ2, st));
  MLP_RC(fbk_mlp_gemm(MLP_H, MLP_H, Bi, dZ2, 1, MLP_H, A1, MLP_H, 1, G + MLP_P_W2, MLP_H, 1, nz, part, st));
  MLP_RC(fbk_mlp_gemm(1, MLP_H, Bi, ones, 0, 0, dZ2, MLP_H, 1, G + MLP_P_B2, MLP_H, 1, nz, part, st));
  MLP_RC(fbk_mlp_gemm(Bi, MLP_H, MLP_H, dZ2, MLP_H, 1, P + MLP_P_W2, MLP_H, 1, dA, MLP_H, 1, 1, nullptr, st));
  MLP_RC(fbk_mlp_act_back(B, dA, Z1, dZ1, st));
  MLP_RC(fbk_mlp_gemm(MLP_H, MLP_IN, Bi, dZ1, 1, MLP_H, X, MLP_IN, 1, G + MLP_P_W1, MLP_IN, 1, nz, part, st));
  MLP_RC(fbk_mlp_gemm(1, MLP_H, Bi, ones, 0, 0, dZ1, MLP_H, 1, G + MLP_P_B1, MLP_H, 1, nz, part, st));
  return FB_OK;
}

double mlp_batch_loss(fb_ctx *ctx, MlpBufs &b, int nme, int64_t B) {
  std::vector<float> rl((size_t)B);
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return -1.0;
  if (hipMemcpy(rl.data(), b.rowloss.p, sizeof(float) * B, hipMemcpyDeviceToHost) != hipSuccess) return -1.0;
  int nvalid = 0;
  for (int o = 0; o < MLP_O; o++) {
    int th, m;
    nvalid += mlp_slot(o, nme, &th, &m) ? 1 : 0;
  }
  double s = 0.0;
  for (float x : rl) s += x;
  return s / ((double)B * nvalid);
}

// weights and biases uniform in +-1/sqrt(fan in): the first layer's units
// then switch at points spread over the input square
void mlp_init_params(std::vector<float> &p, unsigned long long seed) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> u(-1.0, 1.0);
  p.assign(MLP_NP, 0.0f);
  const double l1 = 1.0 / std::sqrt((double)MLP_IN), l2 = 1.0 / std::sqrt((double)MLP_H);
  for (int i = 0; i < MLP_H * MLP_IN; i++) p[MLP_P_W1 + i] = (float)(l1 * u(rng));
  for (int i = 0; i < MLP_H; i++) p[MLP_P_B1 + i] = (float)(l1 * u(rng));
  for (int i = 0; i < MLP_H * MLP_H; i++) p[MLP_P_W2 + i] = (float)(l2 * u(rng));
  for (int i = 0; i < MLP_H; i++) p[MLP_P_B2 + i] = (float)(l2 * u(rng));
  for (int i = 0; i < MLP_O * MLP_H; i++) p[MLP_P_W3 + i] = (float)(l2 * u(rng));
  for (int i = 0; i < MLP_O; i++) p[MLP_P_B3 + i] = (float)(l2 * u(rng));
}

// device image of a network: MFMA operand fragments in the K-slot order of
// mlp_eval16 (fb_mlp.h), the slot scale folded into the last layer
void mlp_image(const std::vector<float> &p, const float *oscale, const DGroup3c &g, DMlp &w) {
  std::memset(&w, 0, sizeof(w));
  for (int o = 0; o < MLP_O; o++) w.oscale[o] = oscale ? oscale[o] : 1.0f;
  w.xmax = (float)g.xmax;
  w.ymax = (float)g.ymax;
  w.nme = g.nME;
  w.nx = g.nx;
  w.ny = g.ny;
  w.ready = 1;
  auto hi = [](float f) { return mlp_bf16(f); };
  auto lo = [](float f) { return mlp_bf16(f - mlp_bf16f(mlp_bf16(f))); };
  auto pack2 = [](const unsigned short h[4]) {
    return uint2{(unsigned)h[0] | ((unsigned)h[1] << 16), (unsigned)h[2] | ((unsigned)h[3] << 16)};
  };
  auto pack4 = [](const unsigned short h[8]) {
    return uint4{(unsigned)h[0] | ((unsigned)h[1] << 16), (unsigned)h[2] | ((unsigned)h[3] << 16),
                 (unsigned)h[4] | ((unsigned)h[5] << 16), (unsigned)h[6] | ((unsigned)h[7] << 16)};
  };
  // hidden unit of K slot (q, i) of K block kb: accumulator tiles 2kb, 2kb+1
  auto unit = [](int kb, int q, int i) { return i < 4 ? 32 * kb + 4 * q + i : 32 * kb + 16 + 4 * q + (i - 4); };
  for (int mt = 0; mt < 4; mt++)
    for (int l = 0; l < 64; l++) {
      const int q = l >> 4, r = 16 * mt + (l & 15);
      // layer 1: slots q0 (w0h, w0h, w1h, w1h), q1 (w2h, w2h, w0l, w1l), q2 (w2l, 0, 0, 0)
      const float *w1 = &p[MLP_P_W1 + r * MLP_IN];
      unsigned short s1[4] = {0, 0, 0, 0};
      if (q == 0) { s1[0] = hi(w1[0]); s1[1] = hi(w1[0]); s1[2] = hi(w1[1]); s1[3] = hi(w1[1]); }
      if (q == 1) { s1[0] = hi(w1[2]); s1[1] = hi(w1[2]); s1[2] = lo(w1[0]); s1[3] = lo(w1[1]); }
      if (q == 2) { s1[0] = lo(w1[2]); }
      w.w.a1[mt][l] = pack2(s1);
      const float sc = w.oscale[r];
      for (int kb = 0; kb < 2; kb++) {
        unsigned short s2[8], s3h[8], s3l[8];
        for (int i = 0; i < 8; i++) {
          const int c = unit(kb, q, i);
          s2[i] = hi(p[MLP_P_W2 + r * MLP_H + c]);
          const float v3 = p[MLP_P_W3 + r * MLP_H + c] * sc;
          s3h[i] = hi(v3);
          s3l[i] = lo(v3);
        }
        w.w.a2[mt][kb][l] = pack4(s2);
        w.w.a3h[mt][kb][l] = pack4(s3h);
        w.w.a3l[mt][kb][l] = pack4(s3l);
      }
    }
  for (int u = 0; u < MLP_H; u++) {
    w.w.b1[u] = p[MLP_P_B1 + u];
    w.w.b2[u] = p[MLP_P_B2 + u];
  }
  for (int o = 0; o < MLP_O; o++) w.w.b3[o] = p[MLP_P_B3 + o] * w.oscale[o];
}

}  // namespace

int fb_mlp_train(fb_ctx *ctx, int64_t nsamples, int nsteps, int batch, double lr, uint64_t seed) {
  if (!ctx || !ctx->tables_ok) return ctx ? ctx->fail(FB_E_ORDER, "mlp_train before commit_tables") : FB_E_ARG;
  if (nsamples < batch || batch < 256 || nsteps < 1 || !(lr > 0.0))
    return ctx->fail(FB_E_ARG, "mlp_train: need nsamples >= batch >= 256, nsteps >= 1, lr > 0");
  if (!ctx->canon_sp) return ctx->fail(FB_E_UNSUPPORTED, "mlp_train: s/p species with canonical tables only");
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("assemble_all"));
  const int ng = (int)ctx->h_grp3c.size();
  ctx->h_mlp.assign(ng, DMlp{});
  ctx->h_mlp_params.assign(ng, {});
  ctx->h_mlp_err.assign(ng, {});
  ctx->mlp_stat.assign(ng, std::array<double, 8>{});
  DView v = ctx->view();
  for (int gid = 0; gid < ng; gid++) {
    if (ctx->h_grpkey[gid][3] != 0) continue;   // the neutral term only
    const DGroup3c &g = ctx->h_grp3c[gid];
    auto t0 = std::chrono::steady_clock::now();
    MlpBufs b;
    hipError_t e = b.X.ensure(sizeof(float) * (size_t)nsamples * MLP_IN);
    if (e == hipSuccess) e = b.T.ensure(sizeof(float) * (size_t)nsamples * MLP_O);
    if (e == hipSuccess) e = b.osc.ensure(sizeof(float) * MLP_O);
    if (e == hipSuccess) e = b.inv.ensure(sizeof(float) * MLP_O);
    if (e == hipSuccess) e = b.M.ensure(sizeof(float) * MLP_NP);
    if (e == hipSuccess) e = b.V.ensure(sizeof(float) * MLP_NP);
    if (e != hipSuccess) return ctx->fail(FB_E_HIP, "network sample buffers", e);
    RC(mlp_alloc_batch(ctx, b, batch));
    MLP_RC(fbk_mlp_samples(v, gid, nsamples, seed + 0x1000ull * gid, b.X.as<float>(), b.T.as<float>(), ctx->stream));
    // envelope exp(c - kx x - ky y): least squares on log max_o |t_o| over a
    // subset of the samples (decay rates clamped at zero)
    double env[3] = {0.0, 0.0, 0.0};
    {
      const int64_t nf = std::min<int64_t>(nsamples, 1 << 18);
      DevBuf dm;
      FB_HIP(dm.ensure(sizeof(float) * (size_t)nf));
      std::vector<float> hx((size_t)nf * MLP_IN), hm((size_t)nf);
      int rc = fbk_mlp_rowmax(nf, b.T.as<float>(), dm.as<float>(), ctx->stream);
      if (rc == 0) rc = (int)hipStreamSynchronize(ctx->stream);
      if (rc == 0) rc = (int)hipMemcpy(hm.data(), dm.p, sizeof(float) * nf, hipMemcpyDeviceToHost);
      if (rc == 0) rc = (int)hipMemcpy(hx.data(), b.X.p, sizeof(float) * nf * MLP_IN, hipMemcpyDeviceToHost);
      dm.release();
      if (rc) return ctx->fail(FB_E_HIP, "envelope fit data", (hipError_t)rc);
      double A[3][3] = {{0}}, r[3] = {0};
      for (int64_t i = 0; i < nf; i++) {
        if (!(hm[i] > 1e-30f)) continue;
        const double f[3] = {1.0, -(hx[i * 3] + 1.0) * 0.5 * g.xmax, -(hx[i * 3 + 1] + 1.0) * 0.5 * g.ymax};
        const double lv = std::log((double)hm[i]);
        for (int a = 0; a < 3; a++) {
          r[a] += f[a] * lv;
          for (int c = 0; c < 3; c++) A[a][c] += f[a] * f[c];
        }
      }
      // 3x3 solve (Cramer)
      auto det3 = [](double M[3][3]) {
        return M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
               M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
      };
      const double D = det3(A);
      if (std::fabs(D) > 1e-300) {
        for (int k = 0; k < 3; k++) {
          double Mk[3][3];
          for (int a = 0; a < 3; a++)
            for (int c = 0; c < 3; c++) Mk[a][c] = (c == k) ? r[a] : A[a][c];
          env[k] = det3(Mk) / D;
        }
      }
      env[1] = std::max(env[1], 0.0);
      env[2] = std::max(env[2], 0.0);
    }
    MLP_RC(fbk_mlp_unenvelope(nsamples, b.X.as<float>(), b.T.as<float>(), (float)env[0], (float)env[1],
                              (float)env[2], (float)(0.5 * g.xmax), (float)(0.5 * g.ymax), ctx->stream));
    MLP_RC(fbk_mlp_colmaxabs(nsamples, MLP_O, b.T.as<float>(), b.osc.as<float>(), ctx->stream));
    std::vector<float> osc(MLP_O), inv(MLP_O);
    FB_HIP(hipStreamSynchronize(ctx->stream));
    FB_HIP(hipMemcpy(osc.data(), b.osc.p, sizeof(float) * MLP_O, hipMemcpyDeviceToHost));
    for (int o = 0; o < MLP_O; o++) {
      if (!(osc[o] > 1e-30f)) osc[o] = 1.0f;
      inv[o] = 1.0f / osc[o];
    }
    FB_HIP(hipMemcpyAsync(b.inv.p, inv.data(), sizeof(float) * MLP_O, hipMemcpyHostToDevice, ctx->stream));
    MLP_RC(fbk_mlp_scale_cols(nsamples, MLP_O, b.T.as<float>(), b.inv.as<float>(), ctx->stream));
    std::vector<float> p;
    mlp_init_params(p, seed * 31ull + (unsigned long long)gid);
    FB_HIP(hipMemcpyAsync(b.P.p, p.data(), sizeof(float) * MLP_NP, hipMemcpyHostToDevice, ctx->stream));
    FB_HIP(hipMemsetAsync(b.M.p, 0, sizeof(float) * MLP_NP, ctx->stream));
    FB_HIP(hipMemsetAsync(b.V.p, 0, sizeof(float) * MLP_NP, ctx->stream));
    const int64_t nb = nsamples / batch;
    const double b1 = 0.9, b2 = 0.999;
    double loss0 = -1.0, loss = -1.0;
    for (int step = 0; step < nsteps; step++) {
      const int64_t off = (step % nb) * batch;
      RC(mlp_fwd_bwd(ctx, b.work(), b.P.as<float>(), b.G.as<float>(), g.nME, batch, b.X.as<float>() + off * MLP_IN,
                     b.T.as<float>() + off * MLP_O));
      if (step == 0) loss0 = mlp_batch_loss(ctx, b, g.nME, batch);
      const double lr_t = lr * (0.01 + 0.99 * 0.5 * (1.0 + std::cos(M_PI * step / nsteps)));
      const double c1 = 1.0 / (1.0 - std::pow(b1, step + 1)), c2 = 1.0 / (1.0 - std::pow(b2, step + 1));
      MLP_RC(fbk_mlp_adam(MLP_NP, b.P.as<float>(), b.G.as<float>(), b.M.as<float>(), b.V.as<float>(), (float)lr_t,
                          (float)b1, (float)b2, (float)c1, (float)c2, 1e-8f, ctx->stream));
    }
    // loss of the trained network on the first batch
    RC(mlp_fwd_bwd(ctx, b.work(), b.P.as<float>(), b.G.as<float>(), g.nME, batch, b.X.as<float>(), b.T.as<float>()));
    loss = mlp_batch_loss(ctx, b, g.nME, batch);
    FB_HIP(hipMemcpy(p.data(), b.P.p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
    ctx->h_mlp_params[gid] = p;
    mlp_image(p, osc.data(), g, ctx->h_mlp[gid]);
    ctx->h_mlp[gid].ec = (float)env[0];
    ctx->h_mlp[gid].ekx = (float)env[1];
    ctx->h_mlp[gid].eky = (float)env[2];
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    ctx->mlp_stat[gid] = {1.0, (double)nsamples, (double)nsteps, loss0, loss, 0.0, ms, 0.0};
  }
  // calibration: per sub-cell of the calibration grid, the largest block
  // error |network - table| (Frobenius) of the sampled points
  FB_HIP(hipSetDevice(ctx->device));
  RC(upload(ctx, ctx->d_mlp, ctx->h_mlp.data(), sizeof(DMlp) * std::max(ng, 1), "upload networks"));
  std::vector<float> errs;
  for (int gid = 0; gid < ng; gid++) {
    if (!ctx->h_mlp[gid].ready) continue;
    const DGroup3c &g = ctx->h_grp3c[gid];
    const size_t nsub = (size_t)(g.nx - 1) * MLP_CSUB * (size_t)(g.ny - 1) * MLP_CSUB;
    DevBuf de;
    FB_HIP(de.ensure(sizeof(unsigned) * nsub));
    FB_HIP(hipMemsetAsync(de.p, 0, sizeof(unsigned) * nsub, ctx->stream));
    const int64_t ncal = std::max<int64_t>(1 << 20, (int64_t)nsub * 256);
    int rc = fbk_mlp_calib(v, gid, ctx->d_mlp.as<DMlp>() + gid, ncal, seed ^ 0xca1b7a7eull, de.as<unsigned>(),
                           ctx->stream);
    std::vector<float> ce(nsub);
    if (rc == 0) rc = (int)hipStreamSynchronize(ctx->stream);
    if (rc == 0) rc = (int)hipMemcpy(ce.data(), de.p, sizeof(float) * nsub, hipMemcpyDeviceToHost);
    de.release();
    if (rc) return ctx->fail(FB_E_HIP, "network calibration", (hipError_t)rc);
    // sub-cells no calibration sample reached never take the network
    float emax = 0.f;
    for (float &x : ce) {
      if (!(x > 0.f)) x = 3.0e38f;
      else emax = std::max(emax, x);
    }
    ctx->mlp_stat[gid][5] = emax;
    ctx->mlp_stat[gid][7] = (double)ncal;
    ctx->h_mlp[gid].err_off = (int64_t)errs.size();
    ctx->h_mlp_err[gid] = ce;
    errs.insert(errs.end(), ce.begin(), ce.end());
  }
  if (errs.empty()) errs.push_back(0.f);
  RC(upload(ctx, ctx->d_mlp, ctx->h_mlp.data(), sizeof(DMlp) * std::max(ng, 1), "upload networks"));
  RC(upload(ctx, ctx->d_mlp_err, errs.data(), sizeof(float) * errs.size(), "upload calibration"));
  FB_HIP(ctx->d_mlp_count.ensure(2 * sizeof(unsigned long long)));
  FB_HIP(ctx->d_queue.ensure(FB_QUEUE_INTS * sizeof(int)));

  FB_HIP(hipMemsetAsync(ctx->d_mlp_count.p, 0, 2 * sizeof(unsigned long long), ctx->stream));
  FB_HIP(hipStreamSynchronize(ctx->stream));
  ctx->mlp_ready = true;
  ctx->mlp_gen++;
  return FB_OK;
}

int fb_mlp_groups(fb_ctx *ctx, int *ngroups) {
  if (!ctx || !ngroups) return FB_E_ARG;
  *ngroups = (int)ctx->h_grp3c.size();
  return FB_OK;
}

int fb_mlp_samples(fb_ctx *ctx, int gid, int64_t n, uint64_t seed, float *X, float *T) {
  if (!ctx || !ctx->tables_ok) return ctx ? ctx->fail(FB_E_ORDER, "mlp_samples before commit_tables") : FB_E_ARG;
  if (gid < 0 || gid >= (int)ctx->h_grp3c.size() || n < 1 || !X || !T)
    return ctx->fail(FB_E_ARG, "mlp_samples arguments");
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("assemble_all"));
  DevBuf dx, dt;
  struct Release {
    DevBuf *b[2];
    ~Release() {
      for (DevBuf *x : b) x->release();
    }
  } release_all{{&dx, &dt}};
  FB_HIP(dx.ensure(sizeof(float) * (size_t)n * MLP_IN));
  FB_HIP(dt.ensure(sizeof(float) * (size_t)n * MLP_O));
  int rc = fbk_mlp_samples(ctx->view(), gid, n, seed, dx.as<float>(), dt.as<float>(), ctx->stream);
  if (rc) return ctx->fail(FB_E_HIP, "mlp_samples launch", (hipError_t)rc);
  FB_HIP(hipStreamSynchronize(ctx->stream));
  FB_HIP(hipMemcpy(X, dx.p, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyDeviceToHost));
  FB_HIP(hipMemcpy(T, dt.p, sizeof(float) * (size_t)n * MLP_O, hipMemcpyDeviceToHost));
  return FB_OK;
}

int fb_mlp_set_tolerance(fb_ctx *ctx, double tau) {
  if (!ctx || !(tau >= 0.0)) return FB_E_ARG;
  ctx->mlp_tau = tau;
  ctx->mlp_gen++;
  return FB_OK;
}

int fb_mlp_info(fb_ctx *ctx, int gid, double *out, int nout) {
  if (!ctx || !out || nout < 1) return FB_E_ARG;
  if (gid < 0 || gid >= (int)ctx->h_grp3c.size()) return ctx->fail(FB_E_ARG, "mlp_info: group index");
  double row[12] = {0};
  for (int k = 0; k < 4; k++) row[k] = ctx->h_grpkey[gid][k];
  if (gid < (int)ctx->mlp_stat.size())
    for (int k = 0; k < 8; k++) row[4 + k] = ctx->mlp_stat[gid][k];
  for (int k = 0; k < nout && k < 12; k++) out[k] = row[k];
  return FB_OK;
}

int fb_mlp_get_params(fb_ctx *ctx, int gid, float *params, float *oscale) {
  if (!ctx || gid < 0 || gid >= (int)ctx->h_mlp_params.size() || ctx->h_mlp_params[gid].empty())
    return ctx ? ctx->fail(FB_E_ARG, "mlp_get_params: no network for this group") : FB_E_ARG;
  if (params) std::memcpy(params, ctx->h_mlp_params[gid].data(), sizeof(float) * MLP_NP);
  if (oscale) std::memcpy(oscale, ctx->h_mlp[gid].oscale, sizeof(float) * MLP_O);
  return FB_OK;
}

int fb_mlp_stats(fb_ctx *ctx, int64_t *n_network, int64_t *n_seen) {
  if (!ctx) return FB_E_ARG;
  unsigned long long c[2] = {0, 0};
  if (ctx->d_mlp_count.p) {
    FB_HIP(hipStreamSynchronize(ctx->stream));
    FB_HIP(hipMemcpy(c, ctx->d_mlp_count.p, sizeof(c), hipMemcpyDeviceToHost));
    FB_HIP(hipMemset(ctx->d_mlp_count.p, 0, sizeof(c)));
  }
  if (n_network) *n_network = (int64_t)c[0];
  if (n_seen) *n_seen = (int64_t)c[1];
  return FB_OK;
}

int fb_mlp_gradients(fb_ctx *ctx, int nme, const float *params, int n, const float *X, const float *T,
                     float *grads, double *loss) {
  if (!ctx || !params || !X || !T || !grads || n < 1 || nme < 1 || nme > 12)
    return ctx ? ctx->fail(FB_E_ARG, "mlp_gradients arguments") : FB_E_ARG;
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("assemble_all"));
  MlpBufs b;
  RC(mlp_alloc_batch(ctx, b, n));
  FB_HIP(b.X.ensure(sizeof(float) * (size_t)n * MLP_IN));
  FB_HIP(b.T.ensure(sizeof(float) * (size_t)n * MLP_O));
  FB_HIP(hipMemcpyAsync(b.P.p, params, sizeof(float) * MLP_NP, hipMemcpyHostToDevice, ctx->stream));
  FB_HIP(hipMemcpyAsync(b.X.p, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice, ctx->stream));
  FB_HIP(hipMemcpyAsync(b.T.p, T, sizeof(float) * (size_t)n * MLP_O, hipMemcpyHostToDevice, ctx->stream));
  RC(mlp_fwd_bwd(ctx, b.work(), b.P.as<float>(), b.G.as<float>(), nme, n, b.X.as<float>(), b.T.as<float>()));
  const double l = mlp_batch_loss(ctx, b, nme, n);
  if (loss) *loss = l;
  FB_HIP(hipMemcpy(grads, b.G.p, sizeof(float) * MLP_NP, hipMemcpyDeviceToHost));
  return FB_OK;
}

// The forward and backward pass of fb_mlp_gradients on device arrays, with
// the work arrays in the scratch array of this file: no allocation, no host
// copy, no synchronization, so it can be captured in a HIP graph.
int fb_mlp_gradients_noalloc(fb_ctx *ctx, int nme, const float *params, int n, const float *X, const float *T,
                             float *grads, float *loss) {
  if (!ctx) return FB_E_ARG;
  if (!params || !X || !T || !grads || !loss || n < 1 || nme < 1 || nme > 12)
    return ctx->fail(FB_E_ARG, "mlp_gradients_noalloc arguments");
  if (n > FB_MLP_NOALLOC_MAX_BATCH)
    return ctx->fail(FB_E_CAPACITY, "mlp_gradients_noalloc: batch larger than FB_MLP_NOALLOC_MAX_BATCH (2048)");
  FB_HIP(hipSetDevice(ctx->device));
  float *const sc = ctx->scratch;
  const MlpWork w{sc, sc + NA_HID, sc + 2 * NA_HID, sc + 3 * NA_HID, sc + NA_OFF_Y, sc + NA_OFF_DY,
                  sc + 4 * NA_HID, sc + 5 * NA_HID, sc + 6 * NA_HID, sc + NA_OFF_ROW, sc + NA_OFF_PART,
                  sc + NA_OFF_ONES};
  // 1.0f, written by a kernel (no host copy, nothing a graph replays wrongly)
  int rc = fbk_fill_i32(reinterpret_cast<int *>(sc + NA_OFF_ONES), 1, 0x3f800000, ctx->stream);
  if (rc) return ctx->fail(FB_E_HIP, "mlp_gradients_noalloc fill", (hipError_t)rc);
  RC(mlp_fwd_bwd(ctx, w, params, grads, nme, n, X, T));
  int nvalid = 0;
  for (int o = 0; o < MLP_O; o++) {
    int th, m;
    nvalid += mlp_slot(o, nme, &th, &m) ? 1 : 0;
  }
  rc = fbk_mlp_loss(n, w.rowloss, 1.0f / ((float)n * (float)nvalid), loss, ctx->stream);
  if (rc) return ctx->fail(FB_E_HIP, "mlp_gradients_noalloc loss", (hipError_t)rc);
  return FB_OK;
}

int fb_mlp_forward(fb_ctx *ctx, int nme, const float *params, const float *oscale, int n, const float *X, float *Y) {
  if (!ctx || !params || !X || !Y || n < 1 || nme < 1 || nme > 12)
    return ctx ? ctx->fail(FB_E_ARG, "mlp_forward arguments") : FB_E_ARG;
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("assemble_all"));
  std::vector<float> p(params, params + MLP_NP);
  DGroup3c g{};
  g.nME = nme;
  g.xmax = g.ymax = 1.0;
  DMlp w;
  mlp_image(p, oscale, g, w);
  DevBuf dw, dx, dy;
  struct Release {
    DevBuf *b[3];
    ~Release() {
      for (DevBuf *x : b) x->release();
    }
  } release_all{{&dw, &dx, &dy}};
  FB_HIP(dw.ensure(sizeof(DMlp)));
  FB_HIP(dx.ensure(sizeof(float) * (size_t)n * MLP_IN));
  FB_HIP(dy.ensure(sizeof(float) * (size_t)n * MLP_O));
  FB_HIP(hipMemcpyAsync(dw.p, &w, sizeof(DMlp), hipMemcpyHostToDevice, ctx->stream));
  FB_HIP(hipMemcpyAsync(dx.p, X, sizeof(float) * (size_t)n * MLP_IN, hipMemcpyHostToDevice, ctx->stream));
  int rc = fbk_mlp_infer(dw.as<DMlp>(), n, dx.as<float>(), dy.as<float>(), ctx->stream);
  if (rc) return ctx->fail(FB_E_HIP, "mlp_forward launch", (hipError_t)rc);
  FB_HIP(hipStreamSynchronize(ctx->stream));
  FB_HIP(hipMemcpy(Y, dy.p, sizeof(float) * (size_t)n * MLP_O, hipMemcpyDeviceToHost));
  return FB_OK;
}

int fb_get_block_len(fb_ctx *ctx, int is, int64_t *len) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || !len) return ctx->fail(FB_E_ARG, "structure index");
  *len = ctx->h_boff_struct[is + 1] - ctx->h_boff_struct[is];
  return FB_OK;
}

int fb_get_blocks(fb_ctx *ctx, int is, int matrix, double *out) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || matrix < 0 || matrix >= FB_NMAT || !out)
    return ctx->fail(FB_E_ARG, "get_blocks arguments");
  if (matrix >= FB_MAT_RHO_IN) {
    // densities: valid once their two-centre assembler has run on these lists
    const bool ok = ctx->rho_gen == ctx->lists_gen && ctx->rdone[matrix >= FB_MAT_RHOW_IN ? 2 : 0];
    if (!ok) return ctx->fail(FB_E_ORDER, "get_blocks: density not assembled");
    FB_HIP(hipStreamSynchronize(ctx->stream));
    const std::vector<int64_t> &o = (matrix >= FB_MAT_RHOW_IN) ? ctx->h_wboff_struct : ctx->h_boff_struct;
    return d2h(ctx, out, ctx->d_mat[matrix].as<double>() + o[is], sizeof(double) * (o[is + 1] - o[is]));
  }
  if (!ctx->done[matrix]) return ctx->fail(FB_E_ORDER, "get_blocks: matrix not assembled");
  FB_HIP(hipStreamSynchronize(ctx->stream));
  RC(ctx->settle("assemble_all"));
  int64_t b0 = ctx->h_boff_struct[is], b1 = ctx->h_boff_struct[is + 1];
  return d2h(ctx, out, ctx->d_mat[matrix].as<double>() + b0, sizeof(double) * (b1 - b0));
}

int fb_get_device_blocks(fb_ctx *ctx, int matrix, void **dptr, int64_t *len) {
  if (!ctx || !ctx->lists_ok || matrix < 0 || matrix > 4) return FB_E_ARG;
  if (dptr) *dptr = ctx->d_mat[matrix].p;
  if (len) *len = ctx->blen;
  return FB_OK;
}

// ------------------------------------------------------------ McWEDA densities
extern "C" {
int fbk_rho_2c(const DView &v, int weighted, int maxsh, int sp, hipStream_t st);
int fbk_rho_3c(const DView &v, int weighted, int maxsh, int sp, hipStream_t st);
int fbk_shell_sizes(const DView &v, int64_t *sz, hipStream_t st);
}

// shell-block offsets and the four density arrays, once per list generation
static int prep_rho(fb_ctx *ctx) {
  if (ctx->rho_gen == ctx->lists_gen) return FB_OK;
  const int64_t P = ctx->npairs;
  hipStream_t st = ctx->stream;
  FB_HIP(ctx->d_wbsz.ensure(sizeof(int64_t) * std::max<int64_t>(P, 1)));
  FB_HIP(ctx->d_wboff.ensure(sizeof(int64_t) * (P + 1)));
  DView v = ctx->view();
  RC(ctx->launch("shell_sizes", [&] { return fbk_shell_sizes(v, ctx->d_wbsz.as<int64_t>(), st); }));
  RC(scan_i64(ctx, ctx->d_wbsz.as<int64_t>(), ctx->d_wboff.as<int64_t>(), P));
  FB_HIP(hipMemcpyAsync(&ctx->wlen, ctx->d_wboff.as<int64_t>() + P, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  ctx->h_wboff_struct.resize(ctx->structs.size() + 1);
  {
    std::vector<int64_t> idx(ctx->structs.size());
    for (size_t k = 0; k < idx.size(); k++) idx[k] = ctx->h_noff[ctx->h_st[k].atom_off];
    RC(gather_to_host(ctx, ctx->d_wboff.as<int64_t>(), idx, ctx->h_wboff_struct.data()));
  }
  ctx->h_wboff_struct[ctx->structs.size()] = ctx->wlen;
  for (int m = FB_MAT_RHO_IN; m <= FB_MAT_RHO_BOND; m++)
    FB_HIP(ctx->d_mat[m].ensure(sizeof(double) * std::max<int64_t>(ctx->blen, 1)));
  for (int m = FB_MAT_RHOW_IN; m <= FB_MAT_RHOW_BOND; m++)
    FB_HIP(ctx->d_mat[m].ensure(sizeof(double) * std::max<int64_t>(ctx->wlen, 1)));
  for (int i = 0; i < 4; i++) ctx->rdone[i] = 0;
  ctx->rho_gen = ctx->lists_gen;
  return FB_OK;
}

static int pre_rho(fb_ctx *ctx, const char *what) {
  RC(pre_assemble(ctx, what));
  for (auto &s : ctx->structs)
    if (!s.has_qin) return ctx->fail(FB_E_ORDER, (std::string(what) + ": shell charges not set (fb_set_qin)").c_str());
  return prep_rho(ctx);
}

static int rho_step(fb_ctx *ctx, int step) {
  static const char *names[4] = {"assemble_rho_2c", "assemble_rho_3c", "assemble_rho_weighted_2c",
                                 "assemble_rho_weighted_3c"};
  // rho_3c adds into the blocks of rho_2c; rho_weighted_3c into those of rho_weighted_2c
  if ((step == 1 || step == 3) && !ctx->rdone[step - 1])
    return ctx->fail(FB_E_ORDER, (std::string(names[step]) + " before " + names[step - 1]).c_str());
  const int weighted = step >= 2;
  DView v = ctx->view();
  const int sp = ctx->canon_rho && ctx->force_kernel;   // FB_TUNE_FORCE_KERNEL 0 selects the generic kernels
  if (step == 0 || step == 2)
    RC(ctx->launch(names[step], [&] { return fbk_rho_2c(v, weighted, ctx->maxsh, sp, ctx->stream); }));
  else
    RC(ctx->launch(names[step], [&] { return fbk_rho_3c(v, weighted, ctx->maxsh, sp, ctx->stream); }));
  ctx->rdone[step] = 1;
  if (step == 0 || step == 2) ctx->rdone[step + 1] = 0;
  return FB_OK;
}

static int rho_entry(fb_ctx *ctx, int step) {
  if (!ctx) return FB_E_ARG;
  static const char *names[4] = {"assemble_rho_2c", "assemble_rho_3c", "assemble_rho_weighted_2c",
                                 "assemble_rho_weighted_3c"};
  RC(pre_rho(ctx, names[step]));
  RC(rho_step(ctx, step));
  return ctx->check_status(names[step]);
}

int fb_assemble_rho_2c(fb_ctx *ctx) { return rho_entry(ctx, 0); }
int fb_assemble_rho_3c(fb_ctx *ctx) { return rho_entry(ctx, 1); }
int fb_assemble_rho_weighted_2c(fb_ctx *ctx) { return rho_entry(ctx, 2); }
int fb_assemble_rho_weighted_3c(fb_ctx *ctx) { return rho_entry(ctx, 3); }

int fb_assemble_rho_all(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_rho(ctx, "assemble_rho_all"));
  for (int step = 0; step < 4; step++) RC(rho_step(ctx, step));
  return ctx->check_status("assemble_rho_all");
}

int fb_get_shell_block_len(fb_ctx *ctx, int is, int64_t *len) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || !len) return ctx->fail(FB_E_ARG, "structure index");
  if (ctx->rho_gen != ctx->lists_gen) {
    FB_HIP(hipSetDevice(ctx->device));
    RC(ctx->settle("previous assemble_all"));
    RC(prep_rho(ctx));
  }
  *len = ctx->h_wboff_struct[is + 1] - ctx->h_wboff_struct[is];
  return FB_OK;
}

// ---------------------------------------------------------- pseudopotential
extern "C" int fbk_fill_i32(int *a, int64_t n, int val, hipStream_t st);

int fb_set_species_pp(fb_ctx *ctx, const int *nssh_pp, const int *lssh_pp, const double *rcutoff_pp,
                      const double *cl) {
  if (!ctx || !nssh_pp || !lssh_pp || !rcutoff_pp || !cl) return ctx ? ctx->fail(FB_E_ARG, "set_species_pp") : FB_E_ARG;
  if (!ctx->species_set) return ctx->fail(FB_E_ORDER, "set_species_pp before set_species");
  DPP h{};
  for (int i = 0; i < ctx->sp.ns; i++) {
    if (nssh_pp[i] < 0 || nssh_pp[i] > FB_MAXSH) return ctx->fail(FB_E_ARG, "set_species_pp: nssh_PP out of range");
    h.nssh_pp[i] = nssh_pp[i];
    for (int k = 0; k < nssh_pp[i]; k++) {
      const int l = lssh_pp[i * FB_MAXSH + k];
      if (l < 0 || l > 1) return ctx->fail(FB_E_UNSUPPORTED, "set_species_pp: PP shells beyond s and p");
      h.lssh_pp[i][k] = l;
      h.cl[i][k] = cl[i * FB_MAXSH + k];
      h.norb_pp[i] += 2 * l + 1;
    }
    if (h.norb_pp[i] > 4) return ctx->fail(FB_E_UNSUPPORTED, "set_species_pp: more than 4 projector orbitals");
    h.rcpp[i] = rcutoff_pp[i];
    int c = 0;   // cl per projector orbital, b.FUNCTIONS/cl.f90:84-93
    for (int k = 0; k < nssh_pp[i]; k++)
      for (int m = 0; m < 2 * h.lssh_pp[i][k] + 1; m++) h.clo[i][c++] = h.cl[i][k];
  }
  ctx->hpp = h;
  ctx->pp_species = true;
  ctx->pp_lists_ok = false;
  ctx->tables_ok = false;   // the projector tables are checked against the PP shells at commit
  return FB_OK;
}

static int pp_list_pass(fb_ctx *ctx, int list) {
  const int N = ctx->natoms;
  hipStream_t st = ctx->stream;
  FB_HIP(ctx->d_ppn[list].ensure(sizeof(int) * std::max(N, 1)));
  FB_HIP(ctx->d_ppoff[list].ensure(sizeof(int64_t) * (N + 1)));
  FB_HIP(ctx->d_ppself[list].ensure(sizeof(int) * std::max(N, 1)));
  DView v = ctx->view();
  DPP q = ctx->pview();
  RC(ctx->launch("pp_neighbors_count", [&] { return fbk_pp_neighbors(v, q, list, 0, v.status, st); }));
  RC(scan_i32(ctx, ctx->d_ppn[list].as<int>(), ctx->d_ppoff[list].as<int64_t>(), N));
  ctx->h_ppoff[list].resize(N + 1);
  FB_HIP(hipMemcpyAsync(ctx->h_ppoff[list].data(), ctx->d_ppoff[list].p, sizeof(int64_t) * (N + 1),
                        hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  ctx->pp_npairs[list] = ctx->h_ppoff[list][N];
  const int64_t P = std::max<int64_t>(ctx->pp_npairs[list], 1);
  FB_HIP(ctx->d_ppj[list].ensure(sizeof(int) * P));
  FB_HIP(ctx->d_ppb[list].ensure(sizeof(int) * P));
  FB_HIP(ctx->d_ppatom[list].ensure(sizeof(int) * P));
  RC(ctx->launch("fill_self", [&] { return fbk_fill_i32(ctx->d_ppself[list].as<int>(), N, -999, st); }));
  q = ctx->pview();
  RC(ctx->launch("pp_neighbors_fill", [&] { return fbk_pp_neighbors(v, q, list, 1, v.status, st); }));
  std::vector<int> self(N);
  FB_HIP(hipMemcpyAsync(self.data(), ctx->d_ppself[list].p, sizeof(int) * N, hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  for (int i = 0; i < N; i++)
    if (self[i] == -999) return ctx->fail(FB_E_NEIGH_SELF, "driver_neighbors_PP: self entry missing");
  return FB_OK;
}

int fb_build_lists_pp(fb_ctx *ctx, int64_t *npp, int64_t *nppp, int64_t *ntrip) {
  if (!ctx) return FB_E_ARG;
  if (!ctx->tables_ok) return ctx->fail(FB_E_ORDER, "build_lists_pp before commit_tables");
  if (!ctx->pp_species) return ctx->fail(FB_E_ORDER, "build_lists_pp without PP shells (fb_set_species_pp)");
  if (ctx->structs.empty()) return ctx->fail(FB_E_ORDER, "build_lists_pp without structures");
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("previous assemble_all"));
  RC(upload_structures(ctx));
  RC(ctx->clear_status());
  const int N = ctx->natoms;
  hipStream_t st = ctx->stream;
  RC(pp_list_pass(ctx, PPL_PP));
  RC(pp_list_pass(ctx, PPL_PPX));
  // neighbors_PPp: candidate ranges, first-seen dedup, copy
  FB_HIP(ctx->d_cand_cnt.ensure(sizeof(int64_t) * std::max(N, 1)));
  FB_HIP(ctx->d_cand_off.ensure(sizeof(int64_t) * (N + 1)));
  DView v = ctx->view();
  DPP q = ctx->pview();
  RC(ctx->launch("ppp_candidates", [&] { return fbk_ppp_candidates(v, q, ctx->d_cand_cnt.as<int64_t>(), st); }));
  RC(scan_i64(ctx, ctx->d_cand_cnt.as<int64_t>(), ctx->d_cand_off.as<int64_t>(), N));
  int64_t ncand = 0;
  FB_HIP(hipMemcpyAsync(&ncand, ctx->d_cand_off.as<int64_t>() + N, sizeof(int64_t), hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  FB_HIP(ctx->d_cand_j.ensure(sizeof(int) * std::max<int64_t>(ncand, 1)));
  FB_HIP(ctx->d_cand_b.ensure(sizeof(int) * std::max<int64_t>(ncand, 1)));
  FB_HIP(ctx->d_ppn[PPL_PPP].ensure(sizeof(int) * std::max(N, 1)));
  FB_HIP(ctx->d_ppoff[PPL_PPP].ensure(sizeof(int64_t) * (N + 1)));
  FB_HIP(ctx->d_ppself[PPL_PPP].ensure(sizeof(int) * std::max(N, 1)));
  q = ctx->pview();
  RC(ctx->launch("ppp_dedup", [&] { return fbk_ppp_build(v, q, 0, v.status, st); }));
  RC(scan_i32(ctx, ctx->d_ppn[PPL_PPP].as<int>(), ctx->d_ppoff[PPL_PPP].as<int64_t>(), N));
  ctx->h_ppoff[PPL_PPP].resize(N + 1);
  FB_HIP(hipMemcpyAsync(ctx->h_ppoff[PPL_PPP].data(), ctx->d_ppoff[PPL_PPP].p, sizeof(int64_t) * (N + 1),
                        hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  ctx->pp_npairs[PPL_PPP] = ctx->h_ppoff[PPL_PPP][N];
  const int64_t Q = std::max<int64_t>(ctx->pp_npairs[PPL_PPP], 1);
  FB_HIP(ctx->d_ppj[PPL_PPP].ensure(sizeof(int) * Q));
  FB_HIP(ctx->d_ppb[PPL_PPP].ensure(sizeof(int) * Q));
  FB_HIP(ctx->d_ppatom[PPL_PPP].ensure(sizeof(int) * Q));
  q = ctx->pview();
  RC(ctx->launch("ppp_fill", [&] { return fbk_ppp_build(v, q, 1, v.status, st); }));
  RC(ctx->check_status("driver_neighbors_PP (PPp)"));
  FB_HIP(ctx->d_ppp_disp.ensure(sizeof(double) * 3 * Q));
  q = ctx->pview();
  RC(ctx->launch("ppp_disp", [&] { return fbk_ppp_disp(v, q, st); }));
  // per-atom key tables of the PPp lists (the common-neighbour match)
  {
    FB_HIP(ctx->d_pk_size.ensure(sizeof(int) * std::max(N, 1)));
    FB_HIP(ctx->d_pk_off.ensure(sizeof(int64_t) * (N + 1)));
    RC(ctx->launch("ppk_size", [&] { return fbk_ppk_size(v, q, ctx->d_pk_size.as<int>(), st); }));
    RC(scan_i32(ctx, ctx->d_pk_size.as<int>(), ctx->d_pk_off.as<int64_t>(), N));
    int64_t nslot = 0;
    FB_HIP(hipMemcpyAsync(&nslot, ctx->d_pk_off.as<int64_t>() + N, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    FB_HIP(hipStreamSynchronize(st));
    FB_HIP(ctx->d_pk_key.ensure(sizeof(unsigned) * std::max<int64_t>(nslot, 1)));
    FB_HIP(ctx->d_pk_idx.ensure(sizeof(int) * std::max<int64_t>(nslot, 1)));
    RC(ctx->launch("ppk_clear", [&] { return fbk_fill_i32(ctx->d_pk_key.as<int>(), nslot, -1, st); }));
    q = ctx->pview();
    RC(ctx->launch("ppk_insert", [&] { return fbk_ppk_insert(v, q, st); }));
  }
  // maps, inverse maps and points
  FB_HIP(ctx->d_pp_map.ensure(sizeof(int) * std::max<int64_t>(ctx->pp_npairs[PPL_PP], 1)));
  FB_HIP(ctx->d_px_map.ensure(sizeof(int) * std::max<int64_t>(ctx->pp_npairs[PPL_PPX], 1)));
  FB_HIP(ctx->d_px_point.ensure(sizeof(int) * std::max<int64_t>(ctx->pp_npairs[PPL_PPX], 1)));
  FB_HIP(ctx->d_inv_pp.ensure(sizeof(int) * Q));
  FB_HIP(ctx->d_inv_px.ensure(sizeof(int) * Q));
  RC(ctx->launch("fill_inv", [&] { return fbk_fill_i32(ctx->d_inv_pp.as<int>(), Q, -1, st); }));
  RC(ctx->launch("fill_inv", [&] { return fbk_fill_i32(ctx->d_inv_px.as<int>(), Q, -1, st); }));
  q = ctx->pview();
  RC(ctx->launch("pp_maps", [&] { return fbk_pp_maps(v, q, v.status, st); }));
  RC(ctx->check_status("driver_neighbors_PP (maps)"));
  // PP common neighbours
  FB_HIP(ctx->d_pc_n.ensure(sizeof(int) * std::max(N, 1)));
  FB_HIP(ctx->d_pc_off.ensure(sizeof(int64_t) * (N + 1)));
  q = ctx->pview();
  RC(ctx->launch("pp_common_count", [&] { return fbk_pp_common(v, q, 0, st); }));
  RC(scan_i32(ctx, ctx->d_pc_n.as<int>(), ctx->d_pc_off.as<int64_t>(), N));
  ctx->h_pcoff.resize(N + 1);
  FB_HIP(hipMemcpyAsync(ctx->h_pcoff.data(), ctx->d_pc_off.p, sizeof(int64_t) * (N + 1), hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  const int64_t T = ctx->h_pcoff[N];
  if (T >= (1ll << 31) || Q >= (1ll << 31)) return ctx->fail(FB_E_CAPACITY, "more than 2^31 PP triplets or pairs");
  ctx->pp_ntrip = T;
  const int64_t T1 = std::max<int64_t>(T, 1);
  FB_HIP(ctx->d_pc_trip.ensure(sizeof(int4) * T1));
  FB_HIP(ctx->d_pc_tpair.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_trip_s.ensure(sizeof(int4) * T1));
  FB_HIP(ctx->d_pc_tpair_s.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_perm_in.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_perm.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_flag.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_segincl.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_seg_off.ensure(sizeof(int64_t) * (T1 + 1)));
  FB_HIP(ctx->d_pc_seg_pair.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_trip_seg.ensure(sizeof(int) * T1));
  FB_HIP(ctx->d_pc_tgt.ensure(sizeof(int) * Q));
  FB_HIP(ctx->d_pc_mp.ensure(sizeof(int2) * T1));
  q = ctx->pview();
  RC(ctx->launch("pp_common_fill", [&] { return fbk_pp_common(v, q, 1, st); }));
  ctx->pp_nseg = 0;
  ctx->pp_ntiles = 0;
  if (T > 0) {
    // stable sort of the triplets by target PPp pair: segments in list order
    int end_bit = 1;
    while ((1ll << end_bit) <= Q) end_bit++;
    RC(ctx->launch("iota", [&] { return fbk_iota(ctx->d_pc_perm_in.as<int>(), T, st); }));
    size_t bytes = 0, b2 = 0;
    int rc = fbk_sort_pairs(ctx->d_pc_tpair.as<int>(), ctx->d_pc_tpair_s.as<int>(), ctx->d_pc_perm_in.as<int>(),
                            ctx->d_pc_perm.as<int>(), T, end_bit, nullptr, &bytes, st);
    if (rc) return ctx->fail(FB_E_HIP, "sort sizing", (hipError_t)rc);
    rc = fbk_scan_incl_i32(ctx->d_pc_flag.as<int>(), ctx->d_pc_segincl.as<int>(), T, nullptr, &b2, st);
    if (rc) return ctx->fail(FB_E_HIP, "scan sizing", (hipError_t)rc);
    RC(ensure_tmp(ctx, std::max(bytes, b2)));
    bytes = ctx->d_tmp.cap;
    RC(ctx->launch("pp_sort_triplets", [&] {
      return fbk_sort_pairs(ctx->d_pc_tpair.as<int>(), ctx->d_pc_tpair_s.as<int>(), ctx->d_pc_perm_in.as<int>(),
                            ctx->d_pc_perm.as<int>(), T, end_bit, ctx->d_tmp.p, &bytes, st);
    }));
    RC(ctx->launch("seg_mark", [&] { return fbk_seg_mark(T, ctx->d_pc_tpair_s.as<int>(), ctx->d_pc_flag.as<int>(), st); }));
    b2 = ctx->d_tmp.cap;
    RC(ctx->launch("seg_scan", [&] {
      return fbk_scan_incl_i32(ctx->d_pc_flag.as<int>(), ctx->d_pc_segincl.as<int>(), T, ctx->d_tmp.p, &b2, st);
    }));
    int nseg = 0;
    FB_HIP(hipMemcpyAsync(&nseg, ctx->d_pc_segincl.as<int>() + (T - 1), sizeof(int), hipMemcpyDeviceToHost, st));
    RC(ctx->launch("seg_fill", [&] {
      return fbk_seg_fill(T, ctx->d_pc_tpair_s.as<int>(), ctx->d_pc_segincl.as<int>(), ctx->d_pc_seg_off.as<int64_t>(),
                          ctx->d_pc_seg_pair.as<int>(), ctx->d_pc_trip_seg.as<int>(), ctx->d_pc_tgt.as<int>(), st);
    }));
    RC(ctx->launch("gather_triplets", [&] {
      return fbk_gather_trip(T, ctx->d_pc_perm.as<int>(), ctx->d_pc_trip.as<int4>(), ctx->d_pc_trip_s.as<int4>(), st);
    }));
    FB_HIP(hipStreamSynchronize(st));
    ctx->pp_nseg = nseg;
    // one-pass tiles of the s/p vnl_3c kernel (whole segments, at most 64 triplets)
    RC(build_tiles(ctx, ctx->d_ppoff[PPL_PPP].as<int64_t>(), ctx->d_pc_seg_off.as<int64_t>(),
                   ctx->d_pc_seg_pair.as<int>(), nseg, 64, ctx->d_pc_tiles, &ctx->pp_ntiles));
    q = ctx->pview();
    RC(ctx->launch("pp_mpair", [&] { return fbk_pp_mpair(v, q, st); }));
  }
  // block offsets: sVNL per PP pair, vnl per PPp pair
  const int64_t P = std::max<int64_t>(ctx->pp_npairs[PPL_PP], 1);
  FB_HIP(ctx->d_sv_bsz.ensure(sizeof(int64_t) * P));
  FB_HIP(ctx->d_sv_off.ensure(sizeof(int64_t) * (P + 1)));
  FB_HIP(ctx->d_vl_bsz.ensure(sizeof(int64_t) * Q));
  FB_HIP(ctx->d_vl_off.ensure(sizeof(int64_t) * (Q + 1)));
  q = ctx->pview();
  RC(ctx->launch("pp_block_sizes", [&] { return fbk_pp_block_sizes(v, q, 0, ctx->d_sv_bsz.as<int64_t>(), st); }));
  RC(scan_i64(ctx, ctx->d_sv_bsz.as<int64_t>(), ctx->d_sv_off.as<int64_t>(), ctx->pp_npairs[PPL_PP]));
  RC(ctx->launch("pp_block_sizes", [&] { return fbk_pp_block_sizes(v, q, 1, ctx->d_vl_bsz.as<int64_t>(), st); }));
  RC(scan_i64(ctx, ctx->d_vl_bsz.as<int64_t>(), ctx->d_vl_off.as<int64_t>(), ctx->pp_npairs[PPL_PPP]));
  FB_HIP(hipMemcpyAsync(&ctx->sv_len, ctx->d_sv_off.as<int64_t>() + ctx->pp_npairs[PPL_PP], sizeof(int64_t),
                        hipMemcpyDeviceToHost, st));
  FB_HIP(hipMemcpyAsync(&ctx->vl_len, ctx->d_vl_off.as<int64_t>() + ctx->pp_npairs[PPL_PPP], sizeof(int64_t),
                        hipMemcpyDeviceToHost, st));
  FB_HIP(hipStreamSynchronize(st));
  FB_HIP(ctx->d_svnl.ensure(sizeof(double) * std::max<int64_t>(ctx->sv_len, 1)));
  FB_HIP(ctx->d_vnl.ensure(sizeof(double) * std::max<int64_t>(ctx->vl_len, 1)));
  const size_t ns = ctx->structs.size();
  ctx->h_sv_struct.assign(ns + 1, 0);
  ctx->h_vl_struct.assign(ns + 1, 0);
  {
    std::vector<int64_t> isv(ns), ivl(ns);
    for (size_t k = 0; k < ns; k++) {
      const int a0 = ctx->h_st[k].atom_off;
      isv[k] = ctx->h_ppoff[PPL_PP][a0];
      ivl[k] = ctx->h_ppoff[PPL_PPP][a0];
    }
    RC(gather_to_host(ctx, ctx->d_sv_off.as<int64_t>(), isv, ctx->h_sv_struct.data()));
    RC(gather_to_host(ctx, ctx->d_vl_off.as<int64_t>(), ivl, ctx->h_vl_struct.data()));
  }
  ctx->h_sv_struct[ns] = ctx->sv_len;
  ctx->h_vl_struct[ns] = ctx->vl_len;
  RC(ctx->check_status("driver_neighbors_PP"));
  for (int i = 0; i < 3; i++) ctx->ppdone[i] = 0;
  ctx->pp_lists_ok = true;
  if (npp) *npp = ctx->pp_npairs[PPL_PP];
  if (nppp) *nppp = ctx->pp_npairs[PPL_PPP];
  if (ntrip) *ntrip = T;
  return FB_OK;
}

int fb_get_pp_list_sizes(fb_ctx *ctx, int is, int64_t *sizes) {
  if (!ctx || !ctx->pp_lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "PP lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || !sizes) return ctx->fail(FB_E_ARG, "structure index");
  const int a0 = ctx->h_st[is].atom_off, a1 = a0 + ctx->h_st[is].natoms;
  for (int l = 0; l < 3; l++) sizes[l] = ctx->h_ppoff[l][a1] - ctx->h_ppoff[l][a0];
  sizes[3] = ctx->h_pcoff[a1] - ctx->h_pcoff[a0];
  sizes[4] = ctx->h_sv_struct[is + 1] - ctx->h_sv_struct[is];
  sizes[5] = ctx->h_vl_struct[is + 1] - ctx->h_vl_struct[is];
  return FB_OK;
}

int fb_get_pp_lists(fb_ctx *ctx, int is, int which, int *n, int *a0, int *a1, int *a2, int *a3, int *a4, int *self) {
  if (!ctx || !ctx->pp_lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "PP lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || which < 0 || which > 3 || !n)
    return ctx->fail(FB_E_ARG, "get_pp_lists arguments");
  FB_HIP(hipSetDevice(ctx->device));
  FB_HIP(hipStreamSynchronize(ctx->stream));
  const int at0 = ctx->h_st[is].atom_off, na = ctx->h_st[is].natoms;
  if (which < 3) {
    const int64_t p0 = ctx->h_ppoff[which][at0], p1 = ctx->h_ppoff[which][at0 + na];
    RC(d2h(ctx, n, ctx->d_ppn[which].as<int>() + at0, sizeof(int) * na));
    if (a0) RC(d2h(ctx, a0, ctx->d_ppj[which].as<int>() + p0, sizeof(int) * (p1 - p0)));
    if (a1) RC(d2h(ctx, a1, ctx->d_ppb[which].as<int>() + p0, sizeof(int) * (p1 - p0)));
    if (a2 && which == PPL_PP) RC(d2h(ctx, a2, ctx->d_pp_map.as<int>() + p0, sizeof(int) * (p1 - p0)));
    if (a2 && which == PPL_PPX) RC(d2h(ctx, a2, ctx->d_px_map.as<int>() + p0, sizeof(int) * (p1 - p0)));
    if (a3 && which == PPL_PPX) RC(d2h(ctx, a3, ctx->d_px_point.as<int>() + p0, sizeof(int) * (p1 - p0)));
    if (self) RC(d2h(ctx, self, ctx->d_ppself[which].as<int>() + at0, sizeof(int) * na));
    return FB_OK;
  }
  // common neighbours: local atom numbers and mneigh (index into iatom's PPp list)
  const int64_t t0 = ctx->h_pcoff[at0], t1 = ctx->h_pcoff[at0 + na];
  RC(d2h(ctx, n, ctx->d_pc_n.as<int>() + at0, sizeof(int) * na));
  std::vector<int4> tr(t1 - t0);
  std::vector<int> tp(t1 - t0);
  if (t1 > t0) {
    RC(d2h(ctx, tr.data(), ctx->d_pc_trip.as<int4>() + t0, sizeof(int4) * (t1 - t0)));
    RC(d2h(ctx, tp.data(), ctx->d_pc_tpair.as<int>() + t0, sizeof(int) * (t1 - t0)));
  }
  for (int64_t t = 0; t < t1 - t0; t++) {
    const int gia = tr[t].y, gja = tr[t].z;
    if (a0) a0[t] = gia - at0 + 1;
    if (a1) a1[t] = tr[t].w & 0xffff;
    if (a2) a2[t] = gja - at0 + 1;
    if (a3) a3[t] = (tr[t].w >> 16) & 0xffff;
    if (a4) a4[t] = (int)(tp[t] - ctx->h_ppoff[PPL_PPP][gia] + 1);
  }
  return FB_OK;
}

static int pp_step(fb_ctx *ctx, int step) {
  static const char *names[3] = {"assemble_svnl", "assemble_vnl_2c", "assemble_vnl_3c"};
  if (!ctx) return FB_E_ARG;
  if (!ctx->pp_lists_ok) return ctx->fail(FB_E_ORDER, (std::string(names[step]) + " before build_lists_pp").c_str());
  if (step > 0 && !ctx->ppdone[step - 1])
    return ctx->fail(FB_E_ORDER, (std::string(names[step]) + " before " + names[step - 1]).c_str());
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("previous assemble_all"));
  if (ctx->structs_dirty) RC(upload_structures(ctx));
  RC(ctx->clear_status());
  DView v = ctx->view();
  DPP q = ctx->pview();
  if (step == 0) RC(ctx->launch(names[0], [&] { return fbk_svnl(v, q, ctx->stream); }));
  if (step == 1) RC(ctx->launch(names[1], [&] { return fbk_vnl_2c(v, q, ctx->stream); }));
  if (step == 2) RC(ctx->launch(names[2], [&] { return fbk_vnl_3c(v, q, ctx->stream); }));
  RC(ctx->check_status(names[step]));
  ctx->ppdone[step] = 1;
  for (int k = step + 1; k < 3; k++) ctx->ppdone[k] = 0;
  return FB_OK;
}

int fb_assemble_svnl(fb_ctx *ctx) { return pp_step(ctx, 0); }
int fb_assemble_vnl_2c(fb_ctx *ctx) { return pp_step(ctx, 1); }
int fb_assemble_vnl_3c(fb_ctx *ctx) { return pp_step(ctx, 2); }

int fb_get_pp_blocks(fb_ctx *ctx, int is, int which, double *out) {
  if (!ctx || !ctx->pp_lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "PP lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || which < 0 || which > 1 || !out)
    return ctx->fail(FB_E_ARG, "get_pp_blocks arguments");
  if (!ctx->ppdone[which]) return ctx->fail(FB_E_ORDER, "get_pp_blocks: not assembled");
  FB_HIP(hipSetDevice(ctx->device));
  FB_HIP(hipStreamSynchronize(ctx->stream));
  const std::vector<int64_t> &o = which == 0 ? ctx->h_sv_struct : ctx->h_vl_struct;
  const double *src = (which == 0 ? ctx->d_svnl : ctx->d_vnl).as<double>();
  return d2h(ctx, out, src + o[is], sizeof(double) * (o[is + 1] - o[is]));
}

// ------------------------------------------------------------------ forces
// Buffers of the force assemblers and the two triplet orderings of the
// per-atom sums (iatom: the target-pair order of the lists; jatom: a stable
// radix sort of the list order by jatom), built once per list generation.
static int prep_forces(fb_ctx *ctx) {
  if (ctx->forces_gen == ctx->lists_gen) return FB_OK;
  hipStream_t st = ctx->stream;
  const int64_t P = ctx->npairs, T = ctx->ntrip, B = ctx->blen;
  const int N = ctx->natoms;
  auto one = [](int64_t n) { return (size_t)std::max<int64_t>(n, 1); };
  FB_HIP(ctx->d_rho.ensure(sizeof(double) * one(B)));
  for (int i = 0; i < 3; i++) FB_HIP(ctx->d_dmat[i].ensure(sizeof(double) * one(3 * B)));
  FB_HIP(ctx->d_fatom.ensure(sizeof(double) * one(3 * P)));
  FB_HIP(ctx->d_fontop.ensure(sizeof(double) * one(3 * P)));
  FB_HIP(ctx->d_f3a.ensure(sizeof(double) * one(3 * T)));
  FB_HIP(ctx->d_f3b.ensure(sizeof(double) * one(3 * T)));
  FB_HIP(ctx->d_f3c.ensure(sizeof(double) * one(3 * T)));
  FB_HIP(ctx->d_inv_perm.ensure(sizeof(int) * one(T)));
  FB_HIP(ctx->d_inv_perm_j.ensure(sizeof(int) * one(T)));
  FB_HIP(ctx->d_f3.ensure(sizeof(double) * one(9 * (int64_t)N)));
  FB_HIP(ctx->d_ioff.ensure(sizeof(int64_t) * (N + 1)));
  FB_HIP(ctx->d_joff.ensure(sizeof(int64_t) * (N + 1)));
  FB_HIP(ctx->d_jkey.ensure(sizeof(int) * one(T)));
  FB_HIP(ctx->d_jkey_s.ensure(sizeof(int) * one(T)));
  FB_HIP(ctx->d_iota_t.ensure(sizeof(int) * one(T)));
  FB_HIP(ctx->d_perm_j.ensure(sizeof(int) * one(T)));
  if (T > 0) {
    RC(ctx->launch("trip_jkey", [&] { return fbk_trip_jkey(T, ctx->d_trip_orig.as<int4>(), ctx->d_jkey.as<int>(), st); }));
    RC(ctx->launch("iota", [&] { return fbk_iota(ctx->d_iota_t.as<int>(), T, st); }));
    int end_bit = 1;
    while ((1ll << end_bit) <= N) end_bit++;
    size_t bytes = 0;
    int rc = fbk_sort_pairs(ctx->d_jkey.as<int>(), ctx->d_jkey_s.as<int>(), ctx->d_iota_t.as<int>(),
                            ctx->d_perm_j.as<int>(), T, end_bit, nullptr, &bytes, st);
    if (rc) return ctx->fail(FB_E_HIP, "jatom sort sizing", (hipError_t)rc);
    RC(ensure_tmp(ctx, bytes));
    bytes = ctx->d_tmp.cap;
    RC(ctx->launch("sort_jatom", [&] {
      return fbk_sort_pairs(ctx->d_jkey.as<int>(), ctx->d_jkey_s.as<int>(), ctx->d_iota_t.as<int>(),
                            ctx->d_perm_j.as<int>(), T, end_bit, ctx->d_tmp.p, &bytes, st);
    }));
    // the inverse orders: where each triplet's iatom / jatom record goes
    RC(ctx->launch("invert_perm", [&] { return fbk_invert_perm(T, ctx->d_perm.as<int>(), ctx->d_inv_perm.as<int>(), st); }));
    RC(ctx->launch("invert_perm", [&] {
      return fbk_invert_perm(T, ctx->d_perm_j.as<int>(), ctx->d_inv_perm_j.as<int>(), st);
    }));
  }
  RC(ctx->launch("f3_offsets", [&] {
    return fbk_f3_offsets(N, T, ctx->d_noff.as<int64_t>(), ctx->d_tpair.as<int>(), ctx->d_jkey_s.as<int>(),
                          ctx->d_ioff.as<int64_t>(), ctx->d_joff.as<int64_t>(), st);
  }));
  FB_HIP(hipStreamSynchronize(st));
  ctx->rho_ok.assign(ctx->structs.size(), 0);
  for (int i = 0; i < 5; i++) ctx->ddone[i] = 0;
  ctx->forces_gen = ctx->lists_gen;
  return FB_OK;
}

int fb_set_denmat(fb_ctx *ctx, int is, const double *rho) {
  if (!ctx) return FB_E_ARG;
  if (!ctx->lists_ok) return ctx->fail(FB_E_ORDER, "set_denmat before build_lists");
  if (is < 0 || is >= (int)ctx->structs.size() || !rho) return ctx->fail(FB_E_ARG, "set_denmat arguments");
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("previous assemble_all"));
  RC(prep_forces(ctx));
  const int64_t b0 = ctx->h_boff_struct[is], b1 = ctx->h_boff_struct[is + 1];
  FB_HIP(hipMemcpyAsync(ctx->d_rho.as<double>() + b0, rho, sizeof(double) * (b1 - b0), hipMemcpyHostToDevice,
                        ctx->stream));
  FB_HIP(hipStreamSynchronize(ctx->stream));
  ctx->rho_ok[is] = 1;
  ctx->ddone[3] = ctx->ddone[4] = 0;
  return FB_OK;
}

static int pre_dassemble(fb_ctx *ctx, const char *what) {
  RC(pre_assemble(ctx, what));
  RC(prep_forces(ctx));
  return FB_OK;
}

static bool all_rho(const fb_ctx *ctx) {
  for (char c : ctx->rho_ok)
    if (!c) return false;
  return !ctx->rho_ok.empty();
}

// s/p species with canonical tables take the register-resident kernels
static int launch_d2c(fb_ctx *ctx, int mask, const char *what) {
  DView v = ctx->view();
  DForce f = ctx->fview();
  if (ctx->canon_sp && ctx->force_kernel == 1)
    return ctx->launch(what, [&] { return fbk_d2c_sp(v, f, mask, ctx->stream); });
  return ctx->launch(what, [&] { return fbk_d2c(v, f, mask, ctx->stream); });
}

static int dassemble_2c(fb_ctx *ctx, int bit, const char *what) {
  if (!ctx) return FB_E_ARG;
  RC(pre_dassemble(ctx, what));
  RC(launch_d2c(ctx, 1 << bit, what));
  RC(ctx->check_status(what));
  ctx->ddone[bit] = 1;
  return FB_OK;
}

int fb_dassemble_S(fb_ctx *ctx) { return dassemble_2c(ctx, 0, "Dassemble_S"); }
int fb_dassemble_T(fb_ctx *ctx) { return dassemble_2c(ctx, 1, "Dassemble_T"); }
int fb_dassemble_dipole_z(fb_ctx *ctx) { return dassemble_2c(ctx, 2, "Dassemble_dipole_z"); }

int fb_dassemble_vna_2c(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_dassemble(ctx, "Dassemble_vna_2c"));
  if (!all_rho(ctx)) return ctx->fail(FB_E_ORDER, "Dassemble_vna_2c needs the density matrix (fb_set_denmat)");
  if (!ctx->done[FB_MAT_S]) return ctx->fail(FB_E_ORDER, "Dassemble_vna_2c needs assemble_S first");
  RC(launch_d2c(ctx, 8, "Dassemble_vna_2c"));
  RC(ctx->check_status("Dassemble_vna_2c"));
  ctx->ddone[3] = 1;
  return FB_OK;
}

static int run_d3c(fb_ctx *ctx) {
  DView v = ctx->view();
  DForce f = ctx->fview();
  if (ctx->canon_sp && ctx->force_kernel == 1)
    RC(ctx->launch("Dassemble_vna_3c", [&] {
      return fbk_d3c_sp(v, f, any_charge(ctx), ctx->precision == FB_PREC_MIXED, ctx->stream);
    }));
  else
    RC(ctx->launch("Dassemble_vna_3c", [&] { return fbk_d3c(v, f, ctx->stream); }));
  RC(ctx->launch("forces_3c_sum", [&] { return fbk_f3_reduce(v, f, ctx->stream); }));
  return FB_OK;
}

int fb_dassemble_vna_3c(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_dassemble(ctx, "Dassemble_vna_3c"));
  if (!all_rho(ctx)) return ctx->fail(FB_E_ORDER, "Dassemble_vna_3c needs the density matrix (fb_set_denmat)");
  if (!ctx->done[FB_MAT_S] || !ctx->done[FB_MAT_DIP] || !ctx->ddone[0] || !ctx->ddone[2])
    return ctx->fail(FB_E_ORDER,
                     "Dassemble_vna_3c needs assemble_S, assemble_dipole_z, Dassemble_S and Dassemble_dipole_z first");
  RC(run_d3c(ctx));
  RC(ctx->check_status("Dassemble_vna_3c"));
  ctx->ddone[4] = 1;
  return FB_OK;
}

int fb_dassemble_all(fb_ctx *ctx) {
  if (!ctx) return FB_E_ARG;
  RC(pre_dassemble(ctx, "Dassemble_all"));
  if (!all_rho(ctx)) return ctx->fail(FB_E_ORDER, "Dassemble_all needs the density matrix (fb_set_denmat)");
  if (!ctx->done[FB_MAT_S] || !ctx->done[FB_MAT_DIP])
    return ctx->fail(FB_E_ORDER, "Dassemble_all needs assemble_S and assemble_dipole_z first");
  RC(launch_d2c(ctx, 15, "Dassemble_2c"));
  RC(run_d3c(ctx));
  RC(ctx->check_status("Dassemble_all"));
  for (int i = 0; i < 5; i++) ctx->ddone[i] = 1;
  return FB_OK;
}

int fb_get_dblocks(fb_ctx *ctx, int is, int matrix, double *out) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size() || matrix < 0 || matrix > 2 || !out)
    return ctx->fail(FB_E_ARG, "get_dblocks arguments");
  if (ctx->forces_gen != ctx->lists_gen || !ctx->ddone[matrix])
    return ctx->fail(FB_E_ORDER, "get_dblocks: Dblocks not assembled");
  FB_HIP(hipStreamSynchronize(ctx->stream));
  const int64_t b0 = ctx->h_boff_struct[is], b1 = ctx->h_boff_struct[is + 1];
  return d2h(ctx, out, ctx->d_dmat[matrix].as<double>() + 3 * b0, sizeof(double) * 3 * (b1 - b0));
}

int fb_get_forces(fb_ctx *ctx, int is, double *vna_atom, double *vna_ontop, double *f3naa, double *f3nab,
                  double *f3nac) {
  if (!ctx || !ctx->lists_ok) return ctx ? ctx->fail(FB_E_ORDER, "lists not built") : FB_E_ARG;
  if (is < 0 || is >= (int)ctx->structs.size()) return ctx->fail(FB_E_ARG, "get_forces arguments");
  if (ctx->forces_gen != ctx->lists_gen) return ctx->fail(FB_E_ORDER, "get_forces: forces not assembled");
  if ((vna_atom || vna_ontop) && !ctx->ddone[3]) return ctx->fail(FB_E_ORDER, "get_forces: Dassemble_vna_2c not run");
  if ((f3naa || f3nab || f3nac) && !ctx->ddone[4]) return ctx->fail(FB_E_ORDER, "get_forces: Dassemble_vna_3c not run");
  FB_HIP(hipStreamSynchronize(ctx->stream));
  const DStruct &d = ctx->h_st[is];
  const int64_t p0 = ctx->h_noff[d.atom_off], p1 = ctx->h_noff[d.atom_off + d.natoms];
  RC(d2h(ctx, vna_atom, ctx->d_fatom.as<double>() + 3 * p0, sizeof(double) * 3 * (p1 - p0)));
  RC(d2h(ctx, vna_ontop, ctx->d_fontop.as<double>() + 3 * p0, sizeof(double) * 3 * (p1 - p0)));
  if (f3naa || f3nab || f3nac) {
    std::vector<double> h((size_t)9 * d.natoms);
    RC(d2h(ctx, h.data(), ctx->d_f3.as<double>() + (size_t)9 * d.atom_off, sizeof(double) * h.size()));
    double *dst[3] = {f3naa, f3nab, f3nac};
    for (int r = 0; r < 3; r++)
      if (dst[r])
        for (int a = 0; a < d.natoms; a++)
          for (int k = 0; k < 3; k++) dst[r][3 * a + k] = h[(size_t)9 * a + 3 * r + k];
  }
  return FB_OK;
}

int fb_set_tuning(fb_ctx *ctx, int param, int value) {
  if (!ctx) return FB_E_ARG;
  if (param == FB_TUNE_CHUNK3C) {
    if (value != 64 && value != 128 && value != 256) return ctx->fail(FB_E_ARG, "chunk must be 64, 128 or 256");
    ctx->chunk3c = value;
    return FB_OK;
  }
  if (param == FB_TUNE_GRAPH) {
    if (value != 0 && value != 1) return ctx->fail(FB_E_ARG, "graph replay must be 0 or 1");
    ctx->use_graph = value != 0;
    return FB_OK;
  }
  if (param == FB_TUNE_2C_KERNEL) {
    if (value != 0 && value != 1) return ctx->fail(FB_E_ARG, "2c kernel must be 0 (LDS-staged) or 1 (register)");
    ctx->k2c = value;
    ctx->mlp_gen++;   // captured graphs hold the old launch
    return FB_OK;
  }
  if (param == FB_TUNE_FORCE_KERNEL) {
    if (value != 0 && value != 1) return ctx->fail(FB_E_ARG, "force kernel must be 0 (generic) or 1 (s/p)");
    ctx->force_kernel = value;
    return FB_OK;
  }
  if (param == FB_TUNE_FUSED) {
    if (value != 0 && value != 3 && value != 4) return ctx->fail(FB_E_ARG, "fused must be 0, 3 or 4");
    ctx->fused = value;
    return FB_OK;
  }
  if (param == FB_TUNE_PERSIST) {
    if (value != 0 && value != 1) return ctx->fail(FB_E_ARG, "persist must be 0 or 1");
    FB_HIP(hipSetDevice(ctx->device));
    if (value) FB_HIP(ctx->d_queue.ensure(FB_QUEUE_INTS * sizeof(int)));
    ctx->persist = value;
    ctx->mlp_gen++;   // captured graphs hold the old launch
    return FB_OK;
  }
  return ctx->fail(FB_E_ARG, "unknown tuning parameter");
}

int fb_get_tuning(fb_ctx *ctx, int param, int *value) {
  if (!ctx || !value) return FB_E_ARG;
  if (param == FB_TUNE_CHUNK3C) {
    *value = ctx->chunk3c;
    return FB_OK;
  }
  if (param == FB_TUNE_2C_KERNEL) {
    *value = ctx->k2c;
    return FB_OK;
  }
  if (param == FB_TUNE_GRAPH) {
    *value = ctx->use_graph ? 1 : 0;
    return FB_OK;
  }
  if (param == FB_TUNE_FORCE_KERNEL) {
    *value = ctx->force_kernel;
    return FB_OK;
  }
  if (param == FB_TUNE_FUSED) {
    *value = ctx->fused;
    return FB_OK;
  }
  if (param == FB_TUNE_PERSIST) {
    *value = ctx->persist;
    return FB_OK;
  }
  return ctx->fail(FB_E_ARG, "unknown tuning parameter");
}

// Times whole H+S builds, run as fb_assemble_all runs them (graph replays
// unless graphs are off), over every
// launch shape that applies to the context and keeps the fastest: the
// one-launch build at 3 and 4 waves per SIMD (neutral mixed-precision s/p
// builds), and separate 2c and 3c launches with every 3c chunk (mixed s/p
// builds; it sets the triplets and the LDS footprint of a workgroup) and both
// s/p 2c kernels, and the one launch with the 3c tile queue.  The last
// candidate leaves the blocks of its own build, so the build is run once more
// with the chosen shape before returning.
int fb_autotune(fb_ctx *ctx, int nrep, int *chosen) {
  if (!ctx) return FB_E_ARG;
  if (!ctx->lists_ok || !ctx->done[FB_MAT_S] || !ctx->done[FB_MAT_DIP] || !ctx->done[FB_MAT_VNA])
    return ctx->fail(FB_E_ORDER, "autotune needs lists and the 2c blocks (run an assemble first)");
  if (nrep < 1) nrep = 5;
  FB_HIP(hipSetDevice(ctx->device));
  RC(ctx->settle("assemble_all"));
  if (ctx->structs_dirty) RC(upload_structures(ctx));
  struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
    }
  } ev;
  FB_HIP(hipEventCreate(&ev.a));
  FB_HIP(hipEventCreate(&ev.b));
  const int charged = any_charge(ctx);
  const bool timing = ctx->timing;
  ctx->timing = false;   // builds then take the graph (or, without graphs, the forked two-stream form)
  struct Cfg {
    int chunk, k2c, fused, persist;
  };
  std::vector<Cfg> cand;
  const bool mixed_sp = ctx->canon_sp && ctx->precision == FB_PREC_MIXED;
  const bool mlp = mixed_sp && ctx->method3c == FB_3C_MLP && ctx->mlp_ready && !charged;
  if (mixed_sp && !charged && !mlp) {
    FB_HIP(ctx->d_queue.ensure(FB_QUEUE_INTS * sizeof(int)));
    for (int w : {3, 4})
      for (int q : {0, 1}) cand.push_back({ctx->chunk3c, ctx->k2c, w, q});
  }
  for (int c : {64, 128, 256}) {
    if (!mixed_sp && c != ctx->chunk3c) continue;   // the chunk only shapes the mixed s/p 3c kernel
    for (int k : {0, 1}) {
      if (!ctx->canon_sp && k != ctx->k2c) continue;   // the s/p 2c kernels
      cand.push_back({c, k, 0, 0});
    }
  }
  Cfg best{ctx->chunk3c, ctx->k2c, ctx->fused, ctx->persist};
  Cfg best_sep = best;   // the fastest separate-launch shape: its chunk and 2c kernel stay set
  float tbest = 1e30f, tsep = 1e30f;
  hipError_t e = hipSuccess;
  int rc = FB_OK;
  // two sweeps over the candidates, the best time of each counting, so that
  // clock ramp-up in the first sweep does not decide
  std::vector<Cfg> sweep(cand);
  sweep.insert(sweep.end(), cand.begin(), cand.end());
  for (const Cfg &c : sweep) {
    ctx->chunk3c = c.chunk;
    ctx->k2c = c.k2c;
    ctx->fused = c.fused;
    ctx->persist = c.persist;
    rc = launch_build(ctx, charged);   // warm-up (and the capture of this shape's graph)
    if (rc == FB_OK) e = hipEventRecord(ev.a, ctx->stream);
    for (int r = 0; r < nrep && rc == FB_OK && e == hipSuccess; r++) rc = launch_build(ctx, charged);
    if (rc == FB_OK && e == hipSuccess) e = hipEventRecord(ev.b, ctx->stream);
    if (rc == FB_OK && e == hipSuccess) e = hipEventSynchronize(ev.b);
    float ms = 0.f;
    if (rc == FB_OK && e == hipSuccess) e = hipEventElapsedTime(&ms, ev.a, ev.b);
    if (rc != FB_OK || e != hipSuccess) break;
    if (ms < tbest) {
      tbest = ms;
      best = c;
    }
    if (c.fused == 0 && ms < tsep) {
      tsep = ms;
      best_sep = c;
    }
  }
  // a one-launch winner leaves chunk and 2c kernel to the separate launches
  // (per-kernel timing, charged builds), which take the best separate shape's
  if (best.fused && tsep < 1e30f) {
    best.chunk = best_sep.chunk;
    best.k2c = best_sep.k2c;
  }
  ctx->chunk3c = best.chunk;
  ctx->k2c = best.k2c;
  ctx->fused = best.fused;
  ctx->persist = best.persist;
  ctx->mlp_gen++;   // captured graphs hold the previous launch shapes
  if (rc == FB_OK && e == hipSuccess) rc = launch_build(ctx, charged);
  ctx->timing = timing;
  if (e != hipSuccess) return ctx->fail(FB_E_HIP, "autotune", e);
  RC(rc);
  ctx->pending = true;
  RC(ctx->settle("autotune"));
  for (int i = 0; i < 5; i++) ctx->done[i] = 1;
  if (chosen) *chosen = best.chunk;
  return FB_OK;
}

int fb_enable_timing(fb_ctx *ctx, int on) {
  if (!ctx) return FB_E_ARG;
  ctx->timing = on != 0;
  return FB_OK;
}

int fb_get_timings(fb_ctx *ctx, int max, char *names, float *ms, int *count) {
  if (!ctx) return FB_E_ARG;
  FB_HIP(hipStreamSynchronize(ctx->stream));
  RC(ctx->settle("assemble_all"));
  int n = 0;
  for (auto &e : ctx->events) {
    if (n >= max) break;
    float t = 0.f;
    FB_HIP(hipEventElapsedTime(&t, e.a, e.b));
    if (ms) ms[n] = t;
    if (names) {
      std::strncpy(names + 32 * n, e.name.c_str(), 31);
      names[32 * n + 31] = 0;
    }
    n++;
  }
  if (count) *count = n;
  ctx->reset_events();
  return FB_OK;
}

}  // extern "C"
