// Internal data layout shared by the host library and the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fireball_hip.h"

#define FB_NIMG 729           // (2*4+1)^3 images, c.SYSTEM/make_cells.f90:126
#define FB_NTHETA 5           // include/interactions_3c.h:4
#define FB_MAXSUB2 10         // include/interactions_2c.h:3
#define FB_MAXSUB3 7          // include/interactions_3c.h:3
#define FB_NSLOT2 13          // 2c interactions used by the assemblers
#define FB_P_EQ2 14.39975     // include/constants.h:4
#define FB_2C_PCAP 256        // pairs per 2c workgroup pass
#define FB_2C_SKIP_TGT 16     // 2c mask bit: leave the Vna blocks of 3c targets to the 3c launch

// 2c interaction slots (device-side dense index of the interaction ids of
// include/interactions_2c.h: overlap 1, kinetic 2, vna_ontopL 3, vna_ontopR 4,
// vna_atom 5, dipole_z 10; McWEDA densities rho_ontopL 15, rho_ontopR 16,
// rho_atom 17, rhoS_ontopL 19, rhoS_ontopR 20, rhoS_atom 21; non-local
// pseudopotential vnl 6)
enum {
  SLOT_S = 0, SLOT_T = 1, SLOT_ONL = 2, SLOT_ONR = 3, SLOT_ATOM = 4, SLOT_DIP = 5,
  SLOT_RHO_ONL = 6, SLOT_RHO_ONR = 7, SLOT_RHO_ATOM = 8, SLOT_RHOS_ONL = 9, SLOT_RHOS_ONR = 10, SLOT_RHOS_ATOM = 11,
  SLOT_VNL = 12
};
#define FB_NMAT 9             // block arrays: enum fb_matrix

// status bits raised by kernels (one word per context)
enum {
  ST_X_RANGE = 1, ST_X_NEG = 2, ST_TOO_CLOSE = 4, ST_SELF = 8, ST_BACK = 16, ST_NO_TABLE = 32,
  ST_CAPACITY = 64
};

struct DSpecies {
  int ns;
  int have_d;
  int nssh[FB_MAXSPECIES];
  int lssh[FB_MAXSPECIES][FB_MAXSH];
  int norb[FB_MAXSPECIES];
  double rcmax[FB_MAXSPECIES];   // max shell rcutoffA (neighbour range)
  double rcmin[FB_MAXSPECIES];   // min shell rcutoffA (smoothing)
};

#define FB_MO 4           // orbitals per species of the s/p kernels
#define FB_MOD 9          // orbitals per species of the generic binary64 kernels (s, p and d shells)
#define FB_NME2_WIDE 25   // 2c matrix elements the generic kernels hold with d shells (19 for s, p, d)
#define FB_NME3_WIDE 48   // 3c matrix elements likewise (45 for s, p, d)

// tile queues of the persistent 3c launches: one counter per XCD, each in a
// 256-byte line of its own (counters sharing a line serialise their atomics)
#define FB_QSTRIDE 64
#define FB_QUEUE_INTS (8 * FB_QSTRIDE)

// some species has more than FB_MO orbitals: every assembler then runs its
// generic kernels at their FB_MOD sizes
inline bool fb_wide_species(const DSpecies &sp) {
  for (int i = 0; i < sp.ns; i++)
    if (sp.norb[i] > FB_MO) return true;
  return false;
}

struct DCell2c {
  int nME, nx;
  double xmax, dx;
  int64_t off;       // first value in the FP64 slab, layout [nx][nME]
  int64_t off32;     // same cell in the FP32 slab
  int munu;          // first entry in mu/nu/mvalue
  int grid;          // distinct (nx, dx) grid, see DGrid2c
};

// Lagrange denominators 1/(xx_is - xx_ip) of every 7-point window of a grid,
// evaluated on the host with the reference's expression
// (e.FDATA/M_Fdata_2c.f90:379-389); layout [ileft-1][is][ip]
struct DGrid2c {
  int nx, pad;
  double dx;
  int64_t pden_off;
};

// theta values per (point, ME) in the FP32 3c slab, packed without padding:
// one float4 + one float per load, dword aligned
#define FB_TH32 5

// one (is, js, ks, isub) group of the five Legendre theta cells
struct DGroup3c {
  int nME, nx, ny, munu;
  int cls;           // s/p pair class of (is, js): 2 if is has p, + 1 if js has p
  int nq;            // ME quads of the FP32 quad slab, ceil(nME / 4)
  double xmax, ymax, dx, dy;
  double rdx, rdy;   // 1/dx, 1/dy
  int64_t off;       // FP64 slab, layout [theta][ny][nx][nME]
  int64_t off32;     // FP32 slab, layout [ny][nx][ME][theta], FB_TH32 values per (point, ME)
  int64_t off32q;    // FP32 quad slab, layout [ny][nx][theta][4 nq] (MEs zero-padded to whole quads)
  int64_t moff;      // method table [ny][nx] (uint8, 1 = 3-point rule), e.FDATA/M_Fdata_3c.f90:383-385
};

struct DStruct {
  int atom_off, natoms, mbeta_max, icluster;
  int64_t pair_off, trip_off;
  double bmin[3], bmax[3];
  double rcmax_all;
  // cell list of the structure's atoms (home positions): cn cells per axis of
  // width cw >= crange from clo; every atom within crange of a point lies in a
  // cell that overlaps the box point +- crange
  double clo[3], cw[3], crange;
  int cn[3], cell_off;
};

// all device pointers of a context, passed by value to kernels
struct DView {
  DSpecies sp;
  // tables
  const DCell2c *cell2c;        // cells
  int ncell2c, ngrid2c;         // cell and grid counts (small: the 2c kernel caches them in LDS)
  const int *idx2c;             // [ns][ns][FB_NSLOT2][FB_MAXSUB2] -> cell or -1
  const DGroup3c *grp3c;
  const int *idx3c;             // [ns][ns][ns][FB_MAXSUB3] -> group or -1 (itype = bcna)
  const int *idx3r;             // [ns][ns][ns][2][FB_MAXSUB3] -> group or -1 (itype rho_3c 1, rhoS_3c 2)
  const double *tab2c;
  const float *tab2c32;
  const double *tab3c;
  const float *tab3c32;
  const float *tab3c32q;        // FP32 quad slab of the bcna groups (DGroup3c::off32q)
  const unsigned char *meth3c;  // per-cell method choice of every 3c group
  int64_t ntab3c32;             // floats in the FP32 3c slab
  int64_t ntab3c32q;            // floats in the FP32 quad slab
  const int *mu, *nu, *mv;
  const DGrid2c *grid2c;
  const double *pden;
  int canon_sp;                 // every species s or sp, every table in canonical ME order
  // structures
  int nstruct, natoms;
  const DStruct *st;
  const double *xl;             // [nstruct][FB_NIMG][3]
  const double *ratom;          // [natoms][3]
  const int *imass;             // [natoms]
  const int *atom_struct;       // [natoms]
  const double *dQ;             // [natoms][FB_MAXSH]
  int any_dq;                   // some shell charge of the batch is non-zero
  const double *qin;            // [natoms][FB_MAXSH] input shell charges Qin (McWEDA densities)
  // neighbour lists (CSR over global atoms, per-structure 1-based atom ids)
  int64_t npairs;
  const int *neighn;            // [natoms]
  const int64_t *noff;          // [natoms+1] global pair offset
  const int *neigh_j, *neigh_b, *neigh_back, *neigh_self;
  const int *pair_atom;         // [npairs] global iatom (0-based)
  const int4 *pair_rec;         // [npairs] global iatom, global jatom (0-based), structure * FB_NIMG + mbeta,
                                //          imass_i | imass_j << 16: the pair in one load
  const unsigned char *pair_perm;   // [npairs] lane order of the 2c workgroups (k_pair_perm), or null
  const int64_t *boff;          // [npairs+1] block offsets (doubles)
  const int64_t *wboff;         // [npairs+1] shell-block offsets (nssh_mu x nssh_nu per pair)
  const int *wg2c;              // [nwg2c+1] first atom of each 2c workgroup
  int nwg2c;
  // triplets, sorted by target pair (segment order = reference accumulation order)
  int64_t ntrip, nseg;
  const int4 *trip;             // [ntrip] (ialpha_g, iatom_g, jatom_g, packed beta: ibeta | jbeta<<16)
  const int *trip_pair;         // [ntrip] target pair (global)
  const int64_t *seg_off;       // [nseg+1]
  const int *seg_pair;          // [nseg] target pair of each segment
  const int *pair_tgt;          // [npairs] 1 where some triplet accumulates into the pair's Vna block
  const int *trip_aux;          // [ntrip] neutral 3c group | structure << 16 (species and cells of the record)
  const int4 *tiles[3];         // per 3c chunk of 64, 128, 256 triplets: segments first..last packed whole,
                                // their triplet range r0, r1 at most one chunk unless one segment is longer
  int ntiles[3];
  const int *trip_seg;          // [ntrip] segment of each sorted triplet
  const double *pair_disp;      // [npairs][3] xl(jbeta) + r_j - r_i of each neighbour pair (common-neighbour search)
  const int64_t *cell_start;    // [ncells+1] first entry of each cell in cell_atoms (cells of all structures)
  const int *cell_atoms;        // [natoms] local atom index, grouped by cell
  int cell_bits;                // 32-bit words of the per-image atom bitmap (0: no cell list, scan all atoms)
  // learned 3c surrogate (one DMlp per 3c group; ready = 0 where untrained)
  const struct DMlp *mlp;
  const float *mlp_err;         // calibrated per-cell error of each network
  float mlp_tau;                // per-triplet bound above which the table path is used
  unsigned long long *mlp_count;  // [2]: triplets on the network, triplets seen
  // outputs
  double *mat[FB_NMAT];
  int *status;
};

#define FB_HIP(call)                                                      \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess) return ctx->fail(FB_E_HIP, #call, e_);          \
  } while (0)
