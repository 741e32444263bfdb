/* fireball_hip.h -- C ABI of the MI355X (gfx950) two- and three-centre
 * Hamiltonian / overlap assembler.
 *
 * The entry points replace, one for one, the reference's Fortran routines
 * listed next to each declaration (paths relative to the reference's src/
 * tree).  Arguments are plain pointers and sizes; arrays that mirror a
 * reference array use the reference's conventions:
 *   - atom, species, neighbour and matrix-element indices are 1-based;
 *   - image indices (neigh_b, *_common_b) run 0..mbeta_max in make_cells order;
 *   - a block is column-major norb_mu x norb_nu, blocks are packed in pair order
 *     (iatom major, ineigh minor), exactly T_assemble_block%block per
 *     (iatom, ineigh) of a.GLOBAL/M_assemble_blocks.f90:44-64;
 *   - table data arrive in Fortran order, Fdata_2c(ix, iME) and
 *     Fdata_3c(ix, iY, iME) (e.FDATA/M_Fdata_2c.f90:92, M_Fdata_3c.f90:89).
 * Every function returns FB_OK (0) or a negative FB_E_* status instead of the
 * reference's `stop`; fb_last_error() gives the message.  A context is
 * single-threaded; work is ordered on the context's HIP stream.
 */
#ifndef FIREBALL_HIP_H
#define FIREBALL_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FB_MAXSH 4          /* shells per species */
#define FB_MAXSPECIES 8

enum fb_status {
  FB_OK = 0,
  FB_E_ARG = -1,
  FB_E_X_NEGATIVE = -2,       /* e.FDATA/M_Fdata_2c.f90:347-348 */
  FB_E_X_OUT_OF_RANGE = -3,   /* e.FDATA/M_Fdata_2c.f90:349-353 */
  FB_E_ATOMS_TOO_CLOSE = -4,  /* c.SYSTEM/M_neighbors.f90:421-430 */
  FB_E_NEIGH_SELF = -5,       /* c.SYSTEM/M_neighbors.f90:462-465, 509-515 */
  FB_E_NEIGH_BACK = -6,       /* c.SYSTEM/M_neighbors.f90:478-506 */
  FB_E_NO_TABLE = -7,         /* interaction cell absent from Fdata */
  FB_E_ORDER = -8,            /* assemble_* called before its inputs exist */
  FB_E_HIP = -9,              /* HIP runtime error */
  FB_E_CAPACITY = -10,        /* a list exceeds a kernel limit */
  FB_E_UNSUPPORTED = -11      /* e.g. a shell with l > 2, or a PP projector shell beyond p */
};

enum fb_matrix {            /* which block array */
  FB_MAT_S = 0,             /* s%overlap(:)%neighbors(:)%block */
  FB_MAT_T = 1,             /* s%kinetic(:)%neighbors(:)%block */
  FB_MAT_DIP = 2,           /* s%dipole_z(:)%neighbors(:)%block */
  FB_MAT_VNA = 3,           /* s%vna(:)%neighbors(:)%block  (atom term on neigh_self, 3c elsewhere) */
  FB_MAT_VNAO = 4,          /* s%vna(:)%neighbors(:)%blocko (ontop terms) */
  FB_MAT_RHO_IN = 5,        /* s%rho_in(:)%neighbors(:)%block            (McWEDA, orbital blocks) */
  FB_MAT_RHO_BOND = 6,      /* s%rho_bond(:)%neighbors(:)%block          (McWEDA, orbital blocks) */
  FB_MAT_RHOW_IN = 7,       /* s%rho_in_weighted(:)%neighbors(:)%block   (McWEDA, shell blocks) */
  FB_MAT_RHOW_BOND = 8      /* s%rho_bond_weighted(:)%neighbors(:)%block (McWEDA, shell blocks) */
};

enum fb_precision {
  FB_PREC_FP64 = 0,         /* plain binary64 kernels (parity anchor) */
  FB_PREC_MIXED = 1         /* FP32 interpolation/rotation, FP64 geometry + accumulation + blocks */
};

enum fb_method3c {
  FB_3C_TABLE = 0,          /* bivariate table interpolation (reference method) */
  FB_3C_MLP = 1             /* learned surrogate with table fallback */
};

typedef struct fb_ctx fb_ctx;

/* --- context ------------------------------------------------------------ */
int fb_create(fb_ctx **ctx, int device);
int fb_destroy(fb_ctx *ctx);
const char *fb_last_error(const fb_ctx *ctx);          /* NULL: this thread's last context-free failure (fb_create, fb_read_positions) */
int fb_set_stream(fb_ctx *ctx, void *hip_stream);      /* NULL: the context's own stream */
int fb_synchronize(fb_ctx *ctx);
int fb_set_precision(fb_ctx *ctx, int mode);          /* enum fb_precision */
int fb_set_3c_method(fb_ctx *ctx, int method);        /* enum fb_method3c */

/* --- species: read_info, c.SYSTEM/M_species.f90:1057-1251 ----------------
 * lssh and rcutoffA are [nspecies][FB_MAXSH]; rcutoffA in Angstrom (the
 * reference's shell(issh)%rcutoffA).  Shells s, p and d (l <= 2), at most 9
 * orbitals per species.  Species with d shells (or more than 4 orbitals) are
 * assembled by the H+S entry points (S, T, dipole_z, vna_2c, vna_3c, their
 * lists and blocks), the McWEDA density, force and pseudopotential entry
 * points in binary64 whatever the precision mode -- the d rotation of twister
 * and its derivative, c.SYSTEM/M_rotations.f90:272-300,
 * M_Drotations.f90:203-298.  Pseudopotential projector shells are s or p. */
int fb_set_species(fb_ctx *ctx, int nspecies, const int *nZ, const int *nssh, const int *lssh,
                   const double *rcutoffA);

/* --- tables --------------------------------------------------------------
 * fb_set_table_2c: one T_Fdata_cell_2c of Fdata_bundle_2c(is, js), as filled by
 *   read_Fdata_2c, e.FDATA/M_Fdata_2c.f90:149-272 (index_2c(itype, isub)).
 * fb_set_table_3c: one T_Fdata_cell_3c of Fdata_bundle_3c(is, js, ks), as filled
 *   by read_Fdata_3c, e.FDATA/M_Fdata_3c.f90:143-264 (index_3c(itype, isub, itheta)).
 * fb_commit_tables packs all cells into device slabs; call it once after the
 * last fb_set_table_* and before any assemble. */
int fb_set_table_2c(fb_ctx *ctx, int is, int js, int itype, int isub, int nME, int nx, double xmax,
                    const int *mu, const int *nu, const int *mvalue, const double *fdata);
int fb_set_table_3c(fb_ctx *ctx, int is, int js, int ks, int itype, int isub, int itheta, int nME,
                    int nx, double xmax, int ny, double ymax, const int *mu, const int *nu,
                    const int *mvalue, const double *fdata);
int fb_commit_tables(fb_ctx *ctx);
/* fb_read_fdata: read_Fdata_2c + read_Fdata_3c (e.FDATA/M_Fdata_2c.f90:149-272,
 * M_Fdata_3c.f90:143-264) from a directory in the reference's text layout
 * (2c.Z1.Z2.dat/.dir, 3c.Z1.Z2.Z3.dat/.dir, data files, TT_munu_*.dat) for
 * the species of fb_set_species (their nZ), loading the interactions of this
 * path through fb_set_table_2c / fb_set_table_3c.  Call fb_commit_tables
 * afterwards.  A missing or malformed file fails with FB_E_NO_TABLE. */
int fb_read_fdata(fb_ctx *ctx, const char *fdata_dir);

/* --- structures -----------------------------------------------------------
 * One call per structure (T_structure after read_positions / make_cells /
 * read_charges, c.SYSTEM/M_configuraciones.f90:604-922, make_cells.f90:51-157).
 * ratom [natoms][3] in Angstrom, already origin-shifted as read_positions
 * leaves them; imass 1-based species; lattice rows a1, a2, a3; dQ
 * [natoms][FB_MAXSH] shell charges (NULL: neutral).  Several structures form
 * one batch: lists and assembly then cover all of them in single launches. */
/* fb_read_positions: the in-scope steps of read_positions
 * (c.SYSTEM/M_configuraciones.f90:604-797) on a structure file in the
 * reference's layout ("natoms icluster"; three lattice vectors; one "Z x y z"
 * record per atom; list-directed): species by atomic number (the first of the
 * nspecies entries of nZ with that Z, 1-based into imass, :666-681); the
 * origin shift, every atom moved by (pi, 1/e, sqrt 2) when some atom lies
 * within 1e-4 A of the origin (:757-775, *shifted = 1 then); norbitals, the
 * sum of norb over the atoms (:703-708); iblock_slot, each atom's 0-based
 * orbital offset (:778-783, may be NULL).  Outputs are ready for
 * fb_add_structure (lattice rows a1, a2, a3; ratom [natoms][3]).  max_atoms
 * bounds ratom/imass/iblock_slot (FB_E_CAPACITY beyond, *natoms still set).
 * An atomic number missing from nZ fails with FB_E_ARG where the reference
 * stops.  Needs no context and no GPU; fb_last_error(NULL) gives the reason. */
int fb_read_positions(const char *basisfile, int nspecies, const int *nZ, const int *norb, int max_atoms,
                      int *natoms, int *icluster, double *lattice, double *ratom, int *imass, int *norbitals,
                      int *iblock_slot, int *shifted);
int fb_clear_structures(fb_ctx *ctx);
int fb_add_structure(fb_ctx *ctx, int natoms, const double *ratom, const int *imass,
                     const double *lattice, int icluster, const double *dQ, int *istruct);
int fb_set_positions(fb_ctx *ctx, int istruct, const double *ratom);   /* new geometry, same atoms */
int fb_set_charges(fb_ctx *ctx, int istruct, const double *dQ);      /* invalidates Vna only */

/* --- lists: driver_neighbors, c.SYSTEM/M_neighbors.f90:93-141 ------------
 * (find_neigh_max/find_neighbors :246-519, find_common_max/
 * find_common_neighbors :548-812), for every structure of the batch. */
int fb_build_lists(fb_ctx *ctx, int64_t *npairs_total, int64_t *ntriplets_total);
int fb_get_list_sizes(fb_ctx *ctx, int istruct, int64_t *npairs, int64_t *ntriplets, int *mbeta_max);
int fb_get_neighbors(fb_ctx *ctx, int istruct, int *neighn, int *neigh_j, int *neigh_b,
                     int *neigh_back, int *neigh_self);
int fb_get_common(fb_ctx *ctx, int istruct, int *ncommon, int *iatom_common_j, int *iatom_common_b,
                  int *jatom_common_j, int *jatom_common_b, int *neigh_common);

/* --- assembly ---------------------------------------------------------- */
int fb_assemble_S(fb_ctx *ctx);          /* j.ASSEMBLERS/M_assemble_2c_DOGS.f90:87-209   */
int fb_assemble_T(fb_ctx *ctx);          /* j.ASSEMBLERS/M_assemble_2c_DOGS.f90:232-353  */
int fb_assemble_dipole_z(fb_ctx *ctx);   /* j.ASSEMBLERS/M_assemble_2c_DOGS.f90:376-499  */
int fb_assemble_vna_2c(fb_ctx *ctx);     /* j.ASSEMBLERS/M_assemble_2c_DOGS.f90:523-829  */
int fb_assemble_vna_3c(fb_ctx *ctx);     /* j.ASSEMBLERS/M_assemble_3c_DOGS.f90:86-340   */
/* S, T, dipole_z and vna_2c in one 2c launch, then vna_3c: the H+S build.
 * Asynchronous: it returns once the build is enqueued (replayed from a
 * captured HIP graph unless timing is on), without waiting for an earlier
 * build; a kernel-detected error of the builds enqueued since the last check
 * is returned by the next fb_synchronize, fb_get_blocks or other call that
 * reads results or assembles separately. */
int fb_assemble_all(fb_ctx *ctx);

/* --- results ------------------------------------------------------------ */
int fb_get_block_len(fb_ctx *ctx, int istruct, int64_t *len);
int fb_get_blocks(fb_ctx *ctx, int istruct, int matrix, double *out);
int fb_get_device_blocks(fb_ctx *ctx, int matrix, void **dptr, int64_t *len);

/* --- forces: the k.DASSEMBLERS force assemblers --------------------------
 * Call order as in the reference's MD driver (n.NAMD/M_nonadiabatic_mdet.f90:
 * 2065-2101): the H+S assembly above, fb_set_denmat for every structure, then
 * fb_dassemble_S, _T, _dipole_z, _vna_2c, _vna_3c (or fb_dassemble_all).
 * The density matrix (s%denmat) comes in the block layout of fb_get_blocks;
 * a new list build discards it.  Outputs:
 *   Dblocks      Dblock(3, norb_mu, norb_nu) per pair, pair order, 3 x the
 *                block length (fb_get_dblocks; matrix FB_MAT_S/T/DIP)
 *   vna_atom     forces%vna_atom(3, ineigh) per pair, [npairs][3]
 *   vna_ontop    forces%vna_ontop(3, ineigh) per pair, [npairs][3]
 *   f3naa/b/c    forces%f3naa / f3nab / f3nac per atom, [natoms][3]
 * The force kernels compute in binary64; under FB_PREC_MIXED the 3c force
 * interpolation (values and derivatives of the table stencil) runs in FP32. */
int fb_set_denmat(fb_ctx *ctx, int istruct, const double *rho);  /* s%denmat(i)%neighbors(k)%block */
int fb_dassemble_S(fb_ctx *ctx);          /* k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:95-266   */
int fb_dassemble_T(fb_ctx *ctx);          /* k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:292-463  */
int fb_dassemble_dipole_z(fb_ctx *ctx);   /* k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:487-655  */
int fb_dassemble_vna_2c(fb_ctx *ctx);     /* k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:679-1199 */
int fb_dassemble_vna_3c(fb_ctx *ctx);     /* k.DASSEMBLERS/M_Dassemble_3c_DOGS.f90:136-793  */
int fb_dassemble_all(fb_ctx *ctx);        /* all five: one 2c launch, the 3c launch, per-atom sums */
int fb_get_dblocks(fb_ctx *ctx, int istruct, int matrix, double *out);
int fb_get_forces(fb_ctx *ctx, int istruct, double *vna_atom, double *vna_ontop, double *f3naa, double *f3nab,
                  double *f3nac);

/* --- McWEDA densities: j.ASSEMBLERS/M_assemble_rho_McWEDA_DOGS.f90 -------
 * The input density of the McWEDA exchange-correlation terms, assembled from
 * the same lists and frames as H and S (the triplet consumers of the 3c lists
 * besides vna_3c).  fb_set_qin gives the input shell charges
 * s%atom(:)%shell(:)%Qin, [natoms][FB_MAXSH], of one structure; every
 * structure of the batch needs them before the first density call.  Call order
 * as in assemble_vxc (j.ASSEMBLERS/M_assemble_vxc_McWEDA_DOGS.f90:148-154):
 * rho_2c, rho_3c, rho_weighted_2c, rho_weighted_3c (rho_3c adds into the
 * rho_in blocks of rho_2c, rho_weighted_3c into those of rho_weighted_2c).
 * Results through fb_get_blocks: FB_MAT_RHO_IN / FB_MAT_RHO_BOND in the block
 * layout (fb_get_block_len doubles), FB_MAT_RHOW_IN / FB_MAT_RHOW_BOND as
 * column-major nssh_mu x nssh_nu shell blocks in pair order
 * (fb_get_shell_block_len doubles).  Binary64, the reference's summation
 * order; a new list build discards them. */
int fb_set_qin(fb_ctx *ctx, int istruct, const double *Qin);
int fb_assemble_rho_2c(fb_ctx *ctx);            /* M_assemble_rho_McWEDA_DOGS.f90:90-341  */
int fb_assemble_rho_3c(fb_ctx *ctx);            /* M_assemble_rho_McWEDA_DOGS.f90:362-511 */
int fb_assemble_rho_weighted_2c(fb_ctx *ctx);   /* M_assemble_rho_McWEDA_DOGS.f90:534-762 */
int fb_assemble_rho_weighted_3c(fb_ctx *ctx);   /* M_assemble_rho_McWEDA_DOGS.f90:783-924 */
int fb_assemble_rho_all(fb_ctx *ctx);           /* the four in order: one 2c launch, two 3c launches */
int fb_get_shell_block_len(fb_ctx *ctx, int istruct, int64_t *len);

/* --- non-local pseudopotential: c.SYSTEM/M_neighbors_PP.f90,
 * j.ASSEMBLERS/M_assemble_PP_2c.f90, M_assemble_PP_3c.f90 ----------------------
 * fb_set_species_pp: the PP shells of every species (species(:)%nssh_PP,
 *   %shell_PP(:)%lssh, %rcutoff_PP as the neighbour code uses it, and the
 *   Kleinman-Bylander coefficients shell_PP(:)%cl of clPP.ZZ.dat); arrays
 *   [nspecies] and [nspecies][FB_MAXSH]; s and p PP shells.  Call it before
 *   fb_commit_tables: the P_vnl tables (interaction 6) are loaded only then.
 *   fb_read_fdata also reads clPP.ZZ.dat into these coefficients.
 * fb_build_lists_pp: driver_neighbors_PP (M_neighbors_PP.f90:98-151) for every
 *   structure: neighbors_PP (rc_i + rcutoff_PP_j), neighbors_PPx (rcutoff_PP_i
 *   + rc_j), neighbors_PPp with the maps and points, and the PP common
 *   neighbours (triplets), all in the reference's order.
 * fb_get_pp_list_sizes: sizes[6] = PP pairs, PPx pairs, PPp pairs, PP triplets,
 *   sVNL length, vnl length (doubles) of one structure.
 * fb_get_pp_lists: which 0 neighbors_PP (n = neighn, a0 neigh_j, a1 neigh_b,
 *   a2 map, self = neighbors_PP_self), 1 neighbors_PPx (a3 point as well),
 *   2 neighbors_PPp (a0, a1, self), 3 common neighbours (n = ncommon, a0..a4 =
 *   iatom_common_j, iatom_common_b, jatom_common_j, jatom_common_b,
 *   neigh_common); 1-based atom and neighbour numbers as in the reference;
 *   unused outputs may be NULL.
 * fb_assemble_svnl / _vnl_2c / _vnl_3c: in this order.  Blocks through
 *   fb_get_pp_blocks: which 0 s%svnl (per PP pair, column-major norb_mu x
 *   norb_PP_nu), 1 s%vnl (per PPp pair, norb_mu x norb_nu); binary64 with the
 *   reference's summation order.  A geometry change discards the PP lists. */
int fb_set_species_pp(fb_ctx *ctx, const int *nssh_pp, const int *lssh_pp, const double *rcutoff_pp,
                      const double *cl);
int fb_build_lists_pp(fb_ctx *ctx, int64_t *npp_total, int64_t *nppp_total, int64_t *ntriplets_total);
int fb_get_pp_list_sizes(fb_ctx *ctx, int istruct, int64_t *sizes);
int fb_get_pp_lists(fb_ctx *ctx, int istruct, int which, int *n, int *a0, int *a1, int *a2, int *a3, int *a4,
                    int *self);
int fb_assemble_svnl(fb_ctx *ctx);      /* j.ASSEMBLERS/M_assemble_PP_2c.f90:95-207   */
int fb_assemble_vnl_2c(fb_ctx *ctx);    /* j.ASSEMBLERS/M_assemble_PP_2c.f90:230-496  */
int fb_assemble_vnl_3c(fb_ctx *ctx);    /* j.ASSEMBLERS/M_assemble_PP_3c.f90:94-237   */
int fb_get_pp_blocks(fb_ctx *ctx, int istruct, int which, double *out);

/* --- consumer: Gamma-point eigenvalues ------------------------------------
 * Folds the blocks of structure istruct at k = 0 into a dense binary64 matrix
 * (build_Hmatrix, h.SOLVESH/M_kspace_Lowdin.f90:180-271, 630-724) -- which = 0:
 * S, which = 1: H = (T + Vna block) + Vna blocko, the in-scope terms -- and
 * returns its eigenvalues in ascending order from rocSOLVER dsyevd on the
 * upper triangle, the triangle dsyev('V','U') reads in the reference
 * (M_diagonalization_Lowdin_gamma.f90:173-178).  *n_out receives the orbital
 * count; w must hold that many values.  rocSOLVER is loaded on first use
 * (FB_E_UNSUPPORTED when absent).  The n x n matrix is allocated for the call
 * and freed before it returns. */
int fb_gamma_eigenvalues(fb_ctx *ctx, int istruct, int which, double *w, int64_t nw, int64_t *n_out);

/* --- learned 3c surrogate ------------------------------------------------
 * No counterpart in the reference.  One network per neutral 3c group (species
 * triple): inputs (x, y, cos theta), two hidden layers of 64 units with the
 * activation z s(z) (s a quadratic sigmoid, the shape of SiLU without a
 * transcendental), outputs the five Legendre theta cells of every matrix
 * element that getMEs_Fdata_3c interpolates (e.FDATA/M_Fdata_3c.f90:391-489).
 * fb_mlp_train samples nsamples points of every neutral group from the
 * binary64 table path, trains in FP32 on the GPU (Adam, batch, nsteps, cosine
 * learning-rate decay from lr), stores BF16 MFMA operand fragments (inputs,
 * first-layer and last-layer weights split into BF16 high and low parts),
 * and calibrates, for every sub-cell of a grid four times finer than the
 * table's in each direction, the largest Frobenius norm of the block error
 * |M_network - M_table| (theta cells folded at the sampled cos theta) over
 * fresh samples; orthogonal rotations keep that norm.  With
 * fb_set_3c_method(FB_3C_MLP) and mixed precision, the 3c launch evaluates
 * the network on MFMA for each triplet whose bound
 *   2 * P_eq2 * (calibrated Frobenius error of the sub-cell's molecular-frame block)
 * is at most the tolerance (fb_mlp_set_tolerance, eV, default 2.5e-5: with it
 * the measured H-block error stays below 1e-4 eV on the test workloads) and
 * the table stencil for every other triplet.  Charged builds use the tables.
 * fb_mlp_info: 12 values per group: is, js, ks, isub, trained, nsamples,
 *   nsteps, loss before, loss after (normalised MSE), max sub-cell error (eV,
 *   Frobenius),
 *   training ms, calibration samples.
 * fb_mlp_stats: triplets that took the network and triplets seen by MLP 3c
 *   launches since the previous call.
 * fb_mlp_groups: number of 3c groups (network slots).
 * fb_mlp_samples: n training samples of one group as fb_mlp_train draws them
 *   (X [n][3] normalised inputs, T [n][64] theta cells per output slot).
 * fb_mlp_gradients / fb_mlp_forward run the training kernels (FP32 forward and
 *   backward of one batch, mean squared error over the valid output slots) and
 *   the BF16 MFMA inference on caller-given parameters (layout: W1[64][3],
 *   b1[64], W2[64][64], b2[64], W3[64][64], b3[64]; X [n][3], T and Y [n][64]). */
int fb_mlp_train(fb_ctx *ctx, int64_t nsamples, int nsteps, int batch, double lr, uint64_t seed);
int fb_mlp_groups(fb_ctx *ctx, int *ngroups);
int fb_mlp_samples(fb_ctx *ctx, int group, int64_t n, uint64_t seed, float *X, float *T);
int fb_mlp_set_tolerance(fb_ctx *ctx, double tau_ev);
int fb_mlp_info(fb_ctx *ctx, int group, double *out, int nout);
int fb_mlp_get_params(fb_ctx *ctx, int group, float *params, float *oscale);
int fb_mlp_stats(fb_ctx *ctx, int64_t *n_network, int64_t *n_seen);
int fb_mlp_gradients(fb_ctx *ctx, int nme, const float *params, int n, const float *X, const float *T,
                     float *grads, double *loss);
int fb_mlp_forward(fb_ctx *ctx, int nme, const float *params, const float *oscale, int n, const float *X,
                   float *Y);

/* --- allocation-free training step -----------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, training loops that forbid
 * hipMalloc.  The forward and backward pass of fb_mlp_gradients on one batch
 * with every argument in device memory: params [8576] (fb_mlp_gradients
 * layout), X [n][3] and T [n][64] in; grads [8576] and loss (one float: the
 * mean squared error over the valid output slots) out.  The work arrays
 * (activations, their gradients, per-row losses, split-K partial products)
 * live in a fixed scratch array of the library, about 5.8 MB per GPU, held
 * until the process ends; the call allocates nothing, copies nothing to or
 * from the host and does not synchronize (work is enqueued on the context's
 * stream).  n is at most FB_MLP_NOALLOC_MAX_BATCH (FB_E_CAPACITY beyond).
 * Every call on one GPU shares that GPU's scratch array, from any context:
 * calls that overlap in time on one GPU race, so order them on one stream or
 * synchronize between them.  The gradients equal those of fb_mlp_gradients
 * on the same batch. */
#define FB_MLP_NOALLOC_MAX_BATCH 2048
int fb_mlp_gradients_noalloc(fb_ctx *ctx, int nme, const float *params, int n, const float *X, const float *T,
                             float *grads, float *loss);

/* --- tuning ------------------------------------------------------------
 * No counterpart in the reference (its assembler has no launch parameters).
 * FB_TUNE_CHUNK3C: triplets per workgroup chunk of the mixed 3c kernel (64,
 * 128 or 256), which sets its LDS footprint.  FB_TUNE_2C_KERNEL: the 2c
 * kernel for s/p species, 0 = table cells staged in LDS, 1 (default) =
 * register-resident with per-lane table gathers; both give the same bits.
 * FB_TUNE_GRAPH: 1 (default) replays fb_assemble_all from a captured HIP
 * graph, 0 enqueues its launches directly.  FB_TUNE_FORCE_KERNEL: the force
 * kernels for s/p species, 1 (default) = compile-time block shapes, 0 = the
 * generic shell loops; both give the same values.  The McWEDA density and
 * pseudopotential kernels follow the same switch.
 * FB_TUNE_FUSED: 0 runs the neutral mixed-precision H+S build as two
 * launches (2c, then 3c, the 2c one on a second stream), 3 or 4 as one
 * launch whose first blocks are the 2c atom groups and the rest the
 * 64-triplet 3c tiles, four per workgroup (register allocation for 3 or 4
 * waves per SIMD).  FB_TUNE_PERSIST: 0 (default) gives that launch one
 * workgroup per four tiles; 1 makes its 3c part persistent, waves x CUs
 * workgroups whose wavefronts take the tiles one at a time from per-XCD
 * queues (same blocks, same bits);
 * charged, binary64 and learned-surrogate builds always take separate
 * launches.
 * fb_autotune times whole H+S builds as fb_assemble_all runs them (graph
 * replays unless FB_TUNE_GRAPH is 0; nrep each, two sweeps, HIP events) over every
 * launch shape that applies to the context -- the fused launch at both
 * occupancies, with and without the queue, and the separate launches with every chunk and both 2c
 * kernels -- and keeps the fastest; when that is the one launch, the chunk
 * and 2c kernel of the separate launches are those of the fastest separate
 * shape.  It returns the chosen chunk in *chosen.
 * The blocks hold a complete build afterwards. */
enum fb_tune { FB_TUNE_CHUNK3C = 0, FB_TUNE_2C_KERNEL = 1, FB_TUNE_GRAPH = 2, FB_TUNE_FORCE_KERNEL = 3, FB_TUNE_FUSED = 4,
              FB_TUNE_PERSIST = 5 };
int fb_set_tuning(fb_ctx *ctx, int param, int value);
int fb_get_tuning(fb_ctx *ctx, int param, int *value);
int fb_autotune(fb_ctx *ctx, int nrep, int *chosen);

/* --- measurement -------------------------------------------------------
 * With timing on, every kernel launch of the library is bracketed by HIP
 * events on the launch stream; fb_get_timings synchronizes and returns the
 * kernels launched since the previous fb_get_timings (name, ms), at most max
 * entries; names is max*32 chars. */
int fb_enable_timing(fb_ctx *ctx, int on);
int fb_get_timings(fb_ctx *ctx, int max, char *names, float *ms, int *count);

#ifdef __cplusplus
}
#endif
#endif
