/* CPU oracle for the FIREBALL 2c/3c H+S assembly path -- TEST INFRASTRUCTURE.
 *
 * A plain-C, binary64 restatement of the reference's algorithm, written to be
 * compared against (tests, smoke check, CPU baseline).  It is never linked
 * into, called by, or used as a fallback of the GPU library.
 *
 * Conventions follow the reference: atoms, species and neighbour indices are
 * 1-based in every array that mirrors a reference array (neigh_j, neigh_back,
 * neigh_self, iatom_common_j, neigh_common, ...); image indices mbeta run
 * 0..mbeta_max.  Blocks are column-major norb_mu x norb_nu, packed per pair in
 * CSR pair order (iatom major, ineigh minor).
 */
#ifndef FB_ORACLE_H
#define FB_ORACLE_H

#include <stdint.h>

#define FBO_MAXSH 4
#define FBO_MAXTHETA 5

enum {
  FBO_OK = 0,
  FBO_E_ARG = -1,
  FBO_E_X_NEGATIVE = -2,        /* getMEs_Fdata_2c: x < 0 */
  FBO_E_X_OUT_OF_RANGE = -3,    /* getMEs_Fdata_2c: x > xmax*(1+1e-5) */
  FBO_E_ATOMS_TOO_CLOSE = -4,   /* find_neighbors: distance < 0.5 A */
  FBO_E_NEIGH_SELF = -5,        /* neigh_self duplicate / missing */
  FBO_E_NEIGH_BACK = -6,        /* neigh_back duplicate / missing */
  FBO_E_NO_TABLE = -7,          /* interaction cell not loaded */
  FBO_E_ORDER = -8              /* assemble called out of order */
};

enum { FBO_MAT_S = 0, FBO_MAT_T = 1, FBO_MAT_DIP = 2, FBO_MAT_VNA = 3, FBO_MAT_VNAO = 4 };

typedef struct fbo_tables fbo_tables;
typedef struct fbo_system fbo_system;

fbo_tables *fbo_tables_new(int nspecies, const int *nssh, const int *lssh, const double *rcutoffA);
void fbo_tables_free(fbo_tables *t);
int fbo_add_cell2c(fbo_tables *t, int is, int js, int itype, int isub, int nME, int nx, double xmax,
                   const int *mu, const int *nu, const int *mvalue, const double *data);
int fbo_add_cell3c(fbo_tables *t, int is, int js, int ks, int itype, int isub, int itheta, int nME,
                   int nx, int ny, double xmax, double ymax, const int *mu, const int *nu,
                   const int *mvalue, const double *data);

/* primitives */
double fbo_distance(const double *a, const double *b);
double fbo_smoother(double r, double rend, double x);
void fbo_epsilon(const double *r1, const double *r2, double *spe);
void fbo_twister(const double *eps, int have_d, double *dmat, double *pmat);
void fbo_rotate(const fbo_tables *t, int in1, int in3, const double *eps, int norb_mu, int norb_nu,
                const double *mmatrix, double *xmatrix);
int fbo_getMEs_2c(const fbo_tables *t, int is, int js, int iint, int isub, double x, int norb_mu,
                  int norb_nu, double *hmbox);
int fbo_getMEs_3c(const fbo_tables *t, int is, int js, int ks, int iint, int isub, double x,
                  double y, int norb_mu, int norb_nu, double cost, double *hmbox);
int fbo_make_cells(const double *lattice, int icluster, double *xl_out);

/* structure, lists, assembly */
fbo_system *fbo_system_new(const fbo_tables *t, int natoms, const double *ratom, const int *imass,
                           const double *lattice, int icluster, const double *dQ);
void fbo_system_free(fbo_system *s);
int fbo_build_lists(fbo_system *s);
int64_t fbo_npairs(const fbo_system *s);
int64_t fbo_ntriplets(const fbo_system *s);
int fbo_get_lists(const fbo_system *s, int *neighn, int *neigh_j, int *neigh_b, int *neigh_back,
                  int *neigh_self, int *ncommon, int *iatom_j, int *iatom_b, int *jatom_j,
                  int *jatom_b, int *mneigh);
int64_t fbo_block_len(const fbo_system *s);
int fbo_assemble_S(fbo_system *s);
int fbo_assemble_T(fbo_system *s);
int fbo_assemble_dipole_z(fbo_system *s);
int fbo_assemble_vna_2c(fbo_system *s);
int fbo_assemble_vna_3c(fbo_system *s);
int fbo_get_blocks(const fbo_system *s, int matrix, double *out);
int fbo_frame_fallbacks(const fbo_system *s);

/* forces (k.DASSEMBLERS): derivative primitives.  deps layouts are Fortran
 * column-major deps(ix, imu, k); Dblocks are Dblock(ix, imu, inu) per pair. */
double fbo_Dsmoother(double r, double rend, double x);
int fbo_Depsilon_2c(const double *r1, const double *r2, double z, const double *eps, double *deps);
int fbo_Depsilon_3c(const double *r1, const double *r2, const double *r21, double z, const double *ratm,
                    const double *rnabc, const double *eps3, double *dera3, double *der13);
void fbo_Dtwister(const double *eps, const double *deps, double *Ddmat, double *Dpmat);
void fbo_Drotate(const fbo_tables *t, int in1, int in2, const double *eps, const double *deps, int norb_mu,
                 int norb_nu, const double *matm, const double *Dmatm, double *Dmatx);
int fbo_getDMEs_2c(const fbo_tables *t, int is, int js, int iint, int isub, double x, int norb_mu,
                   int norb_nu, double *hmbox, double *Dhmbox);
int fbo_getDMEs_3c(const fbo_tables *t, int is, int js, int ks, int iint, int isub, double x, double y,
                   int norb_mu, int norb_nu, double cost, double *hmbox, double *Dphmbox, double *Dxhmbox,
                   double *Dyhmbox);
/* forces: density matrix in, Dblocks and force pieces out */
int fbo_set_denmat(fbo_system *s, const double *rho);
int fbo_Dassemble_S(fbo_system *s);
int fbo_Dassemble_T(fbo_system *s);
int fbo_Dassemble_dipole_z(fbo_system *s);
int fbo_Dassemble_vna_2c(fbo_system *s);
int fbo_Dassemble_vna_3c(fbo_system *s);
int fbo_get_dblocks(const fbo_system *s, int which, double *out);
int fbo_get_forces(const fbo_system *s, double *vna_atom, double *vna_ontop, double *f3naa, double *f3nab,
                   double *f3nac);
int fbo_deps_warnings(const fbo_system *s);

/* McWEDA densities (j.ASSEMBLERS/M_assemble_rho_McWEDA_DOGS.f90).  qin is
 * [natoms][FBO_MAXSH] (s%atom(:)%shell(:)%Qin).  Orbital blocks (rho_in,
 * rho_bond) use the block layout; shell blocks (rho_in_weighted,
 * rho_bond_weighted) are column-major nssh_mu x nssh_nu per pair, pair order. */
int fbo_set_qin(fbo_system *s, const double *qin);
int fbo_assemble_rho_2c(fbo_system *s);
int fbo_assemble_rho_3c(fbo_system *s);
int fbo_assemble_rho_weighted_2c(fbo_system *s);
int fbo_assemble_rho_weighted_3c(fbo_system *s);
int64_t fbo_shell_block_len(const fbo_system *s);
int fbo_get_rho(const fbo_system *s, int which, double *out);

/* pseudopotential (c.SYSTEM/M_neighbors_PP.f90, j.ASSEMBLERS/M_assemble_PP_2c.f90,
 * M_assemble_PP_3c.f90).  Per species: nssh_PP, lssh_PP[FBO_MAXSH], rcutoff_PP
 * (as read from info.dat), cl[FBO_MAXSH] (clPP.ZZ.dat). */
int fbo_tables_set_pp(fbo_tables *t, const int *nssh_pp, const int *lssh_pp, const double *rcutoff_pp,
                      const double *cl);
void fbo_rotate_PP(const fbo_tables *t, int in1, int in2, const double *eps, int norb_mu, int norb_nu,
                   const double *mmatrix, double *xmatrix);
int fbo_build_pp_lists(fbo_system *s);
int fbo_assemble_svnl(fbo_system *s);
int fbo_assemble_vnl_2c(fbo_system *s);
int fbo_assemble_vnl_3c(fbo_system *s);
void fbo_pp_sizes(const fbo_system *s, int64_t *out);
int fbo_get_pp_list(const fbo_system *s, int which, int *n, int *a0, int *a1, int *a2, int *a3, int *a4,
                    int *self);
int fbo_get_pp_blocks(const fbo_system *s, int which, double *out);

#endif
