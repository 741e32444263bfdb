/* CPU oracle for the FIREBALL 2c/3c assembly path -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain C, binary64, compiled with -ffp-contract=off.  Each routine restates
 * the reference routine named in its comment (paths relative to the
 * reference's src/ tree), keeping the reference's evaluation order so that the
 * results agree with the reference's Intel/flang FP64 build to rounding.
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py load this.
 */
#include "fb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define P_EQ2 14.39975            /* include/constants.h:4 */
#define P_OVERLAP 1               /* include/interactions_2c.h:5-14 */
#define P_KINETIC 2
#define P_VNA_ONTOPL 3
#define P_VNA_ONTOPR 4
#define P_VNA_ATOM 5
#define P_DIPOLE_Z 10
#define P_BCNA 3                  /* include/interactions_3c.h:8 */
#define MAXTYPE2 26
#define MAXSUB2 10
#define MAXTYPE3 5
#define MAXSUB3 7
#define P_MBOX 4                  /* c.SYSTEM/M_configuraciones.f90:359 */

typedef struct {
  int nME, nx;
  double xmax, dx;
  int *mu, *nu, *mv;
  double *data; /* [nx][nME] */
} cell2c;

typedef struct {
  int nME, nx, ny;
  double xmax, ymax;
  int *mu, *nu, *mv;
  double *data; /* [ny][nx][nME] */
} cell3c;

struct fbo_tables {
  int ns;
  int nssh[16];
  int lssh[16][FBO_MAXSH];
  double rcA[16][FBO_MAXSH];
  int norb[16];
  int have_d;
  cell2c **c2; /* [ns][ns][MAXTYPE2][MAXSUB2] */
  /* pseudopotential shells (species(:)%shell_PP, %rcutoff_PP; shell_PP(:)%cl from clPP.ZZ.dat) */
  int nssh_pp[16], lssh_pp[16][FBO_MAXSH], norb_pp[16];
  double rc_pp[16], cl_pp[16][FBO_MAXSH];
  cell3c **c3; /* [ns][ns][ns][MAXTYPE3][MAXSUB3][FBO_MAXTHETA+1] */
};

struct fbo_system {
  const fbo_tables *t;
  int natoms, icluster, mbeta_max;
  double *ratom; /* [N][3] */
  int *imass;    /* 1-based species */
  double lattice[9];
  double *dQ;    /* [N][FBO_MAXSH] */
  double *xl;    /* [mbeta_max+1][3] */
  /* neighbour lists (CSR over iatom) */
  int *neighn, *noff, *neigh_j, *neigh_b, *neigh_back, *neigh_self;
  int64_t npairs;
  /* common-neighbour lists (CSR over ialpha) */
  int *ncommon, *coff, *ci_j, *ci_b, *cj_j, *cj_b, *cm;
  int64_t ntrip;
  /* blocks */
  int64_t *boff; /* per pair */
  int64_t blen;
  double *mat[5];
  int done[5];
  int nfallback;
  /* forces: density matrix input and derivative outputs */
  double *rho;      /* [blen] per-pair density blocks, block layout */
  double *dmat[3];  /* [3*blen] Dblock(3, norb_mu, norb_nu) of S, T, dipole_z */
  double *fatom;    /* [npairs][3] forces%vna_atom(:, ineigh) */
  double *fontop;   /* [npairs][3] forces%vna_ontop(:, ineigh) */
  double *f3[3];    /* [natoms][3] forces%f3naa, f3nab, f3nac */
  int ddone[5];
  int ndwarn;       /* Depsilon_2c/3c colinear guards taken (WARNINGS file lines in the reference) */
  /* McWEDA densities: input shell charges and the four density block arrays */
  double *qin;      /* [N][FBO_MAXSH] s%atom(:)%shell(:)%Qin */
  double *rho2[2];  /* [blen] rho_in, rho_bond: orbital blocks, block layout */
  int64_t *wboff;   /* per pair: offset of the nssh_mu x nssh_nu shell block */
  int64_t wlen;
  double *wrho[2];  /* [wlen] rho_in_weighted, rho_bond_weighted */
  int rdone[4];     /* rho_2c, rho_3c, rho_weighted_2c, rho_weighted_3c */
  /* pseudopotential lists (driver_neighbors_PP) and blocks */
  int *pp_n, *pp_off, *pp_j, *pp_b, *pp_map, *pp_self;
  int *px_n, *px_off, *px_j, *px_b, *px_map, *px_point, *px_self;
  int *pq_n, *pq_off, *pq_j, *pq_b, *pq_self;          /* neighbors_PPp */
  int *pc_n, *pc_off, *pc_ij, *pc_ib, *pc_jj, *pc_jb, *pc_m;   /* PP common neighbours */
  int64_t *sv_off, sv_len; double *svnl;              /* per PP pair: norb_mu x norb_PP_nu */
  int64_t *vl_off, vl_len; double *vnl;               /* per PPp pair: norb_mu x norb_nu */
  int ppdone[3];    /* svnl, vnl_2c, vnl_3c */
};

static inline cell2c *C2(const fbo_tables *t, int is, int js, int it, int isub) {
  if (is < 1 || js < 1 || is > t->ns || js > t->ns || it < 0 || it >= MAXTYPE2 || isub < 0 ||
      isub >= MAXSUB2)
    return NULL;
  return t->c2[(((size_t)(is - 1) * t->ns + (js - 1)) * MAXTYPE2 + it) * MAXSUB2 + isub];
}

static inline cell3c *C3(const fbo_tables *t, int is, int js, int ks, int it, int isub, int ith) {
  if (it < 0 || it >= MAXTYPE3 || isub < 0 || isub >= MAXSUB3 || ith < 1 || ith > FBO_MAXTHETA)
    return NULL;
  size_t k = ((((size_t)(is - 1) * t->ns + (js - 1)) * t->ns + (ks - 1)) * MAXTYPE3 + it);
  return t->c3[(k * MAXSUB3 + isub) * (FBO_MAXTHETA + 1) + ith];
}

fbo_tables *fbo_tables_new(int nspecies, const int *nssh, const int *lssh, const double *rcutoffA) {
  if (nspecies < 1 || nspecies > 16) return NULL;
  fbo_tables *t = calloc(1, sizeof(*t));
  t->ns = nspecies;
  for (int i = 0; i < nspecies; i++) {
    t->nssh[i] = nssh[i];
    t->norb[i] = 0;
    for (int k = 0; k < nssh[i]; k++) {
      t->lssh[i][k] = lssh[i * FBO_MAXSH + k];
      t->rcA[i][k] = rcutoffA[i * FBO_MAXSH + k];
      t->norb[i] += 2 * t->lssh[i][k] + 1;
      if (t->lssh[i][k] == 2) t->have_d = 1; /* c.SYSTEM/M_species.f90:1212-1217 */
    }
  }
  size_t n2 = (size_t)nspecies * nspecies * MAXTYPE2 * MAXSUB2;
  size_t n3 = (size_t)nspecies * nspecies * nspecies * MAXTYPE3 * MAXSUB3 * (FBO_MAXTHETA + 1);
  t->c2 = calloc(n2, sizeof(cell2c *));
  t->c3 = calloc(n3, sizeof(cell3c *));
  return t;
}

void fbo_tables_free(fbo_tables *t) {
  if (!t) return;
  size_t n2 = (size_t)t->ns * t->ns * MAXTYPE2 * MAXSUB2;
  size_t n3 = (size_t)t->ns * t->ns * t->ns * MAXTYPE3 * MAXSUB3 * (FBO_MAXTHETA + 1);
  for (size_t i = 0; i < n2; i++)
    if (t->c2[i]) { free(t->c2[i]->mu); free(t->c2[i]->nu); free(t->c2[i]->mv); free(t->c2[i]->data); free(t->c2[i]); }
  for (size_t i = 0; i < n3; i++)
    if (t->c3[i]) { free(t->c3[i]->mu); free(t->c3[i]->nu); free(t->c3[i]->mv); free(t->c3[i]->data); free(t->c3[i]); }
  free(t->c2); free(t->c3); free(t);
}

static int *dupi(const int *a, int n) { int *b = malloc(sizeof(int) * (n ? n : 1)); memcpy(b, a, sizeof(int) * n); return b; }

int fbo_add_cell2c(fbo_tables *t, int is, int js, int itype, int isub, int nME, int nx, double xmax,
                   const int *mu, const int *nu, const int *mvalue, const double *data) {
  if (!t || is < 1 || js < 1 || is > t->ns || js > t->ns || itype < 0 || itype >= MAXTYPE2 ||
      isub < 0 || isub >= MAXSUB2 || nx < 7 || nME < 1)
    return FBO_E_ARG;
  cell2c *c = calloc(1, sizeof(*c));
  c->nME = nME; c->nx = nx; c->xmax = xmax;
  c->dx = xmax / (double)(nx - 1); /* e.FDATA/M_Fdata_2c.f90:222 */
  c->mu = dupi(mu, nME); c->nu = dupi(nu, nME); c->mv = dupi(mvalue, nME);
  c->data = malloc(sizeof(double) * nx * nME);
  memcpy(c->data, data, sizeof(double) * nx * nME);
  t->c2[(((size_t)(is - 1) * t->ns + (js - 1)) * MAXTYPE2 + itype) * MAXSUB2 + isub] = c;
  return FBO_OK;
}

int fbo_add_cell3c(fbo_tables *t, int is, int js, int ks, int itype, int isub, int itheta, int nME,
                   int nx, int ny, double xmax, double ymax, const int *mu, const int *nu,
                   const int *mvalue, const double *data) {
  if (!t || is < 1 || js < 1 || ks < 1 || is > t->ns || js > t->ns || ks > t->ns || itype < 0 ||
      itype >= MAXTYPE3 || isub < 0 || isub >= MAXSUB3 || itheta < 1 || itheta > FBO_MAXTHETA ||
      nx < 4 || ny < 4)
    return FBO_E_ARG;
  cell3c *c = calloc(1, sizeof(*c));
  c->nME = nME; c->nx = nx; c->ny = ny; c->xmax = xmax; c->ymax = ymax;
  c->mu = dupi(mu, nME); c->nu = dupi(nu, nME); c->mv = dupi(mvalue, nME);
  c->data = malloc(sizeof(double) * nx * ny * nME);
  memcpy(c->data, data, sizeof(double) * nx * ny * nME);
  size_t k = ((((size_t)(is - 1) * t->ns + (js - 1)) * t->ns + (ks - 1)) * MAXTYPE3 + itype);
  t->c3[(k * MAXSUB3 + isub) * (FBO_MAXTHETA + 1) + itheta] = c;
  return FBO_OK;
}

/* ---------------------------------------------------------------- helpers */

/* b.FUNCTIONS/distance.f90:72-73 */
double fbo_distance(const double *a, const double *b) {
  double d1 = b[0] - a[0], d2 = b[1] - a[1], d3 = b[2] - a[2];
  return sqrt(d1 * d1 + d2 * d2 + d3 * d3);
}

/* b.FUNCTIONS/magnitude.f90: sqrt(dot_product(a, a)) */
static double magnitude(const double *a) { return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]); }

/* b.FUNCTIONS/a_cross_b.f90 */
static void cross(const double *a, const double *b, double *c) {
  c[0] = a[1] * b[2] - a[2] * b[1];
  c[1] = a[2] * b[0] - a[0] * b[2];
  c[2] = a[0] * b[1] - a[1] * b[0];
}

/* b.FUNCTIONS/smoother.f90:124-136 */
double fbo_smoother(double r, double rend, double x) {
  double rbegin = x * rend;
  if (r > rend) return 0.0;
  if (r < rbegin) return 1.0;
  double ratio = (r - rbegin) / (rend - rbegin);
  double sm = 1.0 - ratio * ratio;
  return sm * sm;
}

/* b.FUNCTIONS/epsilon.f90:73-177.  spe is column-major 3x3: spe[3*j + i] = spe(i+1, j+1). */
static int epsilon_impl(const double *r1, const double *r2, double *spe) {
  double r1_mag = magnitude(r1), r2_mag = magnitude(r2);
  double zp[3], yp[3] = {0, 0, 0}, xp[3], r1h[3];
  int fallback = 0;
  if (r2_mag < 1.0e-4) { /* :123-133 */
    for (int i = 0; i < 9; i++) spe[i] = 0.0;
    spe[0] = spe[4] = spe[8] = 1.0;
    return 1;
  }
  for (int i = 0; i < 3; i++) zp[i] = r2[i] / r2_mag;
  double yp_mag = 0.0;
  if (r1_mag > 1.0e-4) {
    for (int i = 0; i < 3; i++) r1h[i] = r1[i] / r1_mag;
    cross(zp, r1h, yp);
    yp_mag = magnitude(yp);
  }
  if (yp_mag < 1.0e-4) { /* :148-168 */
    fallback = 1;
    if (fabs(zp[0]) > 1.0e-4) {
      yp[0] = -(zp[1] + zp[2]) / zp[0]; yp[1] = 1.0; yp[2] = 1.0;
    } else if (fabs(zp[1]) > 1.0e-4) {
      yp[0] = 1.0; yp[1] = -(zp[0] + zp[2]) / zp[1]; yp[2] = 1.0;
    } else {
      yp[0] = 1.0; yp[1] = 1.0; yp[2] = -(zp[0] + zp[1]) / zp[2];
    }
    yp_mag = magnitude(yp);
  }
  for (int i = 0; i < 3; i++) yp[i] = yp[i] / yp_mag;
  cross(yp, zp, xp);
  for (int i = 0; i < 3; i++) { spe[i] = xp[i]; spe[3 + i] = yp[i]; spe[6 + i] = zp[i]; }
  return fallback;
}

void fbo_epsilon(const double *r1, const double *r2, double *spe) { (void)epsilon_impl(r1, r2, spe); }

/* The d-orbital rotation tensor amat(3,3,5) of include/constants.h:26-35 written
 * out by its non-zero entries: amat(ix,jx,imu) with 1-based ix, jx. */
static double amat(int ix, int jx, int imu) {
  const double h = 0.5, c = 1.0 / (2.0 * sqrt(3.0)), c2 = 2.0 / (2.0 * sqrt(3.0));
  switch (imu) {
    case 1: return ((ix == 1 && jx == 2) || (ix == 2 && jx == 1)) ? h : 0.0;
    case 2: return ((ix == 2 && jx == 3) || (ix == 3 && jx == 2)) ? h : 0.0;
    case 3: return (ix == jx) ? (ix == 3 ? c2 : -c) : 0.0;
    case 4: return ((ix == 1 && jx == 3) || (ix == 3 && jx == 1)) ? h : 0.0;
    case 5: return (ix == jx && ix < 3) ? (ix == 1 ? h : -h) : 0.0;
  }
  return 0.0;
}

/* c.SYSTEM/M_rotations.f90:221-313.  eps, pmat column-major 3x3; dmat column-major 5x5. */
void fbo_twister(const double *eps, int have_d, double *dmat, double *pmat) {
#define E(i, j) eps[3 * ((j) - 1) + (i) - 1]
#define PM(i, j) pmat[3 * ((j) - 1) + (i) - 1]
#define DM(i, j) dmat[5 * ((j) - 1) + (i) - 1]
  PM(1, 1) = E(2, 2); PM(1, 2) = E(2, 3); PM(1, 3) = E(2, 1);
  PM(2, 1) = E(3, 2); PM(2, 2) = E(3, 3); PM(2, 3) = E(3, 1);
  PM(3, 1) = E(1, 2); PM(3, 2) = E(1, 3); PM(3, 3) = E(1, 1);
  if (!have_d) return;
  for (int imu = 1; imu <= 5; imu++) {
    double l11 = 0, l12 = 0, l13 = 0, l32 = 0, l33 = 0;
    for (int jx = 1; jx <= 3; jx++)
      for (int ix = 1; ix <= 3; ix++) {
        double a = amat(ix, jx, imu);
        if (a != 0.0) {
          l11 = l11 + a * E(jx, 1) * E(ix, 1);
          l12 = l12 + a * E(jx, 1) * E(ix, 2);
          l13 = l13 + a * E(jx, 1) * E(ix, 3);
          l32 = l32 + a * E(jx, 3) * E(ix, 2);
          l33 = l33 + a * E(jx, 3) * E(ix, 3);
        }
      }
    DM(imu, 1) = 2.0 * l12;
    DM(imu, 2) = 2.0 * l32;
    DM(imu, 3) = sqrt(3.0) * l33;
    DM(imu, 4) = 2.0 * l13;
    DM(imu, 5) = 2.0 * l11 + l33;
  }
#undef E
#undef PM
#undef DM
}

/* c.SYSTEM/M_rotations.f90:341-382: 5x5 column-major choice by l */
static void chooser(int l, const double *dmat, const double *pmat, double *r) {
  for (int i = 0; i < 25; i++) r[i] = 0.0;
  if (l == 0) r[0] = 1.0;
  else if (l == 1) { for (int j = 0; j < 3; j++) for (int i = 0; i < 3; i++) r[5 * j + i] = pmat[3 * j + i]; }
  else if (l == 2) { for (int i = 0; i < 25; i++) r[i] = dmat[i]; }
}

/* c.SYSTEM/M_rotations.f90:98-177: X = L M R^T per shell pair, loop order m2, m1, k2, k1. */
void fbo_rotate(const fbo_tables *t, int in1, int in3, const double *eps, int norb_mu, int norb_nu,
                const double *M, double *X) {
  double dmat[25] = {0}, pmat[9], left[25], right[25];
  fbo_twister(eps, t->have_d, dmat, pmat);
  for (int i = 0; i < norb_mu * norb_nu; i++) X[i] = 0.0;
  int n1 = 0;
  for (int issh = 0; issh < t->nssh[in1 - 1]; issh++) {
    int l1 = t->lssh[in1 - 1][issh];
    chooser(l1, dmat, pmat, left);
    int n2 = 0;
    for (int jssh = 0; jssh < t->nssh[in3 - 1]; jssh++) {
      int l2 = t->lssh[in3 - 1][jssh];
      chooser(l2, dmat, pmat, right);
      for (int m2 = 1; m2 <= 2 * l2 + 1; m2++)
        for (int m1 = 1; m1 <= 2 * l1 + 1; m1++)
          for (int k2 = 1; k2 <= 2 * l2 + 1; k2++)
            for (int k1 = 1; k1 <= 2 * l1 + 1; k1++) {
              double *x = &X[(n2 + k2 - 1) * norb_mu + (n1 + k1 - 1)];
              *x = *x + left[5 * (m1 - 1) + (k1 - 1)] * M[(n2 + m2 - 1) * norb_mu + (n1 + m1 - 1)] *
                            right[5 * (m2 - 1) + (k2 - 1)];
            }
      n2 += 2 * l2 + 1;
    }
    n1 += 2 * l1 + 1;
  }
}

/* e.FDATA/M_Fdata_2c.f90:302-415: 7-point Lagrange interpolation + recovery. */
int fbo_getMEs_2c(const fbo_tables *t, int is, int js, int iint, int isub, double x, int norb_mu,
                  int norb_nu, double *hmbox) {
  const double P_tolerance = 1.0e-5;
  const cell2c *c = C2(t, is, js, iint, isub);
  if (!c) return FBO_E_NO_TABLE;
  if (x < 0.0) return FBO_E_X_NEGATIVE;
  if (x > c->xmax * (1.0 + P_tolerance)) return FBO_E_X_OUT_OF_RANGE;
  /* the x < P_tolerance fill (:354-362) is overwritten below, as in the reference */
  int imid = (int)(x / c->dx) + 1;
  int ileft = imid - 3, iright = imid + 3;
  if (ileft < 1) { ileft = 1; iright = 7; }
  else if (iright > c->nx) { ileft = c->nx - 6; iright = c->nx; }
  double xx[7], F[64];
  for (int ip = ileft; ip <= iright; ip++) xx[ip - ileft] = (double)(ip - 1) * c->dx;
  for (int k = 0; k < c->nME; k++) F[k] = 0.0;
  for (int isum = ileft; isum <= iright; isum++) {
    double prod = 1.0;
    for (int iprod = ileft; iprod <= iright; iprod++) {
      if (iprod != isum) {
        double pden = 1.0 / (xx[isum - ileft] - xx[iprod - ileft]);
        prod = prod * (x - xx[iprod - ileft]) * pden;
      }
    }
    const double *row = c->data + (size_t)(isum - 1) * c->nME;
    for (int k = 0; k < c->nME; k++) F[k] = F[k] + row[k] * prod;
  }
  for (int i = 0; i < norb_mu * norb_nu; i++) hmbox[i] = 0.0;
  for (int k = 0; k < c->nME; k++) hmbox[(c->nu[k] - 1) * norb_mu + (c->mu[k] - 1)] = F[k];
  return FBO_OK;
}

/* e.FDATA/M_Fdata_3c.f90:293-517: adaptive bivariate interpolation of the five
 * Legendre theta cells, Legendre sum, sin(theta) on mvalue=1 MEs. */
int fbo_getMEs_3c(const fbo_tables *t, int is, int js, int ks, int iint, int isub, double x,
                  double y, int norb_mu, int norb_nu, double cost, double *hmbox) {
  const double P_tiny = 1.0e-6;
  const double P_small = 1.0 - 4.0; /* "1.0d0-4" in the reference: -3.0, so method 2 never runs */
  double p[5];
  double cost2 = cost * cost;
  double sint = sqrt(fmax(1.0e-5, 1.0 - cost2));
  p[0] = 1.0;
  p[1] = cost;
  p[2] = 0.5 * (3.0 * cost2 - 1.0);
  p[3] = 0.5 * cost * (5.0 * cost2 - 3.0);
  p[4] = 0.125 * (3.0 - cost2 * (30.0 - 35.0 * cost2));
  const cell3c *c1 = C3(t, is, js, ks, iint, isub, 1);
  if (!c1) return FBO_E_NO_TABLE;
  int nME = c1->nME, nx = c1->nx, ny = c1->ny;
  double F[64];
  for (int k = 0; k < nME; k++) F[k] = 0.0;
  double dx = c1->xmax / (nx - 1.0), dy = c1->ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
#define T3(c, i, j, m) (c)->data[((size_t)((j) - 1) * nx + ((i) - 1)) * nME + (m)]
  double gx = (T3(c1, ix + 1, iy, 0) - T3(c1, ix, iy, 0)) / dx;
  double gy = (T3(c1, ix, iy + 1, 0) - T3(c1, ix, iy, 0)) / dy;
  double gg = gx * gx + gy * gy;
  const cell3c *c = c1;
  for (int ith = 1; ith <= FBO_MAXTHETA; ith++) {
    c = C3(t, is, js, ks, iint, isub, ith);
    if (!c) return FBO_E_NO_TABLE;
    for (int m = 0; m < nME; m++) {
      double interim;
      if (gg < P_tiny) { /* method 1 */
        interim = ((1.0 - px - py) * T3(c, ix, iy, m) + px * T3(c, ix + 1, iy, m)) +
                  py * T3(c, ix, iy + 1, m);
      } else if (gg <= P_small) { /* method 2: unreachable, kept for fidelity */
        interim = (py * (py - 1) * 0.5 * T3(c, ix, iy - 1, m) + px * (px - 1) * 0.5 * T3(c, ix - 1, iy, m) +
                   (1.0 + px * py - px * px - py * py) * T3(c, ix, iy, m) +
                   px * (px - 2.0 * py + 1.0) * 0.5 * T3(c, ix + 1, iy, m) +
                   py * (py - 2.0 * px + 1.0) * 0.5 * T3(c, ix, iy + 1, m) + px * py * T3(c, ix + 1, iy + 1, m));
      } else { /* method 3: 4-point in y at four x rows, then 4-point in x */
        double g[4];
        for (int k = -1; k <= 2; k++) {
          double f1m1 = T3(c, ix + k, iy - 1, m);
          double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
          double f0p3 = 3.0 * T3(c, ix + k, iy, m), f0p6 = 2.0 * f0p3;
          double f1p3 = 3.0 * T3(c, ix + k, iy + 1, m), f1p6 = 2.0 * f1p3;
          double f2p1 = T3(c, ix + k, iy + 2, m);
          double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
          double bb2 = f1m3 - f0p6 + f1p3;
          double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
          double bb0 = f0p6;
          g[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
        }
        double f1m1 = g[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
        double f0p3 = 3.0 * g[1], f0p6 = 2.0 * f0p3;
        double f1p3 = 3.0 * g[2], f1p6 = 2.0 * f1p3;
        double f2p1 = g[3];
        double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        double bb2 = f1m3 - f0p6 + f1p3;
        double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        double bb0 = f0p6;
        interim = (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
      }
      F[m] = F[m] + p[ith - 1] * interim;
    }
  }
#undef T3
  for (int i = 0; i < norb_mu * norb_nu; i++) hmbox[i] = 0.0;
  for (int k = 0; k < c->nME; k++) {
    double v = (c->mv[k] != 1) ? F[k] : F[k] * sint;
    hmbox[(c->nu[k] - 1) * norb_mu + (c->mu[k] - 1)] = v;
  }
  return FBO_OK;
}

/* c.SYSTEM/make_cells.f90:115-144.  xl_out holds (mbeta_max+1)*3 values. */
int fbo_make_cells(const double *lat, int icluster, double *xl) {
  xl[0] = xl[1] = xl[2] = 0.0;
  if (icluster == 1) return 0;
  int mbeta = 1;
  for (int iz = -P_MBOX; iz <= P_MBOX; iz++)
    for (int iy = -P_MBOX; iy <= P_MBOX; iy++)
      for (int ix = -P_MBOX; ix <= P_MBOX; ix++) {
        if (ix == 0 && iy == 0 && iz == 0) continue;
        for (int k = 0; k < 3; k++)
          xl[3 * mbeta + k] = (double)ix * lat[0 + k] + (double)iy * lat[3 + k] + (double)iz * lat[6 + k];
        mbeta++;
      }
  return mbeta - 1;
}

/* ------------------------------------------------------------- structure */

fbo_system *fbo_system_new(const fbo_tables *t, int natoms, const double *ratom, const int *imass,
                           const double *lattice, int icluster, const double *dQ) {
  if (!t || natoms < 1) return NULL;
  for (int i = 0; i < natoms; i++)
    if (imass[i] < 1 || imass[i] > t->ns) return NULL;
  fbo_system *s = calloc(1, sizeof(*s));
  s->t = t;
  s->natoms = natoms;
  s->icluster = icluster;
  s->ratom = malloc(sizeof(double) * 3 * natoms);
  memcpy(s->ratom, ratom, sizeof(double) * 3 * natoms);
  s->imass = dupi(imass, natoms);
  memcpy(s->lattice, lattice, sizeof(double) * 9);
  s->dQ = calloc((size_t)natoms * FBO_MAXSH, sizeof(double));
  if (dQ) memcpy(s->dQ, dQ, sizeof(double) * natoms * FBO_MAXSH);
  int nimg = (icluster == 1) ? 1 : (2 * P_MBOX + 1) * (2 * P_MBOX + 1) * (2 * P_MBOX + 1);
  s->xl = calloc((size_t)nimg * 3, sizeof(double));
  s->mbeta_max = fbo_make_cells(lattice, icluster, s->xl);
  return s;
}

void fbo_system_free(fbo_system *s) {
  if (!s) return;
  free(s->ratom); free(s->imass); free(s->dQ); free(s->xl);
  free(s->neighn); free(s->noff); free(s->neigh_j); free(s->neigh_b); free(s->neigh_back); free(s->neigh_self);
  free(s->ncommon); free(s->coff); free(s->ci_j); free(s->ci_b); free(s->cj_j); free(s->cj_b); free(s->cm);
  free(s->boff);
  for (int i = 0; i < 5; i++) free(s->mat[i]);
  free(s->rho); free(s->fatom); free(s->fontop);
  for (int i = 0; i < 3; i++) { free(s->dmat[i]); free(s->f3[i]); }
  free(s->qin); free(s->wboff);
  int **pa[] = {&s->pp_n, &s->pp_off, &s->pp_j, &s->pp_b, &s->pp_map, &s->pp_self, &s->px_n, &s->px_off, &s->px_j,
                &s->px_b, &s->px_map, &s->px_point, &s->px_self, &s->pq_n, &s->pq_off, &s->pq_j, &s->pq_b, &s->pq_self,
                &s->pc_n, &s->pc_off, &s->pc_ij, &s->pc_ib, &s->pc_jj, &s->pc_jb, &s->pc_m};
  for (size_t i = 0; i < sizeof(pa) / sizeof(pa[0]); i++) free(*pa[i]);
  free(s->sv_off); free(s->svnl); free(s->vl_off); free(s->vnl);
  for (int i = 0; i < 2; i++) { free(s->rho2[i]); free(s->wrho[i]); }
  free(s);
}

static double rcut_max(const fbo_tables *t, int in) {
  double r = -99.0;
  for (int k = 0; k < t->nssh[in - 1]; k++) r = fmax(r, t->rcA[in - 1][k]);
  return r;
}

static double rcut_min(const fbo_tables *t, int in) {
  double r = 99.0;
  for (int k = 0; k < t->nssh[in - 1]; k++) r = fmin(r, t->rcA[in - 1][k]);
  return r;
}

#define RAT(s, i) (&(s)->ratom[3 * ((i) - 1)])
#define XL(s, b) (&(s)->xl[3 * (b)])

/* c.SYSTEM/M_neighbors.f90:246-519 (find_neigh_max + find_neighbors) and
 * :548-812 (find_common_max + find_common_neighbors). */
int fbo_build_lists(fbo_system *s) {
  const fbo_tables *t = s->t;
  int N = s->natoms;
  s->neighn = calloc(N, sizeof(int));
  s->noff = calloc(N + 1, sizeof(int));
  /* pass 1: count (find_neigh_max) */
  for (int ia = 1; ia <= N; ia++) {
    int in1 = s->imass[ia - 1];
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    double rci = rcut_max(t, in1);
    int num = 0;
    for (int mb = 0; mb <= s->mbeta_max; mb++)
      for (int ja = 1; ja <= N; ja++) {
        int in2 = s->imass[ja - 1];
        double r2[3];
        for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
        double rcj = rcut_max(t, in2);
        double z = fbo_distance(r1, r2);
        double range_max = rci + rcj - 0.01;
        if (ia != ja && z < 0.5) return FBO_E_ATOMS_TOO_CLOSE;
        if (z < range_max) num++;
      }
    s->neighn[ia - 1] = num;
    s->noff[ia] = s->noff[ia - 1] + num;
  }
  s->npairs = s->noff[N];
  s->neigh_j = malloc(sizeof(int) * (s->npairs + 1));
  s->neigh_b = malloc(sizeof(int) * (s->npairs + 1));
  s->neigh_back = malloc(sizeof(int) * (s->npairs + 1));
  s->neigh_self = malloc(sizeof(int) * N);
  /* pass 2: fill */
  for (int ia = 1; ia <= N; ia++) {
    int in1 = s->imass[ia - 1];
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    double rci = rcut_max(t, in1);
    int num = 0, off = s->noff[ia - 1];
    for (int mb = 0; mb <= s->mbeta_max; mb++)
      for (int ja = 1; ja <= N; ja++) {
        int in2 = s->imass[ja - 1];
        double r2[3];
        for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
        double rcj = rcut_max(t, in2);
        double z = fbo_distance(r1, r2);
        double range_max = rci + rcj - 0.01;
        if (z < range_max) {
          s->neigh_j[off + num] = ja;
          s->neigh_b[off + num] = mb;
          num++;
        }
      }
  }
  /* neigh_self and neigh_back (:457-515) */
  for (int ia = 1; ia <= N; ia++) s->neigh_self[ia - 1] = -999;
  for (int ia = 1; ia <= N; ia++) {
    int off = s->noff[ia - 1];
    for (int in = 1; in <= s->neighn[ia - 1]; in++) {
      int mb = s->neigh_b[off + in - 1], ja = s->neigh_j[off + in - 1];
      s->neigh_back[off + in - 1] = -999;
      if (ia == ja && mb == 0) {
        if (s->neigh_self[ia - 1] != -999) return FBO_E_NEIGH_SELF;
        s->neigh_self[ia - 1] = in;
      }
      int joff = s->noff[ja - 1];
      for (int jn = 1; jn <= s->neighn[ja - 1]; jn++) {
        int jb = s->neigh_b[joff + jn - 1], ka = s->neigh_j[joff + jn - 1];
        if (ia != ka || fabs(XL(s, mb)[0] + XL(s, jb)[0]) > 1.0e-3 ||
            fabs(XL(s, mb)[1] + XL(s, jb)[1]) > 1.0e-3 || fabs(XL(s, mb)[2] + XL(s, jb)[2]) > 1.0e-3) {
        } else {
          if (s->neigh_back[off + in - 1] != -999) return FBO_E_NEIGH_BACK;
          s->neigh_back[off + in - 1] = jn;
        }
      }
    }
    for (int in = 1; in <= s->neighn[ia - 1]; in++)
      if (s->neigh_back[off + in - 1] == -999) return FBO_E_NEIGH_BACK;
  }
  for (int ia = 1; ia <= N; ia++)
    if (s->neigh_self[ia - 1] == -999) return FBO_E_NEIGH_SELF;

  /* common neighbours: one pass storing into growable arrays */
  s->ncommon = calloc(N, sizeof(int));
  s->coff = calloc(N + 1, sizeof(int));
  int64_t cap = 1024, n = 0;
  int *a_ij = malloc(sizeof(int) * cap), *a_ib = malloc(sizeof(int) * cap), *a_jj = malloc(sizeof(int) * cap),
      *a_jb = malloc(sizeof(int) * cap), *a_m = malloc(sizeof(int) * cap);
  for (int ial = 1; ial <= N; ial++) {
    int off = s->noff[ial - 1], nn = s->neighn[ial - 1], num = 0;
    for (int in = 1; in <= nn; in++) {
      int ia = s->neigh_j[off + in - 1], ib = s->neigh_b[off + in - 1];
      if (ia == ial && ib == 0) continue;
      int in1 = s->imass[ia - 1];
      double r1[3];
      for (int k = 0; k < 3; k++) r1[k] = RAT(s, ia)[k] + XL(s, ib)[k];
      double rci = rcut_max(t, in1);
      for (int jn = 1; jn <= nn; jn++) {
        int ja = s->neigh_j[off + jn - 1], jb = s->neigh_b[off + jn - 1];
        if ((ja == ial && jb == 0) || in == jn) continue;
        int in2 = s->imass[ja - 1];
        double r2[3];
        for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, jb)[k];
        double rcj = rcut_max(t, in2);
        double z = fbo_distance(r1, r2);
        double range_max = rci + rcj - 0.01;
        if (z < range_max) {
          int m = -9999;
          int ioff = s->noff[ia - 1];
          for (int kn = 1; kn <= s->neighn[ia - 1]; kn++) {
            int ka = s->neigh_j[ioff + kn - 1], kb = s->neigh_b[ioff + kn - 1];
            double r3[3];
            for (int k = 0; k < 3; k++) r3[k] = XL(s, kb)[k] + RAT(s, ka)[k] - RAT(s, ia)[k];
            if (fabs(r3[0] - r2[0] + r1[0]) < 1.0e-4 && fabs(r3[1] - r2[1] + r1[1]) < 1.0e-4 &&
                fabs(r3[2] - r2[2] + r1[2]) < 1.0e-4) {
              m = kn;
              break;
            }
          }
          if (m == -9999) continue; /* dropped (:791-792) */
          if (n == cap) {
            cap *= 2;
            a_ij = realloc(a_ij, sizeof(int) * cap); a_ib = realloc(a_ib, sizeof(int) * cap);
            a_jj = realloc(a_jj, sizeof(int) * cap); a_jb = realloc(a_jb, sizeof(int) * cap);
            a_m = realloc(a_m, sizeof(int) * cap);
          }
          a_ij[n] = ia; a_ib[n] = ib; a_jj[n] = ja; a_jb[n] = jb; a_m[n] = m;
          n++; num++;
        }
      }
    }
    s->ncommon[ial - 1] = num;
    s->coff[ial] = s->coff[ial - 1] + num;
  }
  s->ntrip = n;
  s->ci_j = a_ij; s->ci_b = a_ib; s->cj_j = a_jj; s->cj_b = a_jb; s->cm = a_m;

  /* block offsets */
  s->boff = malloc(sizeof(int64_t) * (s->npairs + 1));
  int64_t o = 0;
  for (int ia = 1; ia <= N; ia++) {
    int nmu = t->norb[s->imass[ia - 1] - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      s->boff[p] = o;
      o += (int64_t)nmu * t->norb[s->imass[s->neigh_j[p] - 1] - 1];
    }
  }
  s->boff[s->npairs] = o;
  s->blen = o;
  for (int i = 0; i < 5; i++) { free(s->mat[i]); s->mat[i] = calloc(o ? o : 1, sizeof(double)); s->done[i] = 0; }
  free(s->rho); s->rho = NULL;
  free(s->fatom); free(s->fontop);
  s->fatom = calloc((size_t)(3 * s->npairs + 1), sizeof(double));
  s->fontop = calloc((size_t)(3 * s->npairs + 1), sizeof(double));
  for (int i = 0; i < 3; i++) {
    free(s->dmat[i]); s->dmat[i] = calloc((size_t)(3 * o + 1), sizeof(double));
    free(s->f3[i]); s->f3[i] = calloc((size_t)(3 * N), sizeof(double));
  }
  for (int i = 0; i < 5; i++) s->ddone[i] = 0;
  /* McWEDA density blocks: orbital blocks in the block layout, shell blocks
   * nssh_mu x nssh_nu per pair */
  free(s->wboff);
  s->wboff = malloc(sizeof(int64_t) * (s->npairs + 1));
  int64_t w = 0;
  for (int ia = 1; ia <= N; ia++) {
    int nsi = t->nssh[s->imass[ia - 1] - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      s->wboff[p] = w;
      w += (int64_t)nsi * t->nssh[s->imass[s->neigh_j[p] - 1] - 1];
    }
  }
  s->wboff[s->npairs] = w;
  s->wlen = w;
  for (int i = 0; i < 2; i++) {
    free(s->rho2[i]); s->rho2[i] = calloc(o ? o : 1, sizeof(double));
    free(s->wrho[i]); s->wrho[i] = calloc(w ? w : 1, sizeof(double));
  }
  for (int i = 0; i < 4; i++) s->rdone[i] = 0;
  return FBO_OK;
}

int64_t fbo_npairs(const fbo_system *s) { return s->npairs; }
int64_t fbo_ntriplets(const fbo_system *s) { return s->ntrip; }
int64_t fbo_block_len(const fbo_system *s) { return s->blen; }
int fbo_frame_fallbacks(const fbo_system *s) { return s->nfallback; }

int fbo_get_lists(const fbo_system *s, int *neighn, int *neigh_j, int *neigh_b, int *neigh_back,
                  int *neigh_self, int *ncommon, int *iatom_j, int *iatom_b, int *jatom_j,
                  int *jatom_b, int *mneigh) {
  if (!s->neighn) return FBO_E_ORDER;
  int N = s->natoms;
  if (neighn) memcpy(neighn, s->neighn, sizeof(int) * N);
  if (neigh_j) memcpy(neigh_j, s->neigh_j, sizeof(int) * s->npairs);
  if (neigh_b) memcpy(neigh_b, s->neigh_b, sizeof(int) * s->npairs);
  if (neigh_back) memcpy(neigh_back, s->neigh_back, sizeof(int) * s->npairs);
  if (neigh_self) memcpy(neigh_self, s->neigh_self, sizeof(int) * N);
  if (ncommon) memcpy(ncommon, s->ncommon, sizeof(int) * N);
  if (iatom_j) memcpy(iatom_j, s->ci_j, sizeof(int) * s->ntrip);
  if (iatom_b) memcpy(iatom_b, s->ci_b, sizeof(int) * s->ntrip);
  if (jatom_j) memcpy(jatom_j, s->cj_j, sizeof(int) * s->ntrip);
  if (jatom_b) memcpy(jatom_b, s->cj_b, sizeof(int) * s->ntrip);
  if (mneigh) memcpy(mneigh, s->cm, sizeof(int) * s->ntrip);
  return FBO_OK;
}

int fbo_get_blocks(const fbo_system *s, int matrix, double *out) {
  if (matrix < 0 || matrix > 4 || !s->mat[matrix]) return FBO_E_ARG;
  memcpy(out, s->mat[matrix], sizeof(double) * s->blen);
  return FBO_OK;
}

/* ------------------------------------------------------------ assemblers */

/* frame of a pair: sighat along r2 - r1 (default z when z < 1e-5), epsilon(r2, sighat) */
static double pair_frame(fbo_system *s, const double *r1, const double *r2, double *eps) {
  double z = fbo_distance(r1, r2), sighat[3];
  if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) sighat[k] = (r2[k] - r1[k]) / z;
  s->nfallback += epsilon_impl(r2, sighat, eps);
  return z;
}

/* j.ASSEMBLERS/M_assemble_2c_DOGS.f90:87-209 (S), :232-353 (T), :376-499 (dipole_z) */
static int assemble_2c_simple(fbo_system *s, int interaction, int matrix) {
  const fbo_tables *t = s->t;
  if (!s->neighn) return FBO_E_ORDER;
  double sm[81], eps[9];
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1], nnu = t->norb[in2 - 1];
      double z = pair_frame(s, r1, r2, eps);
      int rc = fbo_getMEs_2c(t, in1, in2, interaction, 0, z, nmu, nnu, sm);
      if (rc) return rc;
      fbo_rotate(t, in1, in2, eps, nmu, nnu, sm, s->mat[matrix] + s->boff[p]);
    }
  }
  s->done[matrix] = 1;
  return FBO_OK;
}

int fbo_assemble_S(fbo_system *s) { return assemble_2c_simple(s, P_OVERLAP, FBO_MAT_S); }
int fbo_assemble_T(fbo_system *s) { return assemble_2c_simple(s, P_KINETIC, FBO_MAT_T); }
int fbo_assemble_dipole_z(fbo_system *s) { return assemble_2c_simple(s, P_DIPOLE_Z, FBO_MAT_DIP); }

/* j.ASSEMBLERS/M_assemble_2c_DOGS.f90:523-829 */
int fbo_assemble_vna_2c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->done[FBO_MAT_S]) return FBO_E_ORDER;
  double *vb = s->mat[FBO_MAT_VNA], *vo = s->mat[FBO_MAT_VNAO];
  memset(vb, 0, sizeof(double) * s->blen);
  memset(vo, 0, sizeof(double) * s->blen);
  double bm[81], bx[81], eps[9], emnpl[81];
  /* ontop terms into blocko (:592-702) */
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
      double z = pair_frame(s, r1, r2, eps);
      if (ia == ja && mb == 0) continue;
      double *blk = vo + s->boff[p];
      int rc = fbo_getMEs_2c(t, in1, in2, P_VNA_ONTOPL, 0, z, nmu, nnu, bm);
      if (rc) return rc;
      fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
      for (int i = 0; i < nb; i++) blk[i] = blk[i] + bx[i] * P_EQ2;
      for (int isorp = 1; isorp <= t->nssh[in1 - 1]; isorp++) {
        double dQ = s->dQ[(ia - 1) * FBO_MAXSH + isorp - 1];
        rc = fbo_getMEs_2c(t, in1, in2, P_VNA_ONTOPL, isorp, z, nmu, nnu, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
        for (int i = 0; i < nb; i++) blk[i] = blk[i] + dQ * bx[i] * P_EQ2;
      }
      rc = fbo_getMEs_2c(t, in1, in2, P_VNA_ONTOPR, 0, z, nmu, nnu, bm);
      if (rc) return rc;
      fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
      for (int i = 0; i < nb; i++) blk[i] = blk[i] + bx[i] * P_EQ2;
      for (int isorp = 1; isorp <= t->nssh[in2 - 1]; isorp++) {
        double dQ = s->dQ[(ja - 1) * FBO_MAXSH + isorp - 1];
        rc = fbo_getMEs_2c(t, in1, in2, P_VNA_ONTOPR, isorp, z, nmu, nnu, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
        for (int i = 0; i < nb; i++) blk[i] = blk[i] + dQ * bx[i] * P_EQ2;
      }
    }
  }
  /* atom terms into block(neigh_self) (:710-816) */
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1], nb = nmu * nmu;
    int pself = s->noff[ia - 1] + s->neigh_self[ia - 1] - 1;
    double *blk = vb + s->boff[pself];
    const double *sself = s->mat[FBO_MAT_S] + s->boff[pself];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1];
      double z = pair_frame(s, r1, r2, eps);
      double rend = rcut_min(t, in1) + rcut_min(t, in2);
      double smooth = fbo_smoother(z, rend, 0.8);
      for (int i = 0; i < nb; i++) emnpl[i] = 0.0;
      if (z > 1.0e-4)
        for (int i = 0; i < nb; i++) emnpl[i] = sself[i] / z;
      int rc = fbo_getMEs_2c(t, in1, in2, P_VNA_ATOM, 0, z, nmu, nmu, bm);
      if (rc) return rc;
      fbo_rotate(t, in1, in1, eps, nmu, nmu, bm, bx);
      for (int i = 0; i < nb; i++) blk[i] = blk[i] + bx[i] * P_EQ2;
      for (int isorp = 1; isorp <= t->nssh[in2 - 1]; isorp++) {
        double dQ = s->dQ[(ja - 1) * FBO_MAXSH + isorp - 1];
        rc = fbo_getMEs_2c(t, in1, in2, P_VNA_ATOM, isorp, z, nmu, nmu, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in1, eps, nmu, nmu, bm, bx);
        for (int i = 0; i < nb; i++)
          blk[i] = blk[i] + P_EQ2 * dQ * ((1.0 - smooth) * emnpl[i] + smooth * bx[i]);
      }
    }
  }
  s->done[FBO_MAT_VNA] = s->done[FBO_MAT_VNAO] = 1;
  return FBO_OK;
}

/* j.ASSEMBLERS/M_assemble_3c_DOGS.f90:86-340 (serial order: ialpha, then its common pairs) */
int fbo_assemble_vna_3c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->done[FBO_MAT_VNA] || !s->done[FBO_MAT_S] || !s->done[FBO_MAT_DIP]) return FBO_E_ORDER;
  double bm[81], bx[81], eps[9], emnpl[81];
  for (int ial = 1; ial <= s->natoms; ial++) {
    int indna = s->imass[ial - 1];
    const double *rna = RAT(s, ial);
    for (int in = 0; in < s->ncommon[ial - 1]; in++) {
      int c = s->coff[ial - 1] + in;
      int mneigh = s->cm[c];
      if (mneigh == 0) continue;
      int ia = s->ci_j[c], ib = s->ci_b[c], ja = s->cj_j[c], jb = s->cj_b[c];
      double r1[3], r2[3], r21[3], sighat[3], rnabc[3], rhat[3];
      for (int k = 0; k < 3; k++) { r1[k] = RAT(s, ia)[k] + XL(s, ib)[k]; r2[k] = RAT(s, ja)[k] + XL(s, jb)[k]; }
      int in1 = s->imass[ia - 1], in2 = s->imass[ja - 1];
      int nmu = t->norb[in1 - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
      int p = s->noff[ia - 1] + mneigh - 1;
      double *blk = s->mat[FBO_MAT_VNA] + s->boff[p];
      const double *sb = s->mat[FBO_MAT_S] + s->boff[p];
      const double *db = s->mat[FBO_MAT_DIP] + s->boff[p];
      for (int k = 0; k < 3; k++) r21[k] = r2[k] - r1[k];
      double z = fbo_distance(r1, r2);
      if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
      else for (int k = 0; k < 3; k++) sighat[k] = r21[k] / z;
      for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1[k] + r21[k] / 2.0);
      double x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
      double d13 = fbo_distance(rna, r1), d23 = fbo_distance(rna, r2);
      if (x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
      else for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] / x;
      double cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
      s->nfallback += epsilon_impl(rhat, sighat, eps);
      double rc1 = rcut_min(t, in1), rc2 = rcut_min(t, in2), rc3 = rcut_min(t, indna);
      double stinky = fbo_smoother(d13, rc1 + rc3, 0.8);
      stinky = stinky * fbo_smoother(d23, rc2 + rc3, 0.8);
      for (int i = 0; i < nb; i++) {
        double sterm = sb[i] / 2.0, dterm = db[i] / z;
        emnpl[i] = (sterm - dterm) / d13 + (sterm + dterm) / d23;
      }
      int rc = fbo_getMEs_3c(t, in1, in2, indna, P_BCNA, 0, x, z, nmu, nnu, cost, bm);
      if (rc) return rc;
      fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
      for (int i = 0; i < nb; i++) blk[i] = blk[i] + bx[i] * P_EQ2;
      for (int isorp = 1; isorp <= t->nssh[indna - 1]; isorp++) {
        rc = fbo_getMEs_3c(t, in1, in2, indna, P_BCNA, isorp, x, z, nmu, nnu, cost, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
        double dQ = s->dQ[(ial - 1) * FBO_MAXSH + isorp - 1];
        for (int i = 0; i < nb; i++)
          blk[i] = blk[i] + dQ * (stinky * bx[i] + (1.0 - stinky) * emnpl[i]) * P_EQ2;
      }
    }
  }
  return FBO_OK;
}

/* ================================================================ forces
 * Derivative chain of k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90 and
 * M_Dassemble_3c_DOGS.f90.  Layout conventions (Fortran, column-major):
 *   deps(ix, imu, k)         deps[(ix-1) + 3*(imu-1) + 9*(k-1)]
 *   Dpmat(kx, i, j)          Dpmat[(kx-1) + 3*(i-1) + 9*(j-1)]
 *   Ddmat(kx, i, j)          Ddmat[(kx-1) + 3*(i-1) + 15*(j-1)]
 *   Dblock(ix, imu, inu)     D[(ix-1) + 3*((imu-1) + norb_mu*(inu-1))]
 */
#define EP(e, i, j) (e)[3 * ((j) - 1) + (i) - 1]
#define DE(d, ix, imu, k) (d)[((ix) - 1) + 3 * ((imu) - 1) + 9 * ((k) - 1)]

/* include/constants.h: xlevi (Levi-Civita) and delk (Kronecker) */
static double levi(int i, int j, int k) {
  if (i == j || j == k || i == k) return 0.0;
  return ((i == 1 && j == 2) || (i == 2 && j == 3) || (i == 3 && j == 1)) ? 1.0 : -1.0;
}
static double delk(int i, int j) { return i == j ? 1.0 : 0.0; }

/* b.FUNCTIONS/Dsmoother.f90 */
double fbo_Dsmoother(double r, double rend, double x) {
  double rbegin = x * rend;
  if (r > rend) return 0.0;
  if (r < rbegin) return 0.0;
  double ratio = (r - rbegin) / (rend - rbegin);
  return -(2.0 * (1.0 - ratio * ratio) * (2.0 * ratio) / (rend - rbegin));
}

/* b.FUNCTIONS/Depsilon_2c.f90: deps/dr1 of the 2c frame epsilon(r2, sighat),
 * r1 and r2 absolute positions.  Returns 1 when the colinear guard zeroed it. */
int fbo_Depsilon_2c(const double *r1, const double *r2, double z, const double *eps, double *deps) {
  for (int i = 0; i < 27; i++) deps[i] = 0.0;
  if (z < 1.0e-5) return 0;
  double ca[3];
  ca[0] = r2[1] * r1[2] - r2[2] * r1[1];
  ca[1] = r2[2] * r1[0] - r2[0] * r1[2];
  ca[2] = r2[0] * r1[1] - r2[1] * r1[0];
  double crossmag = sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
  if (fabs(crossmag) < 1.0e-3) return 1;
  double r2mag2 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
  double r1dotr2 = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  double iz = 1.0 / z, ic = 1.0 / crossmag;
  for (int ix = 1; ix <= 3; ix++)
    for (int imu = 1; imu <= 3; imu++) {
      double sum = levi(ix, imu, 1) * r2[0] + levi(ix, imu, 2) * r2[1] + levi(ix, imu, 3) * r2[2];
      DE(deps, ix, imu, 1) = (EP(eps, imu, 1) * EP(eps, ix, 3) * iz) -
                             (EP(eps, imu, 1) * (ic * ic)) * (r2mag2 * r1[ix - 1] - r1dotr2 * r2[ix - 1]) +
                             iz * ic *
                                 (delk(imu, ix) * (r2mag2 - r1dotr2) - r1[imu - 1] * r2[ix - 1] -
                                  r2[imu - 1] * r2[ix - 1] + 2.0 * r1[ix - 1] * r2[imu - 1]);
      DE(deps, ix, imu, 2) = ic * (sum + (EP(eps, imu, 2) * ic) * (r1dotr2 * r2[ix - 1] - r2mag2 * r1[ix - 1]));
      DE(deps, ix, imu, 3) = -((delk(imu, ix) - EP(eps, imu, 3) * EP(eps, ix, 3)) * iz);
    }
  return 0;
}

/* b.FUNCTIONS/Depsilon_3c.f90: dera3 = deps/dratm, der13 = deps/dr1 of the 3c
 * frame epsilon(rhat, sighat).  Returns 1 when the colinear guard zeroed them. */
int fbo_Depsilon_3c(const double *r1, const double *r2, const double *r21, double z, const double *ratm,
                    const double *rnabc, const double *eps3, double *dera3, double *der13) {
  for (int i = 0; i < 27; i++) { dera3[i] = 0.0; der13[i] = 0.0; }
  double ca[3];
  ca[0] = r21[1] * rnabc[2] - r21[2] * rnabc[1];
  ca[1] = r21[2] * rnabc[0] - r21[0] * rnabc[2];
  ca[2] = r21[0] * rnabc[1] - r21[1] * rnabc[0];
  double crossmag = sqrt(ca[0] * ca[0] + ca[1] * ca[1] + ca[2] * ca[2]);
  if (fabs(crossmag) < 1.0e-3) return 1;
  double r2mag2 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
  double r1mag2 = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
  double ratmmag2 = ratm[0] * ratm[0] + ratm[1] * ratm[1] + ratm[2] * ratm[2];
  double r1dotr2 = r1[0] * r2[0] + r1[1] * r2[1] + r1[2] * r2[2];
  double r2dotratm = r2[0] * ratm[0] + r2[1] * ratm[1] + r2[2] * ratm[2];
  double r1dotratm = r1[0] * ratm[0] + r1[1] * ratm[1] + r1[2] * ratm[2];
  for (int ix = 1; ix <= 3; ix++)
    for (int imu = 1; imu <= 3; imu++) {
      DE(dera3, ix, imu, 3) = 0.0;
      double sum = levi(ix, imu, 1) * (r2[0] - r1[0]) + levi(ix, imu, 2) * (r2[1] - r1[1]) +
                   levi(ix, imu, 3) * (r2[2] - r1[2]);
      DE(dera3, ix, imu, 2) =
          (1.0 / crossmag) *
          (sum - (EP(eps3, imu, 2) / crossmag) *
                     ((r1mag2 + r2mag2 - 2.0 * r1dotr2) * ratm[ix - 1] +
                      (r1dotr2 - r1mag2 - r2dotratm + r1dotratm) * r2[ix - 1] +
                      (r1dotr2 - r2mag2 + r2dotratm - r1dotratm) * r1[ix - 1]));
    }
  for (int ix = 1; ix <= 3; ix++) {
    DE(dera3, ix, 1, 1) = EP(eps3, 3, 3) * DE(dera3, ix, 2, 2) - EP(eps3, 2, 3) * DE(dera3, ix, 3, 2);
    DE(dera3, ix, 2, 1) = EP(eps3, 1, 3) * DE(dera3, ix, 3, 2) - EP(eps3, 3, 3) * DE(dera3, ix, 1, 2);
    DE(dera3, ix, 3, 1) = EP(eps3, 2, 3) * DE(dera3, ix, 1, 2) - EP(eps3, 1, 3) * DE(dera3, ix, 2, 2);
  }
  for (int ix = 1; ix <= 3; ix++)
    for (int imu = 1; imu <= 3; imu++) {
      DE(der13, ix, imu, 3) = (1.0 / z) * (EP(eps3, imu, 3) * EP(eps3, ix, 3) - delk(imu, ix));
      double sum = levi(ix, imu, 1) * (ratm[0] - r2[0]) + levi(ix, imu, 2) * (ratm[1] - r2[1]) +
                   levi(ix, imu, 3) * (ratm[2] - r2[2]);
      DE(der13, ix, imu, 2) =
          (1.0 / crossmag) *
          (sum - (EP(eps3, imu, 2) / crossmag) *
                     ((r2dotratm - r1dotratm + r1dotr2 - r2mag2) * ratm[ix - 1] +
                      (r2dotratm + r1dotratm - r1dotr2 - ratmmag2) * r2[ix - 1] +
                      (ratmmag2 - 2.0 * r2dotratm + r2mag2) * r1[ix - 1]));
    }
  for (int ix = 1; ix <= 3; ix++) {
    DE(der13, ix, 1, 1) = EP(eps3, 3, 3) * DE(der13, ix, 2, 2) - EP(eps3, 3, 2) * DE(der13, ix, 2, 3) -
                          EP(eps3, 2, 3) * DE(der13, ix, 3, 2) + EP(eps3, 2, 2) * DE(der13, ix, 3, 3);
    DE(der13, ix, 2, 1) = EP(eps3, 1, 3) * DE(der13, ix, 3, 2) - EP(eps3, 1, 2) * DE(der13, ix, 3, 3) -
                          EP(eps3, 3, 3) * DE(der13, ix, 1, 2) + EP(eps3, 3, 2) * DE(der13, ix, 1, 3);
    DE(der13, ix, 3, 1) = EP(eps3, 2, 3) * DE(der13, ix, 1, 2) - EP(eps3, 2, 2) * DE(der13, ix, 1, 3) -
                          EP(eps3, 1, 3) * DE(der13, ix, 2, 2) + EP(eps3, 1, 2) * DE(der13, ix, 2, 3);
  }
  return 0;
}

/* c.SYSTEM/M_Drotations.f90:203-298 (Dtwister): derivatives of the p and d
 * rotation matrices. */
void fbo_Dtwister(const double *eps, const double *deps, double *Ddmat, double *Dpmat) {
#define DPM(kx, i, j) Dpmat[((kx) - 1) + 3 * ((i) - 1) + 9 * ((j) - 1)]
#define DDM(kx, i, j) Ddmat[((kx) - 1) + 3 * ((i) - 1) + 15 * ((j) - 1)]
  for (int kx = 1; kx <= 3; kx++) {
    DPM(kx, 1, 1) = DE(deps, kx, 2, 2); DPM(kx, 1, 2) = DE(deps, kx, 2, 3); DPM(kx, 1, 3) = DE(deps, kx, 2, 1);
    DPM(kx, 2, 1) = DE(deps, kx, 3, 2); DPM(kx, 2, 2) = DE(deps, kx, 3, 3); DPM(kx, 2, 3) = DE(deps, kx, 3, 1);
    DPM(kx, 3, 1) = DE(deps, kx, 1, 2); DPM(kx, 3, 2) = DE(deps, kx, 1, 3); DPM(kx, 3, 3) = DE(deps, kx, 1, 1);
  }
  for (int imu = 1; imu <= 5; imu++)
    for (int kx = 1; kx <= 3; kx++) {
      double a12 = 0.0, a32 = 0.0, a33 = 0.0, a13 = 0.0, a11 = 0.0;
      for (int ix = 1; ix <= 3; ix++)
        for (int jx = 1; jx <= 3; jx++) {
          double a = amat(jx, ix, imu);
          if (a != 0.0) {
            a12 = a12 + a * (DE(deps, kx, ix, 1) * EP(eps, jx, 2) + EP(eps, ix, 1) * DE(deps, kx, jx, 2));
            a32 = a32 + a * (DE(deps, kx, ix, 3) * EP(eps, jx, 2) + EP(eps, ix, 3) * DE(deps, kx, jx, 2));
            a33 = a33 + a * (DE(deps, kx, ix, 3) * EP(eps, jx, 3) + EP(eps, ix, 3) * DE(deps, kx, jx, 3));
            a13 = a13 + a * (DE(deps, kx, ix, 1) * EP(eps, jx, 3) + EP(eps, ix, 1) * DE(deps, kx, jx, 3));
            a11 = a11 + a * (DE(deps, kx, ix, 1) * EP(eps, jx, 1) + EP(eps, ix, 1) * DE(deps, kx, jx, 1));
          }
        }
      DDM(kx, imu, 1) = 2.0 * a12;
      DDM(kx, imu, 2) = 2.0 * a32;
      DDM(kx, imu, 3) = sqrt(3.0) * a33;
      DDM(kx, imu, 4) = 2.0 * a13;
      DDM(kx, imu, 5) = 2.0 * a11 + a33;
    }
#undef DPM
#undef DDM
}

/* c.SYSTEM/M_Drotations.f90:322-365 (Dchooser): 3x5x5 derivative rotation by l */
static void Dchooser(int l, const double *Ddmat, const double *Dpmat, double *r) {
  for (int i = 0; i < 75; i++) r[i] = 0.0;
  if (l == 1) {
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++)
        for (int kx = 0; kx < 3; kx++) r[kx + 3 * i + 15 * j] = Dpmat[kx + 3 * i + 9 * j];
  } else if (l == 2) {
    for (int i = 0; i < 75; i++) r[i] = Ddmat[i];
  }
}

/* c.SYSTEM/M_Drotations.f90:102-153 (Drotate) and :386-478 (Dmatrix):
 * Dmatx(ix) = dL/dr M R^T + L M dR^T/dr + L dM/dr R^T shell pair by shell pair,
 * loop order m2, k2, m1, k1, ix. */
void fbo_Drotate(const fbo_tables *t, int in1, int in2, const double *eps, const double *deps, int norb_mu,
                 int norb_nu, const double *matm, const double *Dmatm, double *Dmatx) {
  double dmat[25] = {0}, pmat[9], Ddmat[75], Dpmat[27];
  double left[25], right[25], Dleft[75], Dright[75];
  fbo_twister(eps, t->have_d, dmat, pmat);
  fbo_Dtwister(eps, deps, Ddmat, Dpmat);
  int n1 = 0;
  for (int issh = 0; issh < t->nssh[in1 - 1]; issh++) {
    int l1 = t->lssh[in1 - 1][issh];
    chooser(l1, dmat, pmat, left);
    Dchooser(l1, Ddmat, Dpmat, Dleft);
    int n2 = 0;
    for (int jssh = 0; jssh < t->nssh[in2 - 1]; jssh++) {
      int l2 = t->lssh[in2 - 1][jssh];
      chooser(l2, dmat, pmat, right);
      Dchooser(l2, Ddmat, Dpmat, Dright);
      for (int k2 = 1; k2 <= 2 * l2 + 1; k2++)
        for (int k1 = 1; k1 <= 2 * l1 + 1; k1++)
          for (int ix = 0; ix < 3; ix++) Dmatx[ix + 3 * ((n1 + k1 - 1) + norb_mu * (n2 + k2 - 1))] = 0.0;
      for (int m2 = 1; m2 <= 2 * l2 + 1; m2++)
        for (int k2 = 1; k2 <= 2 * l2 + 1; k2++)
          for (int m1 = 1; m1 <= 2 * l1 + 1; m1++)
            for (int k1 = 1; k1 <= 2 * l1 + 1; k1++)
              for (int ix = 0; ix < 3; ix++) {
                double *term = &Dmatx[ix + 3 * ((n1 + k1 - 1) + norb_mu * (n2 + k2 - 1))];
                double mm = matm[(n1 + m1 - 1) + norb_mu * (n2 + m2 - 1)];
                double dm = Dmatm[ix + 3 * ((n1 + m1 - 1) + norb_mu * (n2 + m2 - 1))];
                double lf = left[(k1 - 1) + 5 * (m1 - 1)], rt = right[(k2 - 1) + 5 * (m2 - 1)];
                double dl = Dleft[ix + 3 * (k1 - 1) + 15 * (m1 - 1)], dr = Dright[ix + 3 * (k2 - 1) + 15 * (m2 - 1)];
                *term = *term + dl * rt * mm + lf * dr * mm + lf * rt * dm;
              }
      n2 += 2 * l2 + 1;
    }
    n1 += 2 * l1 + 1;
  }
}

/* e.FDATA/M_Fdata_2c.f90:446-600: 7-point Lagrange value and x-derivative.
 * Unlike the value routine, x < 1e-5 returns point 1 with a zero derivative. */
int fbo_getDMEs_2c(const fbo_tables *t, int is, int js, int iint, int isub, double x, int norb_mu,
                   int norb_nu, double *hmbox, double *Dhmbox) {
  const double P_tolerance = 1.0e-5;
  const cell2c *c = C2(t, is, js, iint, isub);
  if (!c) return FBO_E_NO_TABLE;
  if (x < 0.0) return FBO_E_X_NEGATIVE;
  if (x > c->xmax * (1.0 + P_tolerance)) return FBO_E_X_OUT_OF_RANGE;
  for (int i = 0; i < norb_mu * norb_nu; i++) { hmbox[i] = 0.0; Dhmbox[i] = 0.0; }
  if (x < P_tolerance) {
    for (int k = 0; k < c->nME; k++) hmbox[(c->nu[k] - 1) * norb_mu + (c->mu[k] - 1)] = c->data[k];
    return FBO_OK;
  }
  int imid = (int)(x / c->dx) + 1;
  int ileft = imid - 3, iright = imid + 3;
  if (ileft < 1) { ileft = 1; iright = 7; }
  else if (iright > c->nx) { ileft = c->nx - 6; iright = c->nx; }
  double xx[7], pdenom[7], F[64], dF[64];
  for (int ip = ileft; ip <= iright; ip++) xx[ip - ileft] = (double)(ip - 1) * c->dx;
  for (int k = 0; k < c->nME; k++) { F[k] = 0.0; dF[k] = 0.0; }
  for (int isum = ileft; isum <= iright; isum++) {
    double prod = 1.0;
    pdenom[isum - ileft] = 1.0;
    for (int iprod = ileft; iprod <= iright; iprod++) {
      if (iprod != isum) {
        double pden = 1.0 / (xx[isum - ileft] - xx[iprod - ileft]);
        pdenom[isum - ileft] = pdenom[isum - ileft] * pden;
        prod = prod * (x - xx[iprod - ileft]) * pden;
      }
    }
    const double *row = c->data + (size_t)(isum - 1) * c->nME;
    for (int k = 0; k < c->nME; k++) F[k] = F[k] + row[k] * prod;
  }
  for (int k = 0; k < c->nME; k++) hmbox[(c->nu[k] - 1) * norb_mu + (c->mu[k] - 1)] = F[k];
  for (int ipoint = ileft; ipoint <= iright; ipoint++) {
    double xsum = 0.0;
    for (int jpoint = ileft; jpoint <= iright; jpoint++) {
      if (ipoint == jpoint) continue;
      double xprod = 1.0;
      for (int kpoint = ileft; kpoint <= iright; kpoint++) {
        if (kpoint == jpoint || kpoint == ipoint) continue;
        xprod = xprod * (x - xx[kpoint - ileft]);
      }
      xsum = xsum + xprod;
    }
    const double *row = c->data + (size_t)(ipoint - 1) * c->nME;
    for (int k = 0; k < c->nME; k++) dF[k] = dF[k] + row[k] * pdenom[ipoint - ileft] * xsum;
  }
  for (int k = 0; k < c->nME; k++) Dhmbox[(c->nu[k] - 1) * norb_mu + (c->mu[k] - 1)] = dF[k];
  return FBO_OK;
}

/* e.FDATA/M_Fdata_3c.f90:549-846: value and derivatives (cos theta, x, y) of
 * the 3c interpolation.  The 3-point branch takes the x/y slopes of ME 1 of
 * theta cell 1 for every matrix element (:641-642); the 6-point branch cannot
 * be taken (P_small = 1.0d0-4 = -3 while gg >= 0). */
int fbo_getDMEs_3c(const fbo_tables *t, int is, int js, int ks, int iint, int isub, double x, double y,
                   int norb_mu, int norb_nu, double cost, double *hmbox, double *Dphmbox, double *Dxhmbox,
                   double *Dyhmbox) {
  const double P_tiny = 1.0e-6;
  double p[5], Dp[5];
  double cost2 = cost * cost;
  double sint = sqrt(fmax(1.0e-5, 1.0 - cost2));
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
  const cell3c *c1 = C3(t, is, js, ks, iint, isub, 1);
  if (!c1) return FBO_E_NO_TABLE;
  int nME = c1->nME, nx = c1->nx, ny = c1->ny;
  double F[64], DpF[64], DxF[64], DyF[64];
  for (int k = 0; k < nME; k++) { F[k] = 0.0; DpF[k] = 0.0; DxF[k] = 0.0; DyF[k] = 0.0; }
  double dx = c1->xmax / (nx - 1.0), dy = c1->ymax / (ny - 1.0);
  int ix = (int)(x / dx) + 1, iy = (int)(y / dy) + 1;
  if (ix < 2) ix = 2;
  if (iy < 2) iy = 2;
  if (ix > nx - 2) ix = nx - 2;
  if (iy > ny - 2) iy = ny - 2;
  double px = x / dx - ix + 1.0, py = y / dy - iy + 1.0;
#define T3(c, i, j, m) (c)->data[((size_t)((j) - 1) * nx + ((i) - 1)) * nME + (m)]
  double gx = (T3(c1, ix + 1, iy, 0) - T3(c1, ix, iy, 0)) / dx;
  double gy = (T3(c1, ix, iy + 1, 0) - T3(c1, ix, iy, 0)) / dy;
  double gg = gx * gx + gy * gy;
  const cell3c *c = c1;
  for (int ith = 1; ith <= FBO_MAXTHETA; ith++) {
    c = C3(t, is, js, ks, iint, isub, ith);
    if (!c) return FBO_E_NO_TABLE;
    double pt = p[ith - 1], dpt = Dp[ith - 1];
    for (int m = 0; m < nME; m++) {
      if (gg < P_tiny) {
        double interim = ((1.0 - px - py) * T3(c, ix, iy, m) + px * T3(c, ix + 1, iy, m)) + py * T3(c, ix, iy + 1, m);
        F[m] = F[m] + pt * interim;
        DpF[m] = DpF[m] + dpt * interim;
        DxF[m] = DxF[m] + pt * gx;
        DyF[m] = DyF[m] + pt * gy;
      } else {
        double g[4], gp[4];
        for (int k = -1; k <= 2; k++) {
          double f1m1 = T3(c, ix + k, iy - 1, m);
          double f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
          double f0p3 = 3.0 * T3(c, ix + k, iy, m), f0p6 = 2.0 * f0p3;
          double f1p3 = 3.0 * T3(c, ix + k, iy + 1, m), f1p6 = 2.0 * f1p3;
          double f2p1 = T3(c, ix + k, iy + 2, m);
          double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
          double bb2 = f1m3 - f0p6 + f1p3;
          double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
          double bb0 = f0p6;
          g[k + 1] = ((bb3 * py + bb2) * py + bb1) * py + bb0;
          gp[k + 1] = ((3.0 * bb3 * py + 2.0 * bb2) * py + bb1);
        }
        double f1m1 = g[0], f1m2 = 2.0 * f1m1, f1m3 = 3.0 * f1m1;
        double f0p3 = 3.0 * g[1], f0p6 = 2.0 * f0p3;
        double f1p3 = 3.0 * g[2], f1p6 = 2.0 * f1p3;
        double f2p1 = g[3];
        double bb3 = -f1m1 + f0p3 - f1p3 + f2p1;
        double bb2 = f1m3 - f0p6 + f1p3;
        double bb1 = -f1m2 - f0p3 + f1p6 - f2p1;
        double bb0 = f0p6;
        double interim = (((bb3 * px + bb2) * px + bb1) * px + bb0) / 36.0;
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
#undef T3
  for (int i = 0; i < norb_mu * norb_nu; i++) { hmbox[i] = 0.0; Dphmbox[i] = 0.0; Dxhmbox[i] = 0.0; Dyhmbox[i] = 0.0; }
  for (int k = 0; k < c->nME; k++) {
    int o = (c->nu[k] - 1) * norb_mu + (c->mu[k] - 1);
    if (c->mv[k] != 1) {
      hmbox[o] = F[k]; Dphmbox[o] = DpF[k]; Dxhmbox[o] = DxF[k]; Dyhmbox[o] = DyF[k];
    } else {
      hmbox[o] = F[k] * sint;
      Dphmbox[o] = DpF[k] * sint - cost * F[k] / sint;
      Dxhmbox[o] = DxF[k] * sint;
      Dyhmbox[o] = DyF[k] * sint;
    }
  }
  return FBO_OK;
}

int fbo_set_denmat(fbo_system *s, const double *rho) {
  if (!s->neighn) return FBO_E_ORDER;
  free(s->rho);
  s->rho = malloc(sizeof(double) * (size_t)(s->blen ? s->blen : 1));
  memcpy(s->rho, rho, sizeof(double) * (size_t)s->blen);
  return FBO_OK;
}

/* frame and its r1-derivative of a pair (sighat along r2 - r1, epsilon(r2, sighat)) */
static double pair_dframe(fbo_system *s, const double *r1, const double *r2, double *eps, double *deps,
                          double *eta) {
  double z = fbo_distance(r1, r2), sighat[3];
  if (z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) sighat[k] = (r2[k] - r1[k]) / z;
  (void)epsilon_impl(r2, sighat, eps);
  s->ndwarn += fbo_Depsilon_2c(r1, r2, z, eps, deps);
  for (int k = 0; k < 3; k++) eta[k] = eps[6 + k];
  return z;
}

/* k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:95-266 (S), :292-463 (T), :487-655 (dipole_z) */
static int dassemble_2c_simple(fbo_system *s, int interaction, int d) {
  const fbo_tables *t = s->t;
  if (!s->neighn) return FBO_E_ORDER;
  double sm[81], dsm[81], vdsm[243], eps[9], deps[27], eta[3];
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
      double z = pair_dframe(s, r1, r2, eps, deps, eta);
      int rc = fbo_getDMEs_2c(t, in1, in2, interaction, 0, z, nmu, nnu, sm, dsm);
      if (rc) return rc;
      for (int i = 0; i < 3 * nb; i++) vdsm[i] = 0.0;
      for (int i = 0; i < nb; i++)
        if (z > 1.0e-3)
          for (int k = 0; k < 3; k++) vdsm[k + 3 * i] = -eta[k] * dsm[i];
      fbo_Drotate(t, in1, in2, eps, deps, nmu, nnu, sm, vdsm, s->dmat[d] + 3 * s->boff[p]);
    }
  }
  s->ddone[d] = 1;
  return FBO_OK;
}

int fbo_Dassemble_S(fbo_system *s) { return dassemble_2c_simple(s, P_OVERLAP, 0); }
int fbo_Dassemble_T(fbo_system *s) { return dassemble_2c_simple(s, P_KINETIC, 1); }
int fbo_Dassemble_dipole_z(fbo_system *s) { return dassemble_2c_simple(s, P_DIPOLE_Z, 2); }

/* one (interaction, isorp) term of the ontop forces: f -= rho (dQ) vdx P_eq2 */
static int ontop_term(fbo_system *s, int in1, int in2, int iint, int isorp, double z, const double *eps,
                      const double *deps, const double *eta, const double *rho, double dQ, int use_dq, double *f) {
  const fbo_tables *t = s->t;
  int nmu = t->norb[in1 - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
  double bm[81], dbm[81], vdbm[243], vdx[243];
  int rc = fbo_getDMEs_2c(t, in1, in2, iint, isorp, z, nmu, nnu, bm, dbm);
  if (rc) return rc;
  for (int i = 0; i < 3 * nb; i++) vdbm[i] = 0.0;
  for (int i = 0; i < nb; i++)
    if (z > 1.0e-3)
      for (int k = 0; k < 3; k++) vdbm[k + 3 * i] = -eta[k] * dbm[i];
  fbo_Drotate(t, in1, in2, eps, deps, nmu, nnu, bm, vdbm, vdx);
  for (int i = 0; i < nb; i++)
    for (int k = 0; k < 3; k++) {
      if (use_dq) f[k] = f[k] - rho[i] * dQ * vdx[k + 3 * i] * P_EQ2;
      else f[k] = f[k] - rho[i] * vdx[k + 3 * i] * P_EQ2;
    }
  return FBO_OK;
}

/* k.DASSEMBLERS/M_Dassemble_2c_DOGS.f90:679-1199: forces%vna_ontop (non-self
 * pairs, ontop L and R with their charged shells, contracted with the pair's
 * density block) and forces%vna_atom (every pair, contracted with the density
 * block of neigh_self, smoothed monopole for the charged shells). */
int fbo_Dassemble_vna_2c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->rho || !s->done[FBO_MAT_S]) return FBO_E_ORDER;
  double eps[9], deps[27], eta[3];
  memset(s->fatom, 0, sizeof(double) * 3 * s->npairs);
  memset(s->fontop, 0, sizeof(double) * 3 * s->npairs);
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1];
      double z = pair_dframe(s, r1, r2, eps, deps, eta);
      if (ia == ja && mb == 0) continue;
      const double *rho = s->rho + s->boff[p];
      double *f = s->fontop + 3 * p;
      int rc = ontop_term(s, in1, in2, P_VNA_ONTOPL, 0, z, eps, deps, eta, rho, 0.0, 0, f);
      if (rc) return rc;
      for (int isorp = 1; isorp <= t->nssh[in1 - 1]; isorp++) {
        rc = ontop_term(s, in1, in2, P_VNA_ONTOPL, isorp, z, eps, deps, eta, rho,
                        s->dQ[(ia - 1) * FBO_MAXSH + isorp - 1], 1, f);
        if (rc) return rc;
      }
      rc = ontop_term(s, in1, in2, P_VNA_ONTOPR, 0, z, eps, deps, eta, rho, 0.0, 0, f);
      if (rc) return rc;
      for (int isorp = 1; isorp <= t->nssh[in2 - 1]; isorp++) {
        rc = ontop_term(s, in1, in2, P_VNA_ONTOPR, isorp, z, eps, deps, eta, rho,
                        s->dQ[(ja - 1) * FBO_MAXSH + isorp - 1], 1, f);
        if (rc) return rc;
      }
    }
  }
  double bm[81], dbm[81], bx[81], vdbm[243], vdx[243], emnpl[81], vdemnpl[243];
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1], nb = nmu * nmu;
    int pself = s->noff[ia - 1] + s->neigh_self[ia - 1] - 1;
    const double *rho = s->rho + s->boff[pself];
    const double *sself = s->mat[FBO_MAT_S] + s->boff[pself];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1];
      double z = pair_dframe(s, r1, r2, eps, deps, eta);
      double rend = rcut_min(t, in1) + rcut_min(t, in2);
      double smooth = fbo_smoother(z, rend, 0.8);
      double Dsmooth = fbo_Dsmoother(z, rend, 0.8);
      double *f = s->fatom + 3 * p;
      int rc = fbo_getDMEs_2c(t, in1, in2, P_VNA_ATOM, 0, z, nmu, nmu, bm, dbm);
      if (rc) return rc;
      for (int i = 0; i < 3 * nb; i++) vdbm[i] = 0.0;
      for (int i = 0; i < nb; i++)
        if (z > 1.0e-3)
          for (int k = 0; k < 3; k++) vdbm[k + 3 * i] = -eta[k] * dbm[i];
      fbo_Drotate(t, in1, in1, eps, deps, nmu, nmu, bm, vdbm, vdx);
      for (int i = 0; i < nb; i++)
        for (int k = 0; k < 3; k++) f[k] = f[k] - rho[i] * vdx[k + 3 * i] * P_EQ2;
      for (int isorp = 1; isorp <= t->nssh[in2 - 1]; isorp++) {
        double dQ = s->dQ[(ja - 1) * FBO_MAXSH + isorp - 1];
        rc = fbo_getDMEs_2c(t, in1, in2, P_VNA_ATOM, isorp, z, nmu, nmu, bm, dbm);
        if (rc) return rc;
        for (int i = 0; i < nb; i++) emnpl[i] = 0.0;
        for (int i = 0; i < 3 * nb; i++) { vdbm[i] = 0.0; vdemnpl[i] = 0.0; }
        for (int i = 0; i < nb; i++)
          if (z > 1.0e-4) {
            for (int k = 0; k < 3; k++) vdbm[k + 3 * i] = -eta[k] * dbm[i];
            emnpl[i] = sself[i] / z;
            for (int k = 0; k < 3; k++) vdemnpl[k + 3 * i] = eta[k] * sself[i] / (z * z);
          }
        fbo_rotate(t, in1, in1, eps, nmu, nmu, bm, bx);
        fbo_Drotate(t, in1, in1, eps, deps, nmu, nmu, bm, vdbm, vdx);
        for (int i = 0; i < nb; i++)
          for (int k = 0; k < 3; k++)
            f[k] = f[k] - rho[i] * P_EQ2 * dQ *
                              (smooth * vdx[k + 3 * i] - eta[k] * Dsmooth * bx[i] + (1.0 - smooth) * vdemnpl[k + 3 * i] +
                               eta[k] * Dsmooth * emnpl[i]);
      }
    }
  }
  s->ddone[3] = 1;
  return FBO_OK;
}

/* triplet geometry of M_Dassemble_3c_DOGS.f90:305-362 (neutral loop) and
 * :495-562 (charged loop: rhat from rna - (r1 + r2)/2 instead of rnabc) */
typedef struct {
  double r1[3], r2[3], r21[3], rna[3], rnabc[3], sighat[3], rhat[3], amt[3], bmt[3];
  double z, x, cost, d13, d23;
  double eps[9], depsA[27], depsB[27];
} dgeo3;

static void dgeo3_build(fbo_system *s, int ial, int ia, int ib, int ja, int jb, int charged, dgeo3 *G) {
  for (int k = 0; k < 3; k++) {
    G->r1[k] = RAT(s, ia)[k] + XL(s, ib)[k];
    G->r2[k] = RAT(s, ja)[k] + XL(s, jb)[k];
    G->rna[k] = RAT(s, ial)[k];
    G->r21[k] = G->r2[k] - G->r1[k];
  }
  G->z = fbo_distance(G->r1, G->r2);
  if (G->z < 1.0e-5) { G->sighat[0] = 0.0; G->sighat[1] = 0.0; G->sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) G->sighat[k] = G->r21[k] / G->z;
  for (int k = 0; k < 3; k++) G->rnabc[k] = G->rna[k] - (G->r1[k] + G->r21[k] / 2.0);
  G->x = sqrt(G->rnabc[0] * G->rnabc[0] + G->rnabc[1] * G->rnabc[1] + G->rnabc[2] * G->rnabc[2]);
  G->d13 = fbo_distance(G->rna, G->r1);
  G->d23 = fbo_distance(G->rna, G->r2);
  if (G->x < 1.0e-5) { G->rhat[0] = 0.0; G->rhat[1] = 0.0; G->rhat[2] = 0.0; }
  else if (charged) for (int k = 0; k < 3; k++) G->rhat[k] = (G->rna[k] - 0.5 * (G->r1[k] + G->r2[k])) / G->x;
  else for (int k = 0; k < 3; k++) G->rhat[k] = G->rnabc[k] / G->x;
  G->cost = G->sighat[0] * G->rhat[0] + G->sighat[1] * G->rhat[1] + G->sighat[2] * G->rhat[2];
  (void)epsilon_impl(G->rhat, G->sighat, G->eps);
  s->ndwarn += fbo_Depsilon_3c(G->r1, G->r2, G->r21, G->z, G->rna, G->rnabc, G->eps, G->depsA, G->depsB);
  for (int k = 0; k < 3; k++) {
    G->amt[k] = (G->x > 1.0e-5) ? (G->sighat[k] - G->cost * G->rhat[k]) / G->x : 0.0;
    G->bmt[k] = (G->cost * G->sighat[k] - G->rhat[k]) / G->z;
  }
}

/* f3naMa / f3naMb of the matrix elements of the cell (M_Dassemble_3c_DOGS.f90:388-399),
 * rotated with depsA / depsB; f3naXc by Newton's third law */
static int dforce3_terms(fbo_system *s, int in1, int in2, int indna, int isorp, const dgeo3 *G, double *bm,
                         double *bx, double *Xa, double *Xb, double *Xc) {
  const fbo_tables *t = s->t;
  int nmu = t->norb[in1 - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
  double dp[81], dxm[81], dym[81], Ma[243], Mb[243];
  int rc = fbo_getDMEs_3c(t, in1, in2, indna, P_BCNA, isorp, G->x, G->z, nmu, nnu, G->cost, bm, dp, dxm, dym);
  if (rc) return rc;
  if (bx) fbo_rotate(t, in1, in2, G->eps, nmu, nnu, bm, bx);
  for (int i = 0; i < 3 * nb; i++) { Ma[i] = 0.0; Mb[i] = 0.0; }
  const cell3c *c = C3(t, in1, in2, indna, P_BCNA, isorp, 1);
  for (int k = 0; k < c->nME; k++) {
    int o = (c->nu[k] - 1) * nmu + (c->mu[k] - 1);
    for (int q = 0; q < 3; q++) Ma[q + 3 * o] = G->rhat[q] * dxm[o] + G->amt[q] * dp[o];
    for (int q = 0; q < 3; q++) Mb[q + 3 * o] = -G->sighat[q] * dym[o] + G->bmt[q] * dp[o] - Ma[q + 3 * o] / 2.0;
  }
  fbo_Drotate(t, in1, in2, G->eps, G->depsA, nmu, nnu, bm, Ma, Xa);
  fbo_Drotate(t, in1, in2, G->eps, G->depsB, nmu, nnu, bm, Mb, Xb);
  for (int i = 0; i < 3 * nb; i++) Xc[i] = -Xa[i] - Xb[i];
  return FBO_OK;
}

/* k.DASSEMBLERS/M_Dassemble_3c_DOGS.f90:136-793: f3naa (on ialpha), f3nab (on
 * iatom) and f3nac (on jatom), neutral loop over all triplets first, then the
 * charged loop with the smoothed monopole terms. */
int fbo_Dassemble_vna_3c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->rho || !s->done[FBO_MAT_S] || !s->done[FBO_MAT_DIP] || !s->ddone[0] || !s->ddone[2])
    return FBO_E_ORDER;
  double bm[81], bx[81], Xa[243], Xb[243], Xc[243];
  for (int i = 0; i < 3; i++) memset(s->f3[i], 0, sizeof(double) * 3 * s->natoms);
  for (int ial = 1; ial <= s->natoms; ial++) {
    int indna = s->imass[ial - 1];
    for (int in = 0; in < s->ncommon[ial - 1]; in++) {
      int c = s->coff[ial - 1] + in;
      int mneigh = s->cm[c];
      if (mneigh == 0) continue;
      int ia = s->ci_j[c], ja = s->cj_j[c];
      int in1 = s->imass[ia - 1], in2 = s->imass[ja - 1];
      int nmu = t->norb[in1 - 1], nnu = t->norb[in2 - 1];
      const double *rho = s->rho + s->boff[s->noff[ia - 1] + mneigh - 1];
      dgeo3 G;
      dgeo3_build(s, ial, ia, s->ci_b[c], ja, s->cj_b[c], 0, &G);
      int rc = dforce3_terms(s, in1, in2, indna, 0, &G, bm, NULL, Xa, Xb, Xc);
      if (rc) return rc;
      double *fa = s->f3[0] + 3 * (ial - 1), *fb = s->f3[1] + 3 * (ia - 1), *fc = s->f3[2] + 3 * (ja - 1);
      for (int inu = 0; inu < nnu; inu++)
        for (int imu = 0; imu < nmu; imu++) {
          int o = imu + nmu * inu;
          for (int k = 0; k < 3; k++) fa[k] = fa[k] - rho[o] * Xa[k + 3 * o] * P_EQ2;
          for (int k = 0; k < 3; k++) fb[k] = fb[k] - rho[o] * Xb[k + 3 * o] * P_EQ2;
          for (int k = 0; k < 3; k++) fc[k] = fc[k] - rho[o] * Xc[k + 3 * o] * P_EQ2;
        }
    }
  }
  double emnpl[81], dA[243], dB[243], dC[243];
  for (int ial = 1; ial <= s->natoms; ial++) {
    int indna = s->imass[ial - 1];
    for (int in = 0; in < s->ncommon[ial - 1]; in++) {
      int c = s->coff[ial - 1] + in;
      int mneigh = s->cm[c];
      if (mneigh == 0) continue;
      int ia = s->ci_j[c], ja = s->cj_j[c];
      int in1 = s->imass[ia - 1], in2 = s->imass[ja - 1];
      int nmu = t->norb[in1 - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
      int64_t bo = s->boff[s->noff[ia - 1] + mneigh - 1];
      const double *rho = s->rho + bo;
      const double *sb = s->mat[FBO_MAT_S] + bo, *db = s->mat[FBO_MAT_DIP] + bo;
      const double *sD = s->dmat[0] + 3 * bo, *dD = s->dmat[2] + 3 * bo;
      dgeo3 G;
      dgeo3_build(s, ial, ia, s->ci_b[c], ja, s->cj_b[c], 1, &G);
      double rhatA1[3] = {0, 0, 0}, rhatA2[3] = {0, 0, 0};
      if (G.d13 > 1.0e-5) for (int k = 0; k < 3; k++) rhatA1[k] = (G.rna[k] - G.r1[k]) / G.d13;
      if (G.d23 > 1.0e-5) for (int k = 0; k < 3; k++) rhatA2[k] = (G.rna[k] - G.r2[k]) / G.d23;
      double rend1 = rcut_min(t, in1) + rcut_min(t, indna), rend2 = rcut_min(t, in2) + rcut_min(t, indna);
      double st13 = fbo_smoother(G.d13, rend1, 0.8), dst13 = fbo_Dsmoother(G.d13, rend1, 0.8);
      double st23 = fbo_smoother(G.d23, rend2, 0.8), dst23 = fbo_Dsmoother(G.d23, rend2, 0.8);
      double stinky = st13 * st23;
      double dstnA[3], dstnB[3], dstnC[3];
      for (int k = 0; k < 3; k++) {
        dstnB[k] = -dst13 * st23 * rhatA1[k];
        dstnC[k] = -st13 * dst23 * rhatA2[k];
        dstnA[k] = -dstnB[k] - dstnC[k];
      }
      double d13sq = G.d13 * G.d13, d23sq = G.d23 * G.d23;
      for (int i = 0; i < nb; i++) {
        double sterm = sb[i] / 2.0, dterm = db[i] / G.z;
        emnpl[i] = (sterm - dterm) / G.d13 + (sterm + dterm) / G.d23;
        for (int k = 0; k < 3; k++) {
          double spterm = sD[k + 3 * i] / 2.0;
          double dpterm = dD[k + 3 * i] / G.z;
          dpterm = dpterm + dterm * G.sighat[k] / G.z;
          dA[k + 3 * i] = -((sterm - dterm) * rhatA1[k] / d13sq) - (sterm + dterm) * rhatA2[k] / d23sq;
          dB[k + 3 * i] = (sterm - dterm) * rhatA1[k] / d13sq + (spterm - dpterm) / G.d13 + (spterm + dpterm) / G.d23;
          dC[k + 3 * i] = -dA[k + 3 * i] - dB[k + 3 * i];
        }
      }
      double *fa = s->f3[0] + 3 * (ial - 1), *fb = s->f3[1] + 3 * (ia - 1), *fc = s->f3[2] + 3 * (ja - 1);
      for (int isorp = 1; isorp <= t->nssh[indna - 1]; isorp++) {
        int rc = dforce3_terms(s, in1, in2, indna, isorp, &G, bm, bx, Xa, Xb, Xc);
        if (rc) return rc;
        double dQ = s->dQ[(ial - 1) * FBO_MAXSH + isorp - 1];
        for (int inu = 0; inu < nnu; inu++)
          for (int imu = 0; imu < nmu; imu++) {
            int o = imu + nmu * inu;
            double w = P_EQ2 * dQ * rho[o];
            for (int k = 0; k < 3; k++)
              fa[k] = fa[k] - w * (stinky * Xa[k + 3 * o] + dstnA[k] * bx[o] - dstnA[k] * emnpl[o] +
                                   (1.0 - stinky) * dA[k + 3 * o]);
            for (int k = 0; k < 3; k++)
              fb[k] = fb[k] - w * (stinky * Xb[k + 3 * o] + dstnB[k] * bx[o] - dstnB[k] * emnpl[o] +
                                   (1.0 - stinky) * dB[k + 3 * o]);
            for (int k = 0; k < 3; k++)
              fc[k] = fc[k] - w * (stinky * Xc[k + 3 * o] + dstnC[k] * bx[o] - dstnC[k] * emnpl[o] +
                                   (1.0 - stinky) * dC[k + 3 * o]);
          }
      }
    }
  }
  s->ddone[4] = 1;
  return FBO_OK;
}

int fbo_get_dblocks(const fbo_system *s, int which, double *out) {
  if (which < 0 || which > 2 || !s->dmat[which]) return FBO_E_ARG;
  memcpy(out, s->dmat[which], sizeof(double) * 3 * (size_t)s->blen);
  return FBO_OK;
}

int fbo_get_forces(const fbo_system *s, double *vna_atom, double *vna_ontop, double *f3naa, double *f3nab,
                   double *f3nac) {
  if (!s->neighn) return FBO_E_ORDER;
  if (vna_atom) memcpy(vna_atom, s->fatom, sizeof(double) * 3 * (size_t)s->npairs);
  if (vna_ontop) memcpy(vna_ontop, s->fontop, sizeof(double) * 3 * (size_t)s->npairs);
  if (f3naa) memcpy(f3naa, s->f3[0], sizeof(double) * 3 * (size_t)s->natoms);
  if (f3nab) memcpy(f3nab, s->f3[1], sizeof(double) * 3 * (size_t)s->natoms);
  if (f3nac) memcpy(f3nac, s->f3[2], sizeof(double) * 3 * (size_t)s->natoms);
  return FBO_OK;
}

int fbo_deps_warnings(const fbo_system *s) { return s->ndwarn; }

/* ======================================================= McWEDA densities
 * j.ASSEMBLERS/M_assemble_rho_McWEDA_DOGS.f90: the input density rho_in (and
 * rho_bond) in orbital blocks, and its shell-averaged form rho_in_weighted (and
 * rho_bond_weighted) in shell blocks, each a sum over the shells isorp of the
 * contributing atom of Qin(isorp) times an interpolated (and, for orbital
 * blocks, rotated) table.  Serial order: atoms, then neighbours or common
 * pairs, then isorp; every term is added to the block as it is formed. */
#define P_RHO_ONTOPL 15
#define P_RHO_ONTOPR 16
#define P_RHO_ATOM 17
#define P_RHOS_ONTOPL 19
#define P_RHOS_ONTOPR 20
#define P_RHOS_ATOM 21
#define P_RHO_3C 1
#define P_RHOS_3C 2

int fbo_set_qin(fbo_system *s, const double *qin) {
  if (!s || !qin) return FBO_E_ARG;
  if (!s->qin) s->qin = malloc(sizeof(double) * (size_t)s->natoms * FBO_MAXSH);
  memcpy(s->qin, qin, sizeof(double) * (size_t)s->natoms * FBO_MAXSH);
  for (int i = 0; i < 4; i++) s->rdone[i] = 0;
  return FBO_OK;
}

#define QIN(s, ia, isorp) ((s)->qin[((ia) - 1) * FBO_MAXSH + (isorp) - 1])

/* assemble_rho_2c, M_assemble_rho_McWEDA_DOGS.f90:90-341 */
int fbo_assemble_rho_2c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->qin || !s->wboff) return FBO_E_ORDER;
  double *ri = s->rho2[0], *rb = s->rho2[1];
  memset(ri, 0, sizeof(double) * (s->blen ? s->blen : 1));
  memset(rb, 0, sizeof(double) * (s->blen ? s->blen : 1));
  double bm[81], bx[81], eps[9];
  /* ontop terms on every non-self pair (:166-247) */
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
      double z = pair_frame(s, r1, r2, eps);
      if (ia == ja && mb == 0) continue;
      double *bi = ri + s->boff[p], *bb = rb + s->boff[p];
      for (int isorp = 1; isorp <= t->nssh[in1 - 1]; isorp++) {
        double q = QIN(s, ia, isorp);
        int rc = fbo_getMEs_2c(t, in1, in2, P_RHO_ONTOPL, isorp, z, nmu, nnu, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
        for (int i = 0; i < nb; i++) { bi[i] = bi[i] + bx[i] * q; bb[i] = bb[i] + bx[i] * q; }
      }
      for (int isorp = 1; isorp <= t->nssh[in2 - 1]; isorp++) {
        double q = QIN(s, ja, isorp);
        int rc = fbo_getMEs_2c(t, in1, in2, P_RHO_ONTOPR, isorp, z, nmu, nnu, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
        for (int i = 0; i < nb; i++) { bi[i] = bi[i] + bx[i] * q; bb[i] = bb[i] + bx[i] * q; }
      }
    }
  }
  /* atom terms on the self block: every neighbour into rho_in, the atom itself
   * also into rho_bond (:250-335) */
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1], nb = nmu * nmu;
    int pself = s->noff[ia - 1] + s->neigh_self[ia - 1] - 1;
    double *bi = ri + s->boff[pself], *bb = rb + s->boff[pself];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1];
      double z = pair_frame(s, r1, r2, eps);
      int self = (ia == ja && mb == 0);
      for (int isorp = 1; isorp <= t->nssh[in2 - 1]; isorp++) {
        double q = QIN(s, ja, isorp);
        int rc = fbo_getMEs_2c(t, in1, in2, P_RHO_ATOM, isorp, z, nmu, nmu, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in1, eps, nmu, nmu, bm, bx);
        for (int i = 0; i < nb; i++) bi[i] = bi[i] + bx[i] * q;
        if (self)
          for (int i = 0; i < nb; i++) bb[i] = bb[i] + bx[i] * q;
      }
    }
  }
  s->rdone[0] = 1;
  s->rdone[1] = 0;
  return FBO_OK;
}

/* triplet geometry of the density assemblers (M_assemble_rho_McWEDA_DOGS.f90:451-468) */
static void rho3_geometry(fbo_system *s, int ial, int c, double *x, double *z, double *cost, double *eps) {
  const double *rna = RAT(s, ial);
  int ia = s->ci_j[c], ib = s->ci_b[c], ja = s->cj_j[c], jb = s->cj_b[c];
  double r1[3], r2[3], r21[3], sighat[3], rnabc[3], rhat[3];
  for (int k = 0; k < 3; k++) { r1[k] = RAT(s, ia)[k] + XL(s, ib)[k]; r2[k] = RAT(s, ja)[k] + XL(s, jb)[k]; }
  for (int k = 0; k < 3; k++) r21[k] = r2[k] - r1[k];
  *z = fbo_distance(r1, r2);
  if (*z < 1.0e-5) { sighat[0] = 0.0; sighat[1] = 0.0; sighat[2] = 1.0; }
  else for (int k = 0; k < 3; k++) sighat[k] = r21[k] / *z;
  for (int k = 0; k < 3; k++) rnabc[k] = rna[k] - (r1[k] + r21[k] / 2.0);
  *x = sqrt(rnabc[0] * rnabc[0] + rnabc[1] * rnabc[1] + rnabc[2] * rnabc[2]);
  if (*x < 1.0e-5) { rhat[0] = 0.0; rhat[1] = 0.0; rhat[2] = 0.0; }
  else for (int k = 0; k < 3; k++) rhat[k] = rnabc[k] / *x;
  *cost = sighat[0] * rhat[0] + sighat[1] * rhat[1] + sighat[2] * rhat[2];
  s->nfallback += epsilon_impl(rhat, sighat, eps);
}

/* assemble_rho_3c, M_assemble_rho_McWEDA_DOGS.f90:362-511 */
int fbo_assemble_rho_3c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->rdone[0]) return FBO_E_ORDER;
  double bm[81], bx[81], eps[9];
  for (int ial = 1; ial <= s->natoms; ial++) {
    int indna = s->imass[ial - 1];
    for (int in = 0; in < s->ncommon[ial - 1]; in++) {
      int c = s->coff[ial - 1] + in;
      int mneigh = s->cm[c];
      if (mneigh == 0) continue;
      int ia = s->ci_j[c], ja = s->cj_j[c];
      int in1 = s->imass[ia - 1], in2 = s->imass[ja - 1];
      int nmu = t->norb[in1 - 1], nnu = t->norb[in2 - 1], nb = nmu * nnu;
      double x, z, cost;
      rho3_geometry(s, ial, c, &x, &z, &cost, eps);
      double *blk = s->rho2[0] + s->boff[s->noff[ia - 1] + mneigh - 1];
      for (int isorp = 1; isorp <= t->nssh[indna - 1]; isorp++) {
        double q = QIN(s, ial, isorp);
        int rc = fbo_getMEs_3c(t, in1, in2, indna, P_RHO_3C, isorp, x, z, nmu, nnu, cost, bm);
        if (rc) return rc;
        fbo_rotate(t, in1, in2, eps, nmu, nnu, bm, bx);
        for (int i = 0; i < nb; i++) blk[i] = blk[i] + bx[i] * q;
      }
    }
  }
  s->rdone[1] = 1;
  return FBO_OK;
}

/* assemble_rho_weighted_2c, M_assemble_rho_McWEDA_DOGS.f90:534-762 (no rotation:
 * shell blocks are scalars per shell pair) */
int fbo_assemble_rho_weighted_2c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->qin || !s->wboff) return FBO_E_ORDER;
  double *wi = s->wrho[0], *wb = s->wrho[1];
  memset(wi, 0, sizeof(double) * (s->wlen ? s->wlen : 1));
  memset(wb, 0, sizeof(double) * (s->wlen ? s->wlen : 1));
  double bm[16];
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nsi = t->nssh[in1 - 1];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1], nsj = t->nssh[in2 - 1], nb = nsi * nsj;
      double z = fbo_distance(r1, r2);
      if (ia == ja && mb == 0) continue;
      double *bi = wi + s->wboff[p], *bb = wb + s->wboff[p];
      for (int isorp = 1; isorp <= nsi; isorp++) {
        double q = QIN(s, ia, isorp);
        int rc = fbo_getMEs_2c(t, in1, in2, P_RHOS_ONTOPL, isorp, z, nsi, nsj, bm);
        if (rc) return rc;
        for (int i = 0; i < nb; i++) { bi[i] = bi[i] + bm[i] * q; bb[i] = bb[i] + bm[i] * q; }
      }
      for (int isorp = 1; isorp <= nsj; isorp++) {
        double q = QIN(s, ja, isorp);
        int rc = fbo_getMEs_2c(t, in1, in2, P_RHOS_ONTOPR, isorp, z, nsi, nsj, bm);
        if (rc) return rc;
        for (int i = 0; i < nb; i++) { bi[i] = bi[i] + bm[i] * q; bb[i] = bb[i] + bm[i] * q; }
      }
    }
  }
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nsi = t->nssh[in1 - 1], nb = nsi * nsi;
    int pself = s->noff[ia - 1] + s->neigh_self[ia - 1] - 1;
    double *bi = wi + s->wboff[pself], *bb = wb + s->wboff[pself];
    for (int in = 0; in < s->neighn[ia - 1]; in++) {
      int p = s->noff[ia - 1] + in;
      int mb = s->neigh_b[p], ja = s->neigh_j[p];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      int in2 = s->imass[ja - 1];
      double z = fbo_distance(r1, r2);
      int self = (ia == ja && mb == 0);
      for (int isorp = 1; isorp <= t->nssh[in2 - 1]; isorp++) {
        double q = QIN(s, ja, isorp);
        int rc = fbo_getMEs_2c(t, in1, in2, P_RHOS_ATOM, isorp, z, nsi, nsi, bm);
        if (rc) return rc;
        for (int i = 0; i < nb; i++) bi[i] = bi[i] + bm[i] * q;
        if (self)
          for (int i = 0; i < nb; i++) bb[i] = bb[i] + bm[i] * q;
      }
    }
  }
  s->rdone[2] = 1;
  s->rdone[3] = 0;
  return FBO_OK;
}

/* assemble_rho_weighted_3c, M_assemble_rho_McWEDA_DOGS.f90:783-924 */
int fbo_assemble_rho_weighted_3c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn || !s->rdone[2]) return FBO_E_ORDER;
  double bm[16], eps[9];
  for (int ial = 1; ial <= s->natoms; ial++) {
    int indna = s->imass[ial - 1];
    for (int in = 0; in < s->ncommon[ial - 1]; in++) {
      int c = s->coff[ial - 1] + in;
      int mneigh = s->cm[c];
      if (mneigh == 0) continue;
      int ia = s->ci_j[c], ja = s->cj_j[c];
      int in1 = s->imass[ia - 1], in2 = s->imass[ja - 1];
      int nsi = t->nssh[in1 - 1], nsj = t->nssh[in2 - 1], nb = nsi * nsj;
      double x, z, cost;
      rho3_geometry(s, ial, c, &x, &z, &cost, eps);
      double *blk = s->wrho[0] + s->wboff[s->noff[ia - 1] + mneigh - 1];
      for (int isorp = 1; isorp <= t->nssh[indna - 1]; isorp++) {
        double q = QIN(s, ial, isorp);
        int rc = fbo_getMEs_3c(t, in1, in2, indna, P_RHOS_3C, isorp, x, z, nsi, nsj, cost, bm);
        if (rc) return rc;
        for (int i = 0; i < nb; i++) blk[i] = blk[i] + bm[i] * q;
      }
    }
  }
  s->rdone[3] = 1;
  return FBO_OK;
}

int64_t fbo_shell_block_len(const fbo_system *s) { return s->wlen; }

/* which: 0 rho_in, 1 rho_bond (orbital blocks, fbo_block_len doubles),
 *        2 rho_in_weighted, 3 rho_bond_weighted (shell blocks, fbo_shell_block_len) */
int fbo_get_rho(const fbo_system *s, int which, double *out) {
  if (which < 0 || which > 3) return FBO_E_ARG;
  if (which < 2) memcpy(out, s->rho2[which], sizeof(double) * s->blen);
  else memcpy(out, s->wrho[which - 2], sizeof(double) * s->wlen);
  return FBO_OK;
}

/* ========================================================== pseudopotential
 * c.SYSTEM/M_neighbors_PP.f90 (driver_neighbors_PP), b.FUNCTIONS/cl.f90,
 * c.SYSTEM/rotate_PP.f90, j.ASSEMBLERS/M_assemble_PP_2c.f90 (assemble_svnl,
 * assemble_vnl_2c) and j.ASSEMBLERS/M_assemble_PP_3c.f90 (assemble_vnl_3c,
 * mpairnay).  sVNL(mu, ncc) = <phi_mu(iatom) | Psi_ncc(jatom)>, the
 * Kleinman-Bylander projectors of jatom's PP shells; vnl(iatom, k) =
 * sum cl(ncc) sVNL(i, .) sVNL(j, .)^T over the projectors of the shared PP atom. */
#define P_VNL 6   /* include/interactions_2c.h:10 */

int fbo_tables_set_pp(fbo_tables *t, const int *nssh_pp, const int *lssh_pp, const double *rcutoff_pp,
                      const double *cl) {
  for (int i = 0; i < t->ns; i++) {
    if (nssh_pp[i] < 0 || nssh_pp[i] > FBO_MAXSH) return FBO_E_ARG;
    t->nssh_pp[i] = nssh_pp[i];
    t->norb_pp[i] = 0;
    for (int k = 0; k < nssh_pp[i]; k++) {
      t->lssh_pp[i][k] = lssh_pp[i * FBO_MAXSH + k];
      t->cl_pp[i][k] = cl[i * FBO_MAXSH + k];
      t->norb_pp[i] += 2 * t->lssh_pp[i][k] + 1;
    }
    t->rc_pp[i] = rcutoff_pp[i];
  }
  return FBO_OK;
}

/* cl values per projector orbital (b.FUNCTIONS/cl.f90:50-100) */
static void cl_values(const fbo_tables *t, int in, double *cl) {
  int iorb = 0;
  for (int issh = 0; issh < t->nssh_pp[in - 1]; issh++)
    for (int m = 0; m < 2 * t->lssh_pp[in - 1][issh] + 1; m++) cl[iorb++] = t->cl_pp[in - 1][issh];
}

/* c.SYSTEM/rotate_PP.f90: orbitals of in1 on the left, PP shells of in2 on the right */
void fbo_rotate_PP(const fbo_tables *t, int in1, int in2, const double *eps, int norb_mu, int norb_nu,
                   const double *M, double *X) {
  double dmat[25] = {0}, pmat[9], left[25], right[25];
  fbo_twister(eps, t->have_d, dmat, pmat);
  for (int i = 0; i < norb_mu * norb_nu; i++) X[i] = 0.0;
  int n1 = 0;
  for (int issh = 0; issh < t->nssh[in1 - 1]; issh++) {
    int l1 = t->lssh[in1 - 1][issh];
    chooser(l1, dmat, pmat, left);
    int n2 = 0;
    for (int jssh = 0; jssh < t->nssh_pp[in2 - 1]; jssh++) {
      int l2 = t->lssh_pp[in2 - 1][jssh];
      chooser(l2, dmat, pmat, right);
      for (int m2 = 1; m2 <= 2 * l2 + 1; m2++)
        for (int m1 = 1; m1 <= 2 * l1 + 1; m1++)
          for (int k2 = 1; k2 <= 2 * l2 + 1; k2++)
            for (int k1 = 1; k1 <= 2 * l1 + 1; k1++) {
              double *x = &X[(n2 + k2 - 1) * norb_mu + (n1 + k1 - 1)];
              *x = *x + left[5 * (m1 - 1) + (k1 - 1)] * M[(n2 + m2 - 1) * norb_mu + (n1 + m1 - 1)] *
                            right[5 * (m2 - 1) + (k2 - 1)];
            }
      n2 += 2 * l2 + 1;
    }
    n1 += 2 * l1 + 1;
  }
}

/* range list of find_neighbors_PP (mode 0: rc_i + rcPP_j) or find_neighbors_PPx
 * (mode 1: rcPP_i + rc_j), M_neighbors_PP.f90:409-516, 672-778 */
static void pp_range_list(fbo_system *s, int mode, int **n_out, int **off_out, int **j_out, int **b_out,
                          int **self_out, int *err) {
  const fbo_tables *t = s->t;
  int N = s->natoms;
  int *n = calloc(N, sizeof(int)), *off = calloc(N + 1, sizeof(int)), *self = malloc(sizeof(int) * N);
  int cap = 1024, tot = 0;
  int *jj = malloc(sizeof(int) * cap), *bb = malloc(sizeof(int) * cap);
  for (int ia = 1; ia <= N; ia++) {
    int in1 = s->imass[ia - 1];
    const double *r1 = RAT(s, ia);
    double rci = mode == 0 ? rcut_max(t, in1) : t->rc_pp[in1 - 1];
    for (int mb = 0; mb <= s->mbeta_max; mb++)
      for (int ja = 1; ja <= N; ja++) {
        int in2 = s->imass[ja - 1];
        double r2[3];
        for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
        double rcj = mode == 0 ? t->rc_pp[in2 - 1] : rcut_max(t, in2);
        double z = fbo_distance(r1, r2);
        double range = rci + rcj - 0.01;
        if (z < range) {
          if (tot == cap) { cap *= 2; jj = realloc(jj, sizeof(int) * cap); bb = realloc(bb, sizeof(int) * cap); }
          jj[tot] = ja; bb[tot] = mb; tot++;
          n[ia - 1]++;
        }
      }
    off[ia] = tot;
  }
  for (int ia = 1; ia <= N; ia++) {
    self[ia - 1] = -999;
    for (int k = 0; k < n[ia - 1]; k++)
      if (jj[off[ia - 1] + k] == ia && bb[off[ia - 1] + k] == 0) { self[ia - 1] = k + 1; break; }
    if (self[ia - 1] == -999) *err = FBO_E_NEIGH_SELF;
  }
  *n_out = n; *off_out = off; *j_out = jj; *b_out = bb; *self_out = self;
}

/* driver_neighbors_PP, M_neighbors_PP.f90:98-151 (find_neighbors_PP, _PPx,
 * _PPp with the maps and points, find_common_neighbors_PP :1301-1420) */
int fbo_build_pp_lists(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->neighn) return FBO_E_ORDER;
  int N = s->natoms, err = 0;
  pp_range_list(s, 0, &s->pp_n, &s->pp_off, &s->pp_j, &s->pp_b, &s->pp_self, &err);
  pp_range_list(s, 1, &s->px_n, &s->px_off, &s->px_j, &s->px_b, &s->px_self, &err);
  if (err) return err;
  /* neighbors_PPp: every (jatom, image of ibeta + jbeta) reached through a PP
   * neighbour ialpha of iatom and a PPx neighbour jatom of ialpha, first seen first */
  s->pq_n = calloc(N, sizeof(int));
  s->pq_off = calloc(N + 1, sizeof(int));
  s->pq_self = malloc(sizeof(int) * N);
  int cap = 1024, tot = 0;
  s->pq_j = malloc(sizeof(int) * cap);
  s->pq_b = malloc(sizeof(int) * cap);
  for (int ia = 1; ia <= N; ia++) {
    int base = tot, num = 0;
    for (int in = 0; in < s->pp_n[ia - 1]; in++) {
      int ial = s->pp_j[s->pp_off[ia - 1] + in], ib = s->pp_b[s->pp_off[ia - 1] + in];
      for (int jn = 0; jn < s->px_n[ial - 1]; jn++) {
        int ja = s->px_j[s->px_off[ial - 1] + jn], jb = s->px_b[s->px_off[ial - 1] + jn];
        double r3[3];
        for (int k = 0; k < 3; k++) r3[k] = XL(s, jb)[k] + XL(s, ib)[k];
        int mb;
        for (mb = 0; mb <= s->mbeta_max; mb++)
          if (fbo_distance(XL(s, mb), r3) < 1.0e-4) break;   /* none found: mbeta_max + 1, as the DO loop leaves it */
        int flag = 0;
        for (int k = 0; k < num; k++)
          if (s->pq_j[base + k] == ja && s->pq_b[base + k] == mb) flag = 1;
        if (!flag) {
          if (tot == cap) { cap *= 2; s->pq_j = realloc(s->pq_j, sizeof(int) * cap); s->pq_b = realloc(s->pq_b, sizeof(int) * cap); }
          s->pq_j[tot] = ja; s->pq_b[tot] = mb; tot++; num++;
        }
      }
    }
    s->pq_n[ia - 1] = num;
    s->pq_off[ia] = tot;
    s->pq_self[ia - 1] = -999;
    for (int k = 0; k < num; k++)
      if (s->pq_j[base + k] == ia && s->pq_b[base + k] == 0) { s->pq_self[ia - 1] = k + 1; break; }
    if (s->pq_self[ia - 1] == -999) return FBO_E_NEIGH_SELF;
  }
  /* maps into neighbors_PPp, and the PPx point (index of iatom among jatom's PP neighbours) */
  s->pp_map = calloc(s->pp_off[N] + 1, sizeof(int));
  s->px_map = calloc(s->px_off[N] + 1, sizeof(int));
  s->px_point = calloc(s->px_off[N] + 1, sizeof(int));
  for (int ia = 1; ia <= N; ia++) {
    for (int in = 0; in < s->pp_n[ia - 1]; in++) {
      int p = s->pp_off[ia - 1] + in;
      for (int k = 0; k < s->pq_n[ia - 1]; k++)
        if (s->pq_j[s->pq_off[ia - 1] + k] == s->pp_j[p] && s->pq_b[s->pq_off[ia - 1] + k] == s->pp_b[p]) {
          s->pp_map[p] = k + 1;
          break;
        }
    }
    for (int in = 0; in < s->px_n[ia - 1]; in++) {
      int p = s->px_off[ia - 1] + in;
      for (int k = 0; k < s->pq_n[ia - 1]; k++)
        if (s->pq_j[s->pq_off[ia - 1] + k] == s->px_j[p] && s->pq_b[s->pq_off[ia - 1] + k] == s->px_b[p]) {
          s->px_map[p] = k + 1;
          break;
        }
    }
  }
  for (int ia = 1; ia <= N; ia++) {
    const double *r1 = RAT(s, ia);
    for (int in = 0; in < s->px_n[ia - 1]; in++) {
      int p = s->px_off[ia - 1] + in;
      int ja = s->px_j[p], jb = s->px_b[p];
      for (int k = 0; k < s->pp_n[ja - 1]; k++) {
        int ka = s->pp_j[s->pp_off[ja - 1] + k], kb = s->pp_b[s->pp_off[ja - 1] + k];
        double r2[3];
        for (int q = 0; q < 3; q++) r2[q] = RAT(s, ka)[q] + XL(s, kb)[q] + XL(s, jb)[q];
        if (fbo_distance(r1, r2) < 1.0e-4) { s->px_point[p] = k + 1; break; }
      }
    }
  }
  /* common neighbours: ordered pairs of PPx neighbours of ialpha (neither
   * ialpha itself), kept when jatom is found among iatom's PPp neighbours */
  s->pc_n = calloc(N, sizeof(int));
  s->pc_off = calloc(N + 1, sizeof(int));
  cap = 1024; tot = 0;
  s->pc_ij = malloc(sizeof(int) * cap); s->pc_ib = malloc(sizeof(int) * cap);
  s->pc_jj = malloc(sizeof(int) * cap); s->pc_jb = malloc(sizeof(int) * cap); s->pc_m = malloc(sizeof(int) * cap);
  for (int ial = 1; ial <= N; ial++) {
    int num = 0;
    for (int in = 0; in < s->px_n[ial - 1]; in++) {
      int ia = s->px_j[s->px_off[ial - 1] + in], ib = s->px_b[s->px_off[ial - 1] + in];
      if (ia == ial && ib == 0) continue;
      double r1[3];
      for (int k = 0; k < 3; k++) r1[k] = RAT(s, ia)[k] + XL(s, ib)[k];
      for (int jn = 0; jn < s->px_n[ial - 1]; jn++) {
        int ja = s->px_j[s->px_off[ial - 1] + jn], jb = s->px_b[s->px_off[ial - 1] + jn];
        if ((ja == ial && jb == 0) || in == jn) continue;
        double r2[3], r[3];
        for (int k = 0; k < 3; k++) { r2[k] = RAT(s, ja)[k] + XL(s, jb)[k]; r[k] = r2[k] - r1[k]; }
        int m = -9999;
        for (int kn = 0; kn < s->pq_n[ia - 1]; kn++) {
          int ka = s->pq_j[s->pq_off[ia - 1] + kn], kb = s->pq_b[s->pq_off[ia - 1] + kn];
          if (kb > s->mbeta_max) continue;   /* an image outside xl: the reference would read past it */
          double r3[3];
          for (int k = 0; k < 3; k++) r3[k] = XL(s, kb)[k] + RAT(s, ka)[k] - RAT(s, ia)[k];
          if (fbo_distance(r, r3) < 1.0e-4) { m = kn + 1; break; }
        }
        if (m == -9999) continue;
        if (tot == cap) {
          cap *= 2;
          s->pc_ij = realloc(s->pc_ij, sizeof(int) * cap); s->pc_ib = realloc(s->pc_ib, sizeof(int) * cap);
          s->pc_jj = realloc(s->pc_jj, sizeof(int) * cap); s->pc_jb = realloc(s->pc_jb, sizeof(int) * cap);
          s->pc_m = realloc(s->pc_m, sizeof(int) * cap);
        }
        s->pc_ij[tot] = ia; s->pc_ib[tot] = ib; s->pc_jj[tot] = ja; s->pc_jb[tot] = jb; s->pc_m[tot] = m;
        tot++; num++;
      }
    }
    s->pc_n[ial - 1] = num;
    s->pc_off[ial] = tot;
  }
  /* block offsets: svnl per PP pair, vnl per PPp pair */
  s->sv_off = malloc(sizeof(int64_t) * (s->pp_off[N] + 1));
  int64_t o = 0;
  for (int ia = 1; ia <= N; ia++)
    for (int in = 0; in < s->pp_n[ia - 1]; in++) {
      int p = s->pp_off[ia - 1] + in;
      s->sv_off[p] = o;
      o += (int64_t)t->norb[s->imass[ia - 1] - 1] * t->norb_pp[s->imass[s->pp_j[p] - 1] - 1];
    }
  s->sv_off[s->pp_off[N]] = o;
  s->sv_len = o;
  s->svnl = calloc(o ? o : 1, sizeof(double));
  s->vl_off = malloc(sizeof(int64_t) * (s->pq_off[N] + 1));
  o = 0;
  for (int ia = 1; ia <= N; ia++)
    for (int in = 0; in < s->pq_n[ia - 1]; in++) {
      int p = s->pq_off[ia - 1] + in;
      s->vl_off[p] = o;
      o += (int64_t)t->norb[s->imass[ia - 1] - 1] * t->norb[s->imass[s->pq_j[p] - 1] - 1];
    }
  s->vl_off[s->pq_off[N]] = o;
  s->vl_len = o;
  s->vnl = calloc(o ? o : 1, sizeof(double));
  for (int i = 0; i < 3; i++) s->ppdone[i] = 0;
  return FBO_OK;
}

/* assemble_svnl, M_assemble_PP_2c.f90:95-207 */
int fbo_assemble_svnl(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->pp_n) return FBO_E_ORDER;
  double sm[81], eps[9];
  for (int ia = 1; ia <= s->natoms; ia++) {
    double r1[3] = {RAT(s, ia)[0], RAT(s, ia)[1], RAT(s, ia)[2]};
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    for (int in = 0; in < s->pp_n[ia - 1]; in++) {
      int p = s->pp_off[ia - 1] + in;
      int mb = s->pp_b[p], ja = s->pp_j[p], in2 = s->imass[ja - 1], nnu = t->norb_pp[in2 - 1];
      double r2[3];
      for (int k = 0; k < 3; k++) r2[k] = RAT(s, ja)[k] + XL(s, mb)[k];
      double z = pair_frame(s, r1, r2, eps);
      int rc = fbo_getMEs_2c(t, in1, in2, P_VNL, 0, z, nmu, nnu, sm);
      if (rc) return rc;
      fbo_rotate_PP(t, in1, in2, eps, nmu, nnu, sm, s->svnl + s->sv_off[p]);
    }
  }
  s->ppdone[0] = 1;
  s->ppdone[1] = s->ppdone[2] = 0;
  return FBO_OK;
}

/* PPx(imu, inu) = sum_ncc cl(ncc) * A(imu, ncc) * B(inu, ncc), loops inu, imu, ncc */
static void kb_contract(int nmu, int nnu, int ncc, const double *cl, const double *A, const double *B, double *X) {
  for (int inu = 0; inu < nnu; inu++)
    for (int imu = 0; imu < nmu; imu++) {
      double x = 0.0;
      for (int c = 0; c < ncc; c++) x = x + cl[c] * A[c * nmu + imu] * B[c * nnu + inu];
      X[inu * nmu + imu] = x;
    }
}

/* assemble_vnl_2c, M_assemble_PP_2c.f90:230-496 */
int fbo_assemble_vnl_2c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->ppdone[0]) return FBO_E_ORDER;
  memset(s->vnl, 0, sizeof(double) * (s->vl_len ? s->vl_len : 1));
  double X[81], cl[16];
  /* atom case: every PP neighbour's projectors into the PPp self block */
  for (int ia = 1; ia <= s->natoms; ia++) {
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    double *blk = s->vnl + s->vl_off[s->pq_off[ia - 1] + s->pq_self[ia - 1] - 1];
    for (int in = 0; in < s->pp_n[ia - 1]; in++) {
      int p = s->pp_off[ia - 1] + in, in2 = s->imass[s->pp_j[p] - 1];
      const double *sv = s->svnl + s->sv_off[p];
      cl_values(t, in2, cl);
      kb_contract(nmu, nmu, t->norb_pp[in2 - 1], cl, sv, sv, X);
      for (int i = 0; i < nmu * nmu; i++) blk[i] = blk[i] + X[i];
    }
  }
  /* ontop left: PPx neighbours, the projectors of iatom */
  for (int ia = 1; ia <= s->natoms; ia++) {
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    for (int in = 0; in < s->px_n[ia - 1]; in++) {
      int p = s->px_off[ia - 1] + in, ja = s->px_j[p], mb = s->px_b[p];
      int in2 = s->imass[ja - 1], nnu = t->norb[in2 - 1];
      if (ia == ja && mb == 0) {
        if (s->px_self[ia - 1] != in + 1) return FBO_E_NEIGH_SELF;
        continue;
      }
      cl_values(t, in1, cl);
      const double *sv1 = s->svnl + s->sv_off[s->pp_off[ia - 1] + s->pp_self[ia - 1] - 1];
      const double *sv2 = s->svnl + s->sv_off[s->pp_off[ja - 1] + s->px_point[p] - 1];
      kb_contract(nmu, nnu, t->norb_pp[in1 - 1], cl, sv1, sv2, X);
      double *blk = s->vnl + s->vl_off[s->pq_off[ia - 1] + s->px_map[p] - 1];
      for (int i = 0; i < nmu * nnu; i++) blk[i] = blk[i] + X[i];
    }
  }
  /* ontop right: PP neighbours, the projectors of jatom */
  for (int ia = 1; ia <= s->natoms; ia++) {
    int in1 = s->imass[ia - 1], nmu = t->norb[in1 - 1];
    for (int in = 0; in < s->pp_n[ia - 1]; in++) {
      int p = s->pp_off[ia - 1] + in, ja = s->pp_j[p], mb = s->pp_b[p];
      int in2 = s->imass[ja - 1], nnu = t->norb[in2 - 1];
      if (ia == ja && mb == 0) {
        if (s->pp_self[ia - 1] != in + 1) return FBO_E_NEIGH_SELF;
        continue;
      }
      cl_values(t, in2, cl);
      const double *sv1 = s->svnl + s->sv_off[p];
      const double *sv2 = s->svnl + s->sv_off[s->pp_off[ja - 1] + s->pp_self[ja - 1] - 1];
      kb_contract(nmu, nnu, t->norb_pp[in2 - 1], cl, sv1, sv2, X);
      double *blk = s->vnl + s->vl_off[s->pq_off[ia - 1] + s->pp_map[p] - 1];
      for (int i = 0; i < nmu * nnu; i++) blk[i] = blk[i] + X[i];
    }
  }
  s->ppdone[1] = 1;
  s->ppdone[2] = 0;
  return FBO_OK;
}

/* mpairnay, M_assemble_PP_3c.f90:266-364: the PP neighbour of iatom that is
 * jatom at offset r12 from iatom; exactly one match, else the reference stops */
static int mpairnay(const fbo_system *s, int ia, int ja, const double *r12, int *mpair) {
  const double *r1 = RAT(s, ia);
  int imatch = 0;
  *mpair = 0;
  for (int in = 0; in < s->pp_n[ia - 1]; in++) {
    int p = s->pp_off[ia - 1] + in;
    if (s->pp_j[p] != ja) continue;
    double r2[3], r21[3];
    for (int k = 0; k < 3; k++) { r2[k] = RAT(s, ja)[k] + XL(s, s->pp_b[p])[k]; r21[k] = r2[k] - r1[k]; }
    if (fbo_distance(r12, r21) < 0.0001) { imatch++; *mpair = in + 1; }
  }
  return imatch == 1 ? FBO_OK : FBO_E_NEIGH_BACK;
}

/* assemble_vnl_3c, M_assemble_PP_3c.f90:94-237 */
int fbo_assemble_vnl_3c(fbo_system *s) {
  const fbo_tables *t = s->t;
  if (!s->ppdone[1]) return FBO_E_ORDER;
  double X[81], cl[16];
  for (int ial = 1; ial <= s->natoms; ial++) {
    const double *rna = RAT(s, ial);
    int indna = s->imass[ial - 1];
    cl_values(t, indna, cl);
    for (int in = 0; in < s->pc_n[ial - 1]; in++) {
      int c = s->pc_off[ial - 1] + in, mneigh = s->pc_m[c];
      if (mneigh == 0) continue;
      int ia = s->pc_ij[c], ib = s->pc_ib[c], ja = s->pc_jj[c], jb = s->pc_jb[c];
      int nmu = t->norb[s->imass[ia - 1] - 1], nnu = t->norb[s->imass[ja - 1] - 1];
      double r1[3], r2[3], r31[3], r32[3];
      for (int k = 0; k < 3; k++) {
        r1[k] = RAT(s, ia)[k] + XL(s, ib)[k];
        r2[k] = RAT(s, ja)[k] + XL(s, jb)[k];
        r31[k] = rna[k] - r1[k];
        r32[k] = rna[k] - r2[k];
      }
      int m31, m32, rc;
      if ((rc = mpairnay(s, ia, ial, r31, &m31))) return rc;
      if ((rc = mpairnay(s, ja, ial, r32, &m32))) return rc;
      const double *sv1 = s->svnl + s->sv_off[s->pp_off[ia - 1] + m31 - 1];
      const double *sv2 = s->svnl + s->sv_off[s->pp_off[ja - 1] + m32 - 1];
      kb_contract(nmu, nnu, t->norb_pp[indna - 1], cl, sv1, sv2, X);
      double *blk = s->vnl + s->vl_off[s->pq_off[ia - 1] + mneigh - 1];
      for (int i = 0; i < nmu * nnu; i++) blk[i] = blk[i] + X[i];
    }
  }
  s->ppdone[2] = 1;
  return FBO_OK;
}

/* sizes: [0] PP pairs, [1] PPx pairs, [2] PPp pairs, [3] PP triplets, [4] svnl length, [5] vnl length */
void fbo_pp_sizes(const fbo_system *s, int64_t *out) {
  int N = s->natoms;
  out[0] = s->pp_off ? s->pp_off[N] : 0;
  out[1] = s->px_off ? s->px_off[N] : 0;
  out[2] = s->pq_off ? s->pq_off[N] : 0;
  out[3] = s->pc_off ? s->pc_off[N] : 0;
  out[4] = s->sv_len;
  out[5] = s->vl_len;
}

/* which: 0 PP (neighn, j, b, map), 1 PPx (neighn, j, b, map, point), 2 PPp (neighn, j, b),
 * 3 common (ncommon, iatom_j, iatom_b, jatom_j, jatom_b, mneigh); self: the *_self array */
int fbo_get_pp_list(const fbo_system *s, int which, int *n, int *a0, int *a1, int *a2, int *a3, int *a4,
                    int *self) {
  int N = s->natoms;
  const int *cnt[4] = {s->pp_n, s->px_n, s->pq_n, s->pc_n};
  const int *off[4] = {s->pp_off, s->px_off, s->pq_off, s->pc_off};
  if (which < 0 || which > 3 || !cnt[which]) return FBO_E_ARG;
  int tot = off[which][N];
  memcpy(n, cnt[which], sizeof(int) * N);
  const int *src[4][5] = {{s->pp_j, s->pp_b, s->pp_map, NULL, NULL},
                          {s->px_j, s->px_b, s->px_map, s->px_point, NULL},
                          {s->pq_j, s->pq_b, NULL, NULL, NULL},
                          {s->pc_ij, s->pc_ib, s->pc_jj, s->pc_jb, s->pc_m}};
  int *dst[5] = {a0, a1, a2, a3, a4};
  for (int k = 0; k < 5; k++)
    if (src[which][k] && dst[k]) memcpy(dst[k], src[which][k], sizeof(int) * tot);
  const int *sf[4] = {s->pp_self, s->px_self, s->pq_self, NULL};
  if (self && sf[which]) memcpy(self, sf[which], sizeof(int) * N);
  return FBO_OK;
}

/* which: 0 svnl (per PP pair), 1 vnl (per PPp pair) */
int fbo_get_pp_blocks(const fbo_system *s, int which, double *out) {
  if (which == 0 && s->svnl) memcpy(out, s->svnl, sizeof(double) * s->sv_len);
  else if (which == 1 && s->vnl) memcpy(out, s->vnl, sizeof(double) * s->vl_len);
  else return FBO_E_ARG;
  return FBO_OK;
}
