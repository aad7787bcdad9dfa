/*
 * oracle.c -- CPU restatement of SciHPC's projection-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.  The HIP
 * library in scihpc_amd/ never links or calls it.
 *
 * Written from the behaviour of the reference sources cited per function;
 * arithmetic expressions keep the reference's evaluation order so that the
 * compiled reference (oracle/_ref) and this file agree to the last bit where
 * the reference itself is deterministic.  Build: plain C11, no FMA
 * contraction (-ffp-contract=off), like the reference's default x86-64
 * build.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_MAXL 32
#define ORC_GHC 3

static const double orc_eps = 1e-8; /* global.h:14 */

/* ================================================================== */
/* hierarchy: multigrid.cpp:8-41 (min_grid = 3), :81-122               */
static const int orc_min_grid = 3;

int orc_mg_levels(int ndim, int nx, int ny, int nz, int maxl, int *lnx,
                  int *lny, int *lnz, int *ldeg) {
  int L = 0;
  lnx[0] = nx;
  lny[0] = ny;
  lnz[0] = (ndim == 3) ? nz : 1;
  ldeg[0] = 1;
  L = 1;
  while (L < maxl) {
    int cx = lnx[L - 1], cy = lny[L - 1], cz = lnz[L - 1];
    int r = 1;
    if (ndim == 2) {
      if (cx % 2 == 0 && cy % 2 == 0 && cx / 2 > orc_min_grid &&
          cy / 2 > orc_min_grid)
        r = 2;
      else if (cx % 3 == 0 && cy % 3 == 0 && cx / 3 > orc_min_grid &&
               cy / 3 > orc_min_grid)
        r = 3;
    } else {
      if (cx % 2 == 0 && cy % 2 == 0 && cz % 2 == 0 && cx / 2 > orc_min_grid &&
          cy / 2 > orc_min_grid && cz / 2 > orc_min_grid)
        r = 2;
      else if (cx % 3 == 0 && cy % 3 == 0 && cz % 3 == 0 &&
               cx / 3 > orc_min_grid && cy / 3 > orc_min_grid &&
               cz / 3 > orc_min_grid)
        r = 3;
    }
    if (r == 1) break;
    lnx[L] = cx / r;
    lny[L] = cy / r;
    lnz[L] = (ndim == 3) ? cz / r : 1;
    ldeg[L] = ldeg[L - 1] * r;
    L++;
  }
  return L;
}

/* OpenMP threads of the parallel loops (no-op in a serial build).  Only
 * loops whose iterations write disjoint cells run in parallel; every
 * reduction stays serial, so results do not depend on the thread count. */
int orc_set_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
  return omp_get_max_threads();
#else
  (void)n;
  return 1;
#endif
}

/* ================================================================== */
/* multigrid object                                                    */
typedef struct {
  int nx, ny, nz, n, deg;
  double dx, dy, dz;
  int nocompat;
  double *rhs, *sol, *res, *buf;
  /* operator: diag + coefficient to the +x/+y/+z neighbour (0 when absent);
   * the coefficient to the -x neighbour of cell c is cx[c-1]           */
  double *diag, *cx, *cy, *cz;
  /* diagonal used by the red-black algorithm: equal to diag except on
   * Dirichlet coarse levels, whose wall term keeps the level-0 boundary
   * position (see lvl_rb_diag) */
  double *diag_rb;
  /* coarsest-level dense factorisations (reference / red-black operator) */
  double *lu, *lu_rb;
  int *piv, *piv_rb;
  int lu_ready, lu_rb_ready;
} orc_lvl;

typedef struct {
  int ndim, bc, L, variable;
  int rkind; /* restriction: 0 child average, 1 full weighting (P^T/r^d) */
  int pkind; /* prolongation: 0 (bi/tri)linear, 1 piecewise constant     */
  orc_lvl lv[ORC_MAXL];
  /* flexible-CG vectors of level 0 (orc_mg_solve_fcg): x, r, p, q, p_old,
   * q_old, z, b; allocated on first use                                 */
  double *fcg[8];
} orc_mg;

void orc_mg_set_transfer(void *h, int rkind, int pkind) {
  ((orc_mg *)h)->rkind = rkind;
  ((orc_mg *)h)->pkind = pkind;
}

static void lvl_const_operator(orc_mg *m, orc_lvl *l) {
  /* init_NeumannBC (multigrid_base.cpp:236-291) and init_DirichletBC
   * (:293-345): off-diagonals 1/dx/dx; Neumann diag = -(sum of present
   * neighbours) in the order i-, i+, j-, j+, k-, k+; Dirichlet diag =
   * -2/dx/dx-2/dy/dy(-2/dz/dz).                                        */
  const int nx = l->nx, ny = l->ny, nz = l->nz;
  const double ax = 1.0 / l->dx / l->dx, ay = 1.0 / l->dy / l->dy;
  const double az = (m->ndim == 3) ? 1.0 / l->dz / l->dz : 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long c = i + (long)nx * (j + (long)ny * k);
        l->cx[c] = (i < nx - 1) ? ax : 0.0;
        l->cy[c] = (j < ny - 1) ? ay : 0.0;
        l->cz[c] = (m->ndim == 3 && k < nz - 1) ? az : 0.0;
        if (m->bc == 0) {
          double cc = 0.0;
          if (i > 0) cc += 1.0 / l->dx / l->dx;
          if (i < nx - 1) cc += 1.0 / l->dx / l->dx;
          if (j > 0) cc += 1.0 / l->dy / l->dy;
          if (j < ny - 1) cc += 1.0 / l->dy / l->dy;
          if (m->ndim == 3) {
            if (k > 0) cc += 1.0 / l->dz / l->dz;
            if (k < nz - 1) cc += 1.0 / l->dz / l->dz;
          }
          l->diag[c] = -cc;
        } else {
          if (m->ndim == 2)
            l->diag[c] = -2.0 / l->dx / l->dx - 2.0 / l->dy / l->dy;
          else
            l->diag[c] = -2.0 / l->dx / l->dx - 2.0 / l->dy / l->dy -
                         2.0 / l->dz / l->dz;
        }
      }
}

/* Red-black algorithm's diagonal.  The reference's Dirichlet stencil sets
 * the value at the ghost-cell centre to zero, i.e. the boundary sits at
 * distance h0 from the first level-0 cell centre.  A coarse level of
 * spacing H keeps that boundary when its wall term is 1/(d H) with
 * d = (H + h0)/2 (linear extrapolation to zero at the boundary), which is
 * 2/(H^2 (1 + h0/H)); on level 0 this is the reference's 1/h0^2.        */
static double rb_wall(double H, double h0) { return 2.0 / (H * H * (1.0 + h0 / H)); }

static void lvl_rb_diag(orc_mg *m, orc_lvl *l, double dx0, double dy0, double dz0) {
  const int nx = l->nx, ny = l->ny, nz = l->nz;
  if (m->bc != 1 || l->deg == 1) {
    memcpy(l->diag_rb, l->diag, sizeof(double) * l->n);
    return;
  }
  const double ax = 1.0 / l->dx / l->dx, ay = 1.0 / l->dy / l->dy;
  const double az = (m->ndim == 3) ? 1.0 / l->dz / l->dz : 0.0;
  const double wx = rb_wall(l->dx, dx0), wy = rb_wall(l->dy, dy0);
  const double wz = (m->ndim == 3) ? rb_wall(l->dz, dz0) : 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double cc = 0.0;
        cc += (i > 0) ? ax : wx;
        cc += (i < nx - 1) ? ax : wx;
        cc += (j > 0) ? ay : wy;
        cc += (j < ny - 1) ? ay : wy;
        if (m->ndim == 3) {
          cc += (k > 0) ? az : wz;
          cc += (k < nz - 1) ? az : wz;
        }
        l->diag_rb[i + (long)nx * (j + (long)ny * k)] = -cc;
      }
}

void *orc_mg_create(int ndim, int nx, int ny, int nz, double dx, double dy,
                    double dz, int bc) {
  orc_mg *m = (orc_mg *)calloc(1, sizeof(orc_mg));
  int lnx[ORC_MAXL], lny[ORC_MAXL], lnz[ORC_MAXL], ldeg[ORC_MAXL];
  m->ndim = ndim;
  m->bc = bc;
  m->L = orc_mg_levels(ndim, nx, ny, nz, ORC_MAXL, lnx, lny, lnz, ldeg);
  for (int l = 0; l < m->L; l++) {
    orc_lvl *v = &m->lv[l];
    v->nx = lnx[l];
    v->ny = lny[l];
    v->nz = lnz[l];
    v->deg = ldeg[l];
    v->n = v->nx * v->ny * v->nz;
    v->dx = dx * ldeg[l];
    v->dy = dy * ldeg[l];
    v->dz = (ndim == 3) ? dz * ldeg[l] : 1.0; /* multigrid_base.cpp:425-434 */
    v->nocompat = (bc == 1);
    v->rhs = (double *)calloc(v->n, sizeof(double));
    v->sol = (double *)calloc(v->n, sizeof(double));
    v->res = (double *)calloc(v->n, sizeof(double));
    v->buf = (double *)calloc(v->n, sizeof(double));
    v->diag = (double *)calloc(v->n, sizeof(double));
    v->cx = (double *)calloc(v->n, sizeof(double));
    v->cy = (double *)calloc(v->n, sizeof(double));
    v->cz = (double *)calloc(v->n, sizeof(double));
    v->diag_rb = (double *)calloc(v->n, sizeof(double));
    lvl_const_operator(m, v);
    lvl_rb_diag(m, v, dx, dy, dz);
  }
  return m;
}

void orc_mg_destroy(void *h) {
  orc_mg *m = (orc_mg *)h;
  if (!m) return;
  for (int l = 0; l < m->L; l++) {
    orc_lvl *v = &m->lv[l];
    free(v->rhs);
    free(v->sol);
    free(v->res);
    free(v->buf);
    free(v->diag);
    free(v->cx);
    free(v->cy);
    free(v->cz);
    free(v->lu);
    free(v->piv);
    free(v->diag_rb);
    free(v->lu_rb);
    free(v->piv_rb);
  }
  for (int q = 0; q < 8; q++) free(m->fcg[q]);
  free(m);
}

int orc_mg_nlevels(void *h) { return ((orc_mg *)h)->L; }

void orc_mg_level_shape(void *h, int level, int *shape) {
  orc_lvl *v = &((orc_mg *)h)->lv[level];
  shape[0] = v->nx;
  shape[1] = v->ny;
  shape[2] = v->nz;
  shape[3] = v->deg;
  shape[4] = v->n;
}

double *orc_mg_vec(void *h, int level, int which) {
  orc_lvl *v = &((orc_mg *)h)->lv[level];
  return which == 0 ? v->rhs : which == 1 ? v->sol : v->res;
}

void orc_mg_set_nocompat(void *h, int level, int flag) {
  ((orc_mg *)h)->lv[level].nocompat = flag;
}

/* diag from face arrays in the order of init_ppe_proj_coefficients:
 * CC = -(CR + CL + CU + CD); CC -= (CF + CB)   (projection_method.cpp:456-460) */
static void lvl_diag_from_faces(orc_mg *m, orc_lvl *l) {
  const int nx = l->nx, ny = l->ny, nz = l->nz;
  const long sxy = (long)nx * ny;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long c = i + (long)nx * (j + (long)ny * k);
        double CR = l->cx[c], CL = (i > 0) ? l->cx[c - 1] : 0.0;
        double CU = l->cy[c], CD = (j > 0) ? l->cy[c - nx] : 0.0;
        double cc = -(CR + CL + CU + CD);
        if (m->ndim == 3) {
          double CF = l->cz[c], CB = (k > 0) ? l->cz[c - sxy] : 0.0;
          cc -= (CF + CB);
        }
        l->diag[c] = cc;
      }
}

void orc_mg_set_rho(void *h, const double *rho) {
  orc_mg *m = (orc_mg *)h;
  m->variable = 1;
  /* level 0: projection_method.cpp:411-460 */
  orc_lvl *l = &m->lv[0];
  const int nx = l->nx, ny = l->ny, nz = l->nz;
  const long sxy = (long)nx * ny;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long c = i + (long)nx * (j + (long)ny * k);
        if (i < nx - 1) {
          double R = 0.5 * (rho[c] + rho[c + 1]);
          l->cx[c] = 1.0 / R / l->dx / l->dx;
        } else
          l->cx[c] = 0.0;
        if (j < ny - 1) {
          double U = 0.5 * (rho[c] + rho[c + nx]);
          l->cy[c] = 1.0 / U / l->dy / l->dy;
        } else
          l->cy[c] = 0.0;
        if (m->ndim == 3 && k < nz - 1) {
          double F = 0.5 * (rho[c] + rho[c + sxy]);
          l->cz[c] = 1.0 / F / l->dz / l->dz;
        } else
          l->cz[c] = 0.0;
      }
  lvl_diag_from_faces(m, l);
  memcpy(l->diag_rb, l->diag, sizeof(double) * l->n);
  /* coarse levels: mean of the fine faces on the coarse face / r^2      */
  for (int lv = 1; lv < m->L; lv++) {
    orc_lvl *f = &m->lv[lv - 1], *g = &m->lv[lv];
    const int r = g->deg / f->deg;
    const int rz = (m->ndim == 3) ? r : 1;
    const double scale = 1.0 / (double)(r * r);
    const double cnt = (m->ndim == 3) ? (double)(r * r) : (double)r;
    for (int K = 0; K < g->nz; K++)
      for (int J = 0; J < g->ny; J++)
        for (int I = 0; I < g->nx; I++) {
          long C = I + (long)g->nx * (J + (long)g->ny * K);
          double sx = 0.0, sy = 0.0, sz = 0.0;
          for (int b = 0; b < rz; b++)
            for (int a = 0; a < r; a++) {
              /* x face: fine i = (I+1)r-1, j = Jr+a, k = Kr+b */
              long fxc = ((I + 1) * r - 1) +
                         (long)f->nx * ((J * r + a) + (long)f->ny * (K * rz + b));
              sx += f->cx[fxc];
              long fyc = (I * r + a) +
                         (long)f->nx * (((J + 1) * r - 1) + (long)f->ny * (K * rz + b));
              sy += f->cy[fyc];
              if (m->ndim == 3) {
                long fzc = (I * r + a) +
                           (long)f->nx * ((J * r + b) + (long)f->ny * ((K + 1) * r - 1));
                sz += f->cz[fzc];
              }
            }
          g->cx[C] = (I < g->nx - 1) ? sx / cnt * scale : 0.0;
          g->cy[C] = (J < g->ny - 1) ? sy / cnt * scale : 0.0;
          g->cz[C] = (m->ndim == 3 && K < g->nz - 1) ? sz / cnt * scale : 0.0;
        }
    lvl_diag_from_faces(m, g);
    g->lu_ready = 0;
  }
  for (int lv = 0; lv < m->L; lv++) {
    memcpy(m->lv[lv].diag_rb, m->lv[lv].diag, sizeof(double) * m->lv[lv].n);
    m->lv[lv].lu_ready = 0;
    m->lv[lv].lu_rb_ready = 0;
  }
}

void orc_mg_get_operator(void *h, int level, double *diag, double *cx,
                         double *cy, double *cz) {
  orc_lvl *l = &((orc_mg *)h)->lv[level];
  memcpy(diag, l->diag, sizeof(double) * l->n);
  memcpy(cx, l->cx, sizeof(double) * l->n);
  memcpy(cy, l->cy, sizeof(double) * l->n);
  memcpy(cz, l->cz, sizeof(double) * l->n);
}

/* y = A x with the accumulation order of Eigen's column-major sparse
 * product (columns ascending): k-, j-, i-, diag, i+, j+, k+.            */
static void lvl_apply_d(const orc_lvl *l, const double *diag, const double *x, double *y) {
  const int nx = l->nx, ny = l->ny, nz = l->nz;
  const long sxy = (long)nx * ny;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long c = i + (long)nx * (j + (long)ny * k);
        double t = 0.0;
        if (k > 0 && l->cz[c - sxy] != 0.0) t += l->cz[c - sxy] * x[c - sxy];
        if (j > 0 && l->cy[c - nx] != 0.0) t += l->cy[c - nx] * x[c - nx];
        if (i > 0 && l->cx[c - 1] != 0.0) t += l->cx[c - 1] * x[c - 1];
        t += diag[c] * x[c];
        if (i < nx - 1 && l->cx[c] != 0.0) t += l->cx[c] * x[c + 1];
        if (j < ny - 1 && l->cy[c] != 0.0) t += l->cy[c] * x[c + nx];
        if (k < nz - 1 && l->cz[c] != 0.0) t += l->cz[c] * x[c + sxy];
        y[c] = t;
      }
}

static void lvl_apply(const orc_lvl *l, const double *x, double *y) { lvl_apply_d(l, l->diag, x, y); }

void orc_mg_apply(void *h, int level, const double *x, double *y) {
  lvl_apply(&((orc_mg *)h)->lv[level], x, y);
}

static double lvl_residual(orc_lvl *l) {
  /* multigrid_base.cpp:69-72: res = rhs - A*sol; return res.norm().  Eigen
   * evaluates this as res = rhs; res -= A*sol, accumulating the column-major
   * product column by column (ascending) into res.                      */
  const int nx = l->nx, ny = l->ny, nz = l->nz;
  const long sxy = (long)nx * ny;
  const double *x = l->sol;
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        long c = i + (long)nx * (j + (long)ny * k);
        double t = l->rhs[c];
        if (k > 0 && l->cz[c - sxy] != 0.0) t += l->cz[c - sxy] * -x[c - sxy];
        if (j > 0 && l->cy[c - nx] != 0.0) t += l->cy[c - nx] * -x[c - nx];
        if (i > 0 && l->cx[c - 1] != 0.0) t += l->cx[c - 1] * -x[c - 1];
        t += l->diag[c] * -x[c];
        if (i < nx - 1 && l->cx[c] != 0.0) t += l->cx[c] * -x[c + 1];
        if (j < ny - 1 && l->cy[c] != 0.0) t += l->cy[c] * -x[c + nx];
        if (k < nz - 1 && l->cz[c] != 0.0) t += l->cz[c] * -x[c + sxy];
        l->res[c] = t;
      }
  double s = 0.0;
  for (int c = 0; c < l->n; c++) s += l->res[c] * l->res[c];
  return sqrt(s);
}

double orc_mg_residual(void *h, int level) {
  return lvl_residual(&((orc_mg *)h)->lv[level]);
}

static void lvl_compat(orc_lvl *l, double *f) {
  /* multigrid_base.cpp:47-67 */
  if (l->nocompat) return;
  double sum = 0.0;
  for (int c = 0; c < l->n; c++) sum += f[c];
  sum = sum / l->n;
  for (int c = 0; c < l->n; c++) f[c] -= sum;
}

void orc_mg_compat(void *h, int level, int which) {
  orc_lvl *l = &((orc_mg *)h)->lv[level];
  lvl_compat(l, which == 0 ? l->rhs : which == 1 ? l->sol : l->res);
}

static void lvl_relax(orc_lvl *l, int iter) {
  /* multigrid_base.cpp:16-45: steepest descent on the negative definite A */
  for (int cnt = 0; cnt < iter; ++cnt) {
    if (lvl_residual(l) < 1e-16) break;
    double rdot = 0.0;
    for (int c = 0; c < l->n; c++) rdot += l->res[c] * l->res[c];
    lvl_apply(l, l->res, l->buf);
    double alpha = 0.0;
    for (int c = 0; c < l->n; c++) alpha += l->res[c] * l->buf[c];
    alpha = rdot / alpha;
#pragma omp parallel for schedule(static)
    for (int c = 0; c < l->n; c++) l->sol[c] += alpha * l->res[c];
    lvl_compat(l, l->sol);
  }
}

void orc_mg_relax(void *h, int level, int iter) {
  lvl_relax(&((orc_mg *)h)->lv[level], iter);
}

static const int orc_base_step = 10; /* multigrid_base.h:35 */

void orc_mg_restriction(void *h, int level) {
  /* multigrid_base.cpp:74-113 */
  orc_mg *m = (orc_mg *)h;
  orc_lvl *f = &m->lv[level], *g = &m->lv[level + 1];
  lvl_relax(f, orc_base_step);
  lvl_residual(f);
  const int r = g->deg / f->deg;
  if (m->ndim == 2) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < g->nx; ++i)
      for (int j = 0; j < g->ny; ++j) {
        long C = i + (long)g->nx * j;
        g->rhs[C] = 0.0;
        for (int ii = 0; ii < r; ++ii)
          for (int jj = 0; jj < r; ++jj)
            g->rhs[C] += f->res[(i * r + ii) + (long)f->nx * (j * r + jj)];
        g->rhs[C] /= r * r;
        g->sol[C] = 0.0;
      }
  } else {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < g->nx; ++i)
      for (int j = 0; j < g->ny; ++j)
        for (int k = 0; k < g->nz; ++k) {
          long C = i + (long)g->nx * (j + (long)g->ny * k);
          g->rhs[C] = 0.0;
          for (int ii = 0; ii < r; ++ii)
            for (int jj = 0; jj < r; ++jj)
              for (int kk = 0; kk < r; ++kk)
                g->rhs[C] += f->res[(i * r + ii) +
                                    (long)f->nx * ((j * r + jj) +
                                                   (long)f->ny * (k * r + kk))];
          g->rhs[C] /= r * r * r;
          g->sol[C] = 0.0;
        }
  }
  lvl_compat(g, g->rhs);
}

void orc_mg_prolongation(void *h, int level) {
  /* multigrid_base.cpp:115-208: slope reconstruction added to the fine sol */
  orc_mg *m = (orc_mg *)h;
  orc_lvl *c = &m->lv[level], *d = &m->lv[level - 1];
  const int r = c->deg / d->deg;
  const int nx = c->nx, ny = c->ny, nz = c->nz;
  const long sxy = (long)nx * ny;
  const double *s = c->sol;
  if (m->ndim == 2) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j) {
        long o = i + (long)nx * j;
        double fx, fy;
        if (i > 0 && i < nx - 1)
          fx = 0.5 * (s[o + 1] - s[o - 1]) / c->dx;
        else if (i == 0)
          fx = (s[o + 1] - s[o]) / c->dx;
        else
          fx = (s[o] - s[o - 1]) / c->dx;
        if (j > 0 && j < ny - 1)
          fy = 0.5 * (s[o + nx] - s[o - nx]) / c->dy;
        else if (j == 0)
          fy = (s[o + nx] - s[o]) / c->dy;
        else
          fy = (s[o] - s[o - nx]) / c->dy;
        for (int ii = 0; ii < r; ++ii)
          for (int jj = 0; jj < r; ++jj) {
            double x = (0.5 + ii) * d->dx;
            double y = (0.5 + jj) * d->dy;
            d->sol[(i * r + ii) + (long)d->nx * (j * r + jj)] +=
                s[o] + fx * (x - 0.5 * c->dx) + fy * (y - 0.5 * c->dy);
          }
      }
  } else {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          long o = i + (long)nx * (j + (long)ny * k);
          double fx, fy, fz;
          if (i > 0 && i < nx - 1)
            fx = 0.5 * (s[o + 1] - s[o - 1]) / c->dx;
          else if (i == 0)
            fx = (s[o + 1] - s[o]) / c->dx;
          else
            fx = (s[o] - s[o - 1]) / c->dx;
          if (j > 0 && j < ny - 1)
            fy = 0.5 * (s[o + nx] - s[o - nx]) / c->dy;
          else if (j == 0)
            fy = (s[o + nx] - s[o]) / c->dy;
          else
            fy = (s[o] - s[o - nx]) / c->dy;
          if (k > 0 && k < nz - 1)
            fz = 0.5 * (s[o + sxy] - s[o - sxy]) / c->dz;
          else if (k == 0)
            fz = (s[o + sxy] - s[o]) / c->dz;
          else
            fz = (s[o] - s[o - sxy]) / c->dz;
          for (int ii = 0; ii < r; ++ii)
            for (int jj = 0; jj < r; ++jj)
              for (int kk = 0; kk < r; ++kk) {
                double x = (0.5 + ii) * d->dx;
                double y = (0.5 + jj) * d->dy;
                double z = (0.5 + kk) * d->dz;
                d->sol[(i * r + ii) +
                       (long)d->nx * ((j * r + jj) + (long)d->ny * (k * r + kk))] +=
                    s[o] + fx * (x - 0.5 * c->dx) + fy * (y - 0.5 * c->dy) +
                    fz * (z - 0.5 * c->dz);
              }
        }
  }
  lvl_relax(d, orc_base_step);
}

/* ---------------- dense LU with partial pivoting ------------------ */
static int dense_lu(double *a, int *piv, int n) {
  for (int k = 0; k < n; k++) {
    int p = k;
    double best = fabs(a[(long)k * n + k]);
    for (int i = k + 1; i < n; i++) {
      double v = fabs(a[(long)i * n + k]);
      if (v > best) {
        best = v;
        p = i;
      }
    }
    piv[k] = p;
    if (best == 0.0) return -1;
    if (p != k)
      for (int j = 0; j < n; j++) {
        double t = a[(long)k * n + j];
        a[(long)k * n + j] = a[(long)p * n + j];
        a[(long)p * n + j] = t;
      }
    double inv = 1.0 / a[(long)k * n + k];
    for (int i = k + 1; i < n; i++) {
      double f = a[(long)i * n + k] * inv;
      a[(long)i * n + k] = f;
      if (f != 0.0)
        for (int j = k + 1; j < n; j++) a[(long)i * n + j] -= f * a[(long)k * n + j];
    }
  }
  return 0;
}

static void dense_lu_solve(const double *a, const int *piv, int n,
                           const double *b, double *x) {
  for (int i = 0; i < n; i++) x[i] = b[i];
  for (int k = 0; k < n; k++)
    if (piv[k] != k) {
      double t = x[k];
      x[k] = x[piv[k]];
      x[piv[k]] = t;
    }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i; j++) x[i] -= a[(long)i * n + j] * x[j];
  for (int i = n - 1; i >= 0; i--) {
    for (int j = i + 1; j < n; j++) x[i] -= a[(long)i * n + j] * x[j];
    x[i] /= a[(long)i * n + i];
  }
}

/* Dense matrix of the level operator; for the all-Neumann (singular)
 * operator the constant null space is pinned: M = A - (s/n) 1 1^T with
 * s = max|diag|, which equals A on mean-free vectors and is nonsingular. */
static void lvl_dense_matrix(const orc_mg *m, const orc_lvl *l, const double *diag, double *a) {
  const int n = l->n;
  double *e = (double *)calloc(n, sizeof(double));
  double *col = (double *)calloc(n, sizeof(double));
  for (int j = 0; j < n; j++) {
    e[j] = 1.0;
    lvl_apply_d(l, diag, e, col);
    e[j] = 0.0;
    for (int i = 0; i < n; i++) a[(long)i * n + j] = col[i];
  }
  if (m->bc == 0) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s = fmax(s, fabs(l->diag[i]));
    for (long i = 0; i < (long)n * n; i++) a[i] -= s / n;
  }
  free(e);
  free(col);
}

static void lvl_factor(const orc_mg *m, orc_lvl *l) {
  if (l->lu_ready) return;
  free(l->lu);
  free(l->piv);
  l->lu = (double *)malloc(sizeof(double) * (long)l->n * l->n);
  l->piv = (int *)malloc(sizeof(int) * l->n);
  lvl_dense_matrix(m, l, l->diag, l->lu);
  dense_lu(l->lu, l->piv, l->n);
  l->lu_ready = 1;
}

static void lvl_factor_rb(const orc_mg *m, orc_lvl *l) {
  if (l->lu_rb_ready) return;
  free(l->lu_rb);
  free(l->piv_rb);
  l->lu_rb = (double *)malloc(sizeof(double) * (long)l->n * l->n);
  l->piv_rb = (int *)malloc(sizeof(int) * l->n);
  lvl_dense_matrix(m, l, l->diag_rb, l->lu_rb);
  dense_lu(l->lu_rb, l->piv_rb, l->n);
  l->lu_rb_ready = 1;
}

void orc_mg_coarse_solve(void *h) {
  /* multigrid_base.cpp:210-234: factorise once, solve, compatibility,
   * iterative refinement while the residual drops by more than 10 %.    */
  orc_mg *m = (orc_mg *)h;
  orc_lvl *l = &m->lv[m->L - 1];
  lvl_factor(m, l);
  dense_lu_solve(l->lu, l->piv, l->n, l->rhs, l->sol);
  lvl_compat(l, l->sol);
  double old_res = -1.0, new_res = 0.0;
  double *dx = (double *)malloc(sizeof(double) * l->n);
  do {
    if (old_res < 0.0)
      old_res = lvl_residual(l);
    else
      old_res = new_res;
    dense_lu_solve(l->lu, l->piv, l->n, l->res, dx);
    for (int c = 0; c < l->n; c++) l->sol[c] += dx[c];
    lvl_compat(l, l->sol);
    new_res = lvl_residual(l);
  } while (1.0 - new_res / old_res > 0.1);
  free(dx);
}

void orc_mg_vcycle_ref(void *h) {
  /* multigrid.cpp:43-55 */
  orc_mg *m = (orc_mg *)h;
  for (int i = 0; i < m->L - 1; ++i) orc_mg_restriction(h, i);
  orc_mg_coarse_solve(h);
  for (int i = m->L - 1; i > 0; --i) orc_mg_prolongation(h, i);
}

void orc_mg_fullcycle_ref(void *h) {
  /* multigrid.cpp:57-79; the inner prolongation loop at :66 never runs
   * (i starts at L-1 and requires i < L-1-iter), so it is omitted.     */
  orc_mg *m = (orc_mg *)h;
  const int L = m->L;
  for (int i = 0; i < L - 1; ++i) orc_mg_restriction(h, i);
  orc_mg_coarse_solve(h);
  for (int iter = 1; iter < L; ++iter) {
    for (int i = L - 1 - iter; i < L - 1; ++i) orc_mg_restriction(h, i);
    orc_mg_coarse_solve(h);
  }
  for (int i = L - 1; i > 0; --i) orc_mg_prolongation(h, i);
}

/* ================================================================== */
/* build's algorithm: red-black GS V-cycle                             */
static void lvl_gs_color(const orc_lvl *l, int color) {
  const int nx = l->nx, ny = l->ny, nz = l->nz;
  const long sxy = (long)nx * ny;
  double *x = l->sol;
  /* cells of one colour only couple to the other colour: rows in parallel */
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = (color + j + k) & 1; i < nx; i += 2) {
        long c = i + (long)nx * (j + (long)ny * k);
        double t = 0.0;
        if (i > 0) t += l->cx[c - 1] * x[c - 1];
        if (i < nx - 1) t += l->cx[c] * x[c + 1];
        if (j > 0) t += l->cy[c - nx] * x[c - nx];
        if (j < ny - 1) t += l->cy[c] * x[c + nx];
        if (k > 0) t += l->cz[c - sxy] * x[c - sxy];
        if (k < nz - 1) t += l->cz[c] * x[c + sxy];
        x[c] = (l->rhs[c] - t) / l->diag_rb[c];
      }
}

void orc_mg_smooth_rb(void *h, int level, int sweeps, int black_first) {
  orc_lvl *l = &((orc_mg *)h)->lv[level];
  for (int s = 0; s < sweeps; s++) {
    lvl_gs_color(l, black_first ? 1 : 0);
    lvl_gs_color(l, black_first ? 0 : 1);
  }
}

/* weight of coarse cell C in the linear interpolation to fine cell i
 * (transpose of interp_axis below, same ghost rule)                    */
static double interp_weight(int i, int r, int C, int nc, int dirichlet);

/* full weighting: coarse rhs = (1/r^d) P^T res, P the linear prolongation
 * of orc_mg_prolong_rb (so the V-cycle is symmetric).                   */
void orc_mg_restrict_rb(void *h, int level) {
  orc_mg *m = (orc_mg *)h;
  orc_lvl *f = &m->lv[level], *g = &m->lv[level + 1];
  lvl_apply_d(f, f->diag_rb, f->sol, f->buf);
#pragma omp parallel for schedule(static)
  for (int c = 0; c < f->n; c++) f->res[c] = f->rhs[c] - f->buf[c];
  const int r = g->deg / f->deg;
  const int dir = m->bc == 1;
  const double inv = 1.0 / (double)(m->ndim == 3 ? r * r * r : r * r);
  if (m->rkind == 0) {
    const int rz = (m->ndim == 3) ? r : 1;
#pragma omp parallel for collapse(2) schedule(static)
    for (int K = 0; K < g->nz; K++)
      for (int J = 0; J < g->ny; J++)
        for (int I = 0; I < g->nx; I++) {
          long C = I + (long)g->nx * (J + (long)g->ny * K);
          double s = 0.0;
          for (int kk = 0; kk < rz; kk++)
            for (int jj = 0; jj < r; jj++)
              for (int ii = 0; ii < r; ii++)
                s += f->res[(I * r + ii) + (long)f->nx * ((J * r + jj) + (long)f->ny * (K * rz + kk))];
          g->rhs[C] = s * inv;
          g->sol[C] = 0.0;
        }
    return;
  }
  for (int K = 0; K < g->nz; K++)
    for (int J = 0; J < g->ny; J++)
      for (int I = 0; I < g->nx; I++) {
        long C = I + (long)g->nx * (J + (long)g->ny * K);
        int k0 = 0, k1 = 0;
        if (m->ndim == 3) {
          k0 = K * r - r;
          k1 = K * r + 2 * r - 1;
          if (k0 < 0) k0 = 0;
          if (k1 > f->nz - 1) k1 = f->nz - 1;
        }
        int j0 = J * r - r, j1 = J * r + 2 * r - 1;
        int i0 = I * r - r, i1 = I * r + 2 * r - 1;
        if (j0 < 0) j0 = 0;
        if (j1 > f->ny - 1) j1 = f->ny - 1;
        if (i0 < 0) i0 = 0;
        if (i1 > f->nx - 1) i1 = f->nx - 1;
        double s = 0.0;
        for (int k = k0; k <= k1; k++) {
          double wz = (m->ndim == 3) ? interp_weight(k, r, K, g->nz, dir) : 1.0;
          if (wz == 0.0) continue;
          for (int j = j0; j <= j1; j++) {
            double wy = interp_weight(j, r, J, g->ny, dir);
            if (wy == 0.0) continue;
            for (int i = i0; i <= i1; i++) {
              double wx = interp_weight(i, r, I, g->nx, dir);
              if (wx == 0.0) continue;
              s += (wx * wy * wz) * f->res[i + (long)f->nx * (j + (long)f->ny * k)];
            }
          }
        }
        g->rhs[C] = s * inv;
        g->sol[C] = 0.0;
      }
}

/* Cell-centred linear interpolation weight along one axis for fine index
 * i with ratio r: coarse position X = (2i+1-r)/(2r); I0 = floor(X),
 * t = X - I0 in exact integer arithmetic.                               */
static void interp_axis(int i, int r, int *I0, double *t) {
  int num = 2 * i + 1 - r, den = 2 * r;
  int q = (num >= 0) ? num / den : -((-num + den - 1) / den);
  *I0 = q;
  *t = (double)(num - q * den) / (double)den;
}

static double interp_weight(int i, int r, int C, int nc, int dirichlet) {
  int I0;
  double t, w = 0.0;
  interp_axis(i, r, &I0, &t);
  int a = I0, b = I0 + 1;
  if (!dirichlet) {
    if (a < 0) a = 0;
    if (a > nc - 1) a = nc - 1;
    if (b < 0) b = 0;
    if (b > nc - 1) b = nc - 1;
  }
  if (a == C) w += 1.0 - t;
  if (b == C) w += t;
  return w;
}

static double coarse_at(const orc_mg *m, const orc_lvl *g, int I, int J, int K) {
  /* ghost rule: Neumann -> nearest interior value; Dirichlet -> 0        */
  int out = (I < 0 || I >= g->nx || J < 0 || J >= g->ny || K < 0 || K >= g->nz);
  if (out && m->bc == 1) return 0.0;
  if (I < 0) I = 0;
  if (I >= g->nx) I = g->nx - 1;
  if (J < 0) J = 0;
  if (J >= g->ny) J = g->ny - 1;
  if (K < 0) K = 0;
  if (K >= g->nz) K = g->nz - 1;
  return g->sol[I + (long)g->nx * (J + (long)g->ny * K)];
}

void orc_mg_prolong_rb(void *h, int level) {
  orc_mg *m = (orc_mg *)h;
  orc_lvl *g = &m->lv[level], *f = &m->lv[level - 1];
  const int r = g->deg / f->deg;
  if (m->pkind == 1) {
    for (int k = 0; k < f->nz; k++)
      for (int j = 0; j < f->ny; j++)
        for (int i = 0; i < f->nx; i++)
          f->sol[i + (long)f->nx * (j + (long)f->ny * k)] +=
              g->sol[i / r + (long)g->nx * (j / r + (long)g->ny * (m->ndim == 3 ? k / r : 0))];
    return;
  }
#pragma omp parallel for collapse(2) schedule(static)
  for (int k = 0; k < f->nz; k++)
    for (int j = 0; j < f->ny; j++)
      for (int i = 0; i < f->nx; i++) {
        int I0, J0, K0 = 0;
        double tx, ty, tz = 0.0;
        interp_axis(i, r, &I0, &tx);
        interp_axis(j, r, &J0, &ty);
        if (m->ndim == 3) interp_axis(k, r, &K0, &tz);
        double v;
        double v00 = (1.0 - tx) * coarse_at(m, g, I0, J0, K0) + tx * coarse_at(m, g, I0 + 1, J0, K0);
        double v10 = (1.0 - tx) * coarse_at(m, g, I0, J0 + 1, K0) + tx * coarse_at(m, g, I0 + 1, J0 + 1, K0);
        double v0 = (1.0 - ty) * v00 + ty * v10;
        if (m->ndim == 3) {
          double v01 = (1.0 - tx) * coarse_at(m, g, I0, J0, K0 + 1) + tx * coarse_at(m, g, I0 + 1, J0, K0 + 1);
          double v11 = (1.0 - tx) * coarse_at(m, g, I0, J0 + 1, K0 + 1) + tx * coarse_at(m, g, I0 + 1, J0 + 1, K0 + 1);
          double v1 = (1.0 - ty) * v01 + ty * v11;
          v = (1.0 - tz) * v0 + tz * v1;
        } else
          v = v0;
        f->sol[i + (long)f->nx * (j + (long)f->ny * k)] += v;
      }
}

void orc_mg_coarse_direct(void *h) {
  orc_mg *m = (orc_mg *)h;
  orc_lvl *l = &m->lv[m->L - 1];
  lvl_factor_rb(m, l);
  dense_lu_solve(l->lu_rb, l->piv_rb, l->n, l->rhs, l->sol);
  if (m->bc == 0) {
    double s = 0.0;
    for (int c = 0; c < l->n; c++) s += l->sol[c];
    s /= l->n;
    for (int c = 0; c < l->n; c++) l->sol[c] -= s;
  }
}

void orc_mg_vcycle_rb(void *h, int nu1, int nu2) {
  orc_mg *m = (orc_mg *)h;
  for (int l = 0; l < m->L - 1; l++) {
    orc_mg_smooth_rb(h, l, nu1, 0);
    orc_mg_restrict_rb(h, l);
  }
  orc_mg_coarse_direct(h);
  for (int l = m->L - 1; l > 0; l--) {
    orc_mg_prolong_rb(h, l);
    orc_mg_smooth_rb(h, l - 1, nu2, 1);
  }
}

static double vec_norm(const double *v, long n) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s += v[i] * v[i];
  return sqrt(s);
}

int orc_mg_solve_rb(void *h, double rel_tol, int max_cycles, int nu1, int nu2,
                    double *hist) {
  orc_mg *m = (orc_mg *)h;
  orc_lvl *l = &m->lv[0];
  double bn = vec_norm(l->rhs, l->n);
  if (bn == 0.0) bn = 1.0;
  int c = 0;
  double rr = lvl_residual(l) / bn;
  while (rr > rel_tol && c < max_cycles) {
    orc_mg_vcycle_rb(h, nu1, nu2);
    rr = lvl_residual(l) / bn;
    if (hist) hist[c] = rr;
    c++;
  }
  if (m->bc == 0) {
    double s = 0.0;
    for (int i = 0; i < l->n; i++) s += l->sol[i];
    s /= l->n;
    for (int i = 0; i < l->n; i++) l->sol[i] -= s;
  }
  return c;
}

static double vec_dot(const double *a, const double *b, long n) {
  double s = 0.0;
  for (long i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

/* Flexible CG, FCG(1) (Notay, SIAM J. Sci. Comput. 22 (2000) 1444-1460),
 * preconditioned by one red-black V(nu1,nu2) from zero on level 0:
 *   x = 0, r = b
 *   repeat: z = V(r); p = z (first) or z + beta p_old with
 *           beta = -(z.q_old)/(p_old.q_old); q = A p;
 *           alpha = (p.r)/(p.q); x += alpha p; r -= alpha q
 *   until ||r||/||b|| <= rel_tol (the recurrence residual) or max_iter.
 * On the singular all-Neumann operator solved below 1e-9, p.q and p.r are
 * taken with p's constant component removed, (p - mean p).q and
 * (p - mean p).r (A maps constants to zero; left in, that component grows
 * at the FP64 floor and its product with the rounding-level sum of r drives
 * the iteration away); the corrected p.q is also the next beta's
 * denominator.  When the
 * recurrence says converged, the true residual b - A x is formed; above
 * rel_tol with iterations left, the iteration restarts from x with
 * r = b - A x.  beta and alpha are 0 when their denominator is exactly 0
 * (only after exact convergence).  hist[k] = the relative residual after
 * iteration k (the recurrence's).  Neumann: the result is made mean-free.
 * b is level 0's rhs (restored on return), the result goes to level 0's
 * sol.  The build's outer loop (no reference counterpart; the closest
 * analogue is the FPC loop, projection_method.cpp:694-706).             */
int orc_mg_solve_fcg(void *h, double rel_tol, int max_iter, int nu1, int nu2,
                     double *hist) {
  orc_mg *m = (orc_mg *)h;
  orc_lvl *l = &m->lv[0];
  const long n = l->n;
  const int corr = m->bc == 0 && rel_tol < 1e-9;
  for (int q = 0; q < 8; q++)
    if (!m->fcg[q]) m->fcg[q] = (double *)calloc(n, sizeof(double));
  double *x = m->fcg[0], *r = m->fcg[1], *p = m->fcg[2], *q = m->fcg[3];
  double *po = m->fcg[4], *qo = m->fcg[5], *z = m->fcg[6], *b = m->fcg[7];
  memcpy(b, l->rhs, sizeof(double) * n);
  memcpy(r, b, sizeof(double) * n);
  memset(x, 0, sizeof(double) * n);
  double bn = vec_norm(b, n);
  const double scale = bn > 0.0 ? bn : 1.0;
  double rr = bn > 0.0 ? 1.0 : 0.0;
  double pq_prev = 0.0;
  int it = 0, first = 1;
  for (;;) {
    while (rr > rel_tol && it < max_iter) {
      memcpy(l->rhs, r, sizeof(double) * n);
      memset(l->sol, 0, sizeof(double) * n);
      orc_mg_vcycle_rb(h, nu1, nu2);
      memcpy(z, l->sol, sizeof(double) * n);
      if (first) {
        memcpy(p, z, sizeof(double) * n);
      } else {
        memcpy(po, p, sizeof(double) * n);
        memcpy(qo, q, sizeof(double) * n);
        const double beta = pq_prev != 0.0 ? -vec_dot(z, qo, n) / pq_prev : 0.0;
        for (long c = 0; c < n; c++) p[c] = z[c] + beta * po[c];
      }
      first = 0;
      lvl_apply(l, p, q);
      double pq = vec_dot(p, q, n), pr = vec_dot(p, r, n);
      if (corr) {
        double sp = 0.0, sr = 0.0, sq = 0.0;
        for (long c = 0; c < n; c++) {
          sp += p[c];
          sr += r[c];
          sq += q[c];
        }
        pq -= sp * sq / (double)n;
        pr -= sp * sr / (double)n;
      }
      pq_prev = pq;
      const double alpha = pq != 0.0 ? pr / pq : 0.0;
      for (long c = 0; c < n; c++) {
        x[c] += alpha * p[c];
        r[c] -= alpha * q[c];
      }
      rr = vec_norm(r, n) / scale;
      if (hist) hist[it] = rr;
      it++;
    }
    /* the true residual of x */
    const int recurrence_converged = rr <= rel_tol;
    lvl_apply(l, x, z);
    for (long c = 0; c < n; c++) z[c] = b[c] - z[c];
    rr = vec_norm(z, n) / scale;
    if (!recurrence_converged || rr <= rel_tol || it >= max_iter) break;
    memcpy(r, z, sizeof(double) * n);
    first = 1;
  }
  memcpy(l->rhs, b, sizeof(double) * n);
  memcpy(l->sol, x, sizeof(double) * n);
  if (m->bc == 0) {
    double s = 0.0;
    for (long i = 0; i < n; i++) s += l->sol[i];
    s /= (double)n;
    for (long i = 0; i < n; i++) l->sol[i] -= s;
  }
  return it;
}

/* the flexible-CG vectors after orc_mg_solve_fcg: 0 x (before the mean
 * removal), 1 r, 2 p, 3 q, 4 p_old, 5 q_old, 6 b - A x of the final x */
double *orc_mg_fcg_vec(void *h, int which) {
  orc_mg *m = (orc_mg *)h;
  return (which >= 0 && which < 8) ? m->fcg[which] : NULL;
}

int orc_mg_dense_solve_level0(void *h, const double *b, double *x) {
  orc_mg *m = (orc_mg *)h;
  orc_lvl *l = &m->lv[0];
  if (l->n > 4096) return -1;
  double *a = (double *)malloc(sizeof(double) * (long)l->n * l->n);
  int *piv = (int *)malloc(sizeof(int) * l->n);
  lvl_dense_matrix(m, l, l->diag, a);
  int rc = dense_lu(a, piv, l->n);
  if (rc == 0) {
    dense_lu_solve(a, piv, l->n, b, x);
    if (m->bc == 0) {
      double s = 0.0;
      for (int i = 0; i < l->n; i++) s += x[i];
      s /= l->n;
      for (int i = 0; i < l->n; i++) x[i] -= s;
    }
  }
  free(a);
  free(piv);
  return rc;
}

/* ================================================================== */
/* padded fields                                                       */
typedef struct {
  int ndim, n[3], N[3], off[3];
  long s[3];
} orc_pf;

static orc_pf pf_make(int ndim, int nx, int ny, int nz) {
  orc_pf p;
  int n[3] = {nx, ny, nz};
  p.ndim = ndim;
  for (int a = 0; a < 3; a++) {
    if (a < ndim) {
      p.n[a] = n[a];
      p.N[a] = n[a] + 2 * ORC_GHC;
      p.off[a] = ORC_GHC;
    } else {
      p.n[a] = 1;
      p.N[a] = 1;
      p.off[a] = 0;
    }
  }
  p.s[2] = 1;
  p.s[1] = p.N[2];
  p.s[0] = (long)p.N[1] * p.N[2];
  return p;
}

int orc_padded_size(int ndim, int nx, int ny, int nz, int *N) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  N[0] = p.N[0];
  N[1] = p.N[1];
  N[2] = p.N[2];
  return p.N[0] * p.N[1] * p.N[2];
}

/* index of padded position (I,J,K) */
#define PIDX(p, I, J, K) ((I) * (p).s[0] + (J) * (p).s[1] + (K) * (p).s[2])

/* one boundary line along axis a; `b` points to padded position 0 of the
 * axis in this line; st = stride along the axis.  Formulas follow
 * neumann.cpp:7-164, dirichlet.cpp:7-118, periodic.cpp:7-113,
 * no_slip.cpp:7-103, slip.cpp:7-103 (same pattern on every axis).      */
static void bc_line(double *b, long st, int n, int N, int type, int right,
                    int face, double U, double h) {
  const int g = ORC_GHC;
#define F(q) b[(long)(q) * st]
  if (right) {
    const int I = n - 1 + g; /* index_mapping(n, ...) */
    switch (type) {
      case 0: /* neumann */
        if (face) F(I) = F(I - 1) + U * h;
        for (int i = 1; i <= g; ++i) F(I + i) = F(I + i - 1) + U * h;
        break;
      case 1: /* dirichlet */
        if (face)
          F(I) = U;
        else
          F(I) = F(I - 1) / 3.0 + 2.0 * U / 3.0;
        for (int i = 1; i <= g; ++i) F(I + i) = 2.0 * F(I + i - 1) - F(I + i - 2);
        break;
      case 2: /* periodic (node and face identical) */
        for (int i = 1; i <= g; ++i) F(I + i) = F(g + (i - 1));
        break;
      case 3: /* no slip */
        if (face) {
          for (int i = 1; i <= g; ++i) F(I + i) = -F(I - i);
          F(I) = 0.0;
        } else
          for (int i = 1; i <= g; ++i) F(I + i) = -F(I + 1 - i);
        break;
      case 4: /* slip */
        if (face) {
          for (int i = 1; i <= g; ++i) F(I + i) = F(I - i);
          F(I) = 0.0;
        } else
          for (int i = 1; i <= g; ++i) F(I + i) = F(I + 1 - i);
        break;
    }
  } else {
    switch (type) {
      case 0: { /* neumann */
        if (face) {
          const int I = g - 1; /* index_mapping(0, ...) */
          F(I) = F(I + 1) - U * h;
          for (int i = 1; i <= g - 1; ++i) F(I - i) = F(I - i + 1) - U * h;
        } else {
          const int I = g;
          for (int i = 1; i <= g; ++i) F(I - i) = F(I - i + 1) - U * h;
        }
        break;
      }
      case 1: { /* dirichlet */
        const int I = g;
        if (face) {
          F(I - 1) = U;
          for (int i = 2; i <= g; ++i) F(I - i) = 2.0 * F(I - i + 1) - F(I - i + 2);
        } else {
          F(I) = F(I + 1) / 3.0 + 2.0 * U / 3.0;
          for (int i = 1; i <= g; ++i) F(I - i) = 2.0 * F(I - i + 1) - F(I - i + 2);
        }
        break;
      }
      case 2: {
        const int I = g;
        for (int i = 1; i <= g; ++i) F(I - i) = F(N - g - 1 - (i - 1));
        break;
      }
      case 3: {
        const int I = g;
        if (face) {
          for (int i = 1; i <= g; ++i) F(I - i) = -F(I - 2 + i);
          F(I - 1) = 0.0;
        } else
          for (int i = 1; i <= g; ++i) F(I - i) = -F(I - 1 + i);
        break;
      }
      case 4: {
        const int I = g;
        if (face) {
          for (int i = 1; i <= g; ++i) F(I - i) = F(I - 2 + i);
          F(I - 1) = 0.0;
        } else
          for (int i = 1; i <= g; ++i) F(I - i) = F(I - 1 + i);
        break;
      }
    }
  }
#undef F
}

static void bc_side(const orc_pf *p, double *f, int a, int right, int type,
                    int face, double U, double h) {
  int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
  for (int u = 0; u < p->n[t1]; u++)
    for (int v = 0; v < p->n[t2]; v++) {
      long base = (long)(u + p->off[t1]) * p->s[t1] + (long)(v + p->off[t2]) * p->s[t2];
      bc_line(f + base, p->s[a], p->n[a], p->N[a], type, right, face, U, h);
    }
}

void orc_bc_apply(int ndim, int nx, int ny, int nz, double dx, double dy,
                  double dz, const int *types, const double *values, int stag,
                  double *f) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  double hh[3] = {dx, dy, dz};
  /* order: x-right, x-left, y-right, y-left, z-right, z-left
   * (bc_factory.cpp:61-138)                                         */
  for (int a = 0; a < ndim; a++) {
    bc_side(&p, f, a, 1, types[2 * a + 1], stag == a, values[2 * a + 1], hh[a]);
    bc_side(&p, f, a, 0, types[2 * a], stag == a, values[2 * a], hh[a]);
  }
}

/* ---------------- WENO-Z (weno_solver.cpp:16-66) ------------------ */
static void weno_ind_p(double a, double b, double c, double d, double e,
                       double *o) {
  o[0] = 13.0 * (c - 2.0 * d + e) * (c - 2.0 * d + e) +
         3.0 * (3.0 * c - 4.0 * d + e) * (3.0 * c - 4.0 * d + e);
  o[1] = 13.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 3.0 * (b - d) * (b - d);
  o[2] = 13.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) +
         3.0 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
}
static void weno_ind_m(double a, double b, double c, double d, double e,
                       double *o) {
  o[0] = 13.0 * (a - 2.0 * b + c) * (a - 2.0 * b + c) +
         3.0 * (a - 4.0 * b + 3.0 * c) * (a - 4.0 * b + 3.0 * c);
  o[1] = 13.0 * (b - 2.0 * c + d) * (b - 2.0 * c + d) + 3.0 * (b - d) * (b - d);
  o[2] = 13.0 * (c - 2.0 * d + e) * (c - 2.0 * d + e) +
         3.0 * (3.0 * c - 4.0 * d + e) * (3.0 * c - 4.0 * d + e);
}
static void weno_int_p(double a, double b, double c, double d, double e,
                       double *o) {
  o[2] = (-a + 5.0 * b + 2.0 * c) / 6.0;
  o[1] = (2.0 * b + 5.0 * c - d) / 6.0;
  o[0] = (11.0 * c - 7.0 * d + 2.0 * e) / 6.0;
}
static void weno_int_m(double a, double b, double c, double d, double e,
                       double *o) {
  o[2] = (-e + 5.0 * d + 2.0 * c) / 6.0;
  o[1] = (2.0 * d + 5.0 * c - b) / 6.0;
  o[0] = (11.0 * c - 7.0 * b + 2.0 * a) / 6.0;
}
static void weno_z_weights(double b0, double b1, double b2, double *w) {
  double a0 = 1.0 * (1.0 + fabs(b0 - b2) / (orc_eps + b0));
  double a1 = 6.0 * (1.0 + fabs(b0 - b2) / (orc_eps + b1));
  double a2 = 3.0 * (1.0 + fabs(b0 - b2) / (orc_eps + b2));
  w[0] = a0 / (a0 + a1 + a2);
  w[1] = a1 / (a0 + a1 + a2);
  w[2] = a2 / (a0 + a1 + a2);
}

/* derivative along axis a (weno5_flux_x + weno5_find_fx, :68-105, :185-210) */
static void weno_axis(const orc_pf *p, int a, double h, const double *flux,
                      const double *vel, double *out) {
  const int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
  const long st = p->s[a];
  const int g = ORC_GHC;
  double *fh = (double *)malloc(sizeof(double) * (p->n[a] + 2 * g + 2));
  for (int u = 0; u < p->n[t1]; u++)
    for (int v = 0; v < p->n[t2]; v++) {
      long base = (long)(u + p->off[t1]) * p->s[t1] + (long)(v + p->off[t2]) * p->s[t2];
      const double *F = flux + base;
      const double *V = vel + base;
      /* faces P = g-1 .. g+n-1 */
      for (int P = g - 1; P <= g + p->n[a] - 1; P++) {
        double in[3], ind[3], w[3], fm, fp;
        double A = F[(P - 2) * st], B = F[(P - 1) * st], C = F[P * st],
               D = F[(P + 1) * st], E = F[(P + 2) * st];
        weno_int_m(A, B, C, D, E, in);
        weno_ind_m(A, B, C, D, E, ind);
        weno_z_weights(ind[0], ind[1], ind[2], w);
        fm = w[0] * in[0] + w[1] * in[1] + w[2] * in[2];
        A = F[(P - 1) * st];
        B = F[P * st];
        C = F[(P + 1) * st];
        D = F[(P + 2) * st];
        E = F[(P + 3) * st];
        weno_int_p(A, B, C, D, E, in);
        weno_ind_p(A, B, C, D, E, ind);
        weno_z_weights(ind[0], ind[1], ind[2], w);
        fp = w[0] * in[0] + w[1] * in[1] + w[2] * in[2];
        fh[P] = (V[P * st] + V[(P + 1) * st] > 0.0) ? fm : fp;
      }
      for (int P = g; P <= g + p->n[a] - 1; P++)
        out[base + P * st] = (fh[P] - fh[P - 1]) / h;
    }
  free(fh);
}

void orc_weno5_derivatives(int ndim, int nx, int ny, int nz, double dx,
                           double dy, double dz, const double *flux,
                           const double *u, const double *v, const double *w,
                           double *fx, double *fy, double *fz) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  weno_axis(&p, 0, dx, flux, u, fx);
  if (ndim > 1) weno_axis(&p, 1, dy, flux, v, fy);
  if (ndim > 2) weno_axis(&p, 2, dz, flux, w, fz);
}

void orc_convection_weno(int ndim, int nx, int ny, int nz, double dx,
                         double dy, double dz, const double *phi,
                         const double *u, const double *v, const double *w,
                         double *fx, double *fy, double *fz, double *s) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  long tot = (long)p.N[0] * p.N[1] * p.N[2];
  /* identity_flux: flux <- data (flux.cpp:7-17); phi is used as flux */
  orc_weno5_derivatives(ndim, nx, ny, nz, dx, dy, dz, phi, u, v, w, fx, fy, fz);
  for (long i = 0; i < tot; i++) s[i] = -u[i] * fx[i];
  if (ndim > 1)
    for (long i = 0; i < tot; i++) s[i] -= v[i] * fy[i];
  if (ndim > 2)
    for (long i = 0; i < tot; i++) s[i] -= w[i] * fz[i];
}

void orc_tvd_rk3_weno(int ndim, int nx, int ny, int nz, double dx, double dy,
                      double dz, double dt, const int *types,
                      const double *values, double *phi, const double *u,
                      const double *v, const double *w, double *fx,
                      double *fy, double *fz, int nsteps) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  long tot = (long)p.N[0] * p.N[1] * p.N[2];
  double *s1 = (double *)calloc(tot, sizeof(double));
  double *s2 = (double *)calloc(tot, sizeof(double));
  double *s3 = (double *)calloc(tot, sizeof(double));
  for (int st = 0; st < nsteps; st++) {
    /* runge_kutta.cpp:13-49 */
    orc_convection_weno(ndim, nx, ny, nz, dx, dy, dz, phi, u, v, w, fx, fy, fz, s1);
    for (long i = 0; i < tot; i++) phi[i] += dt * s1[i];
    orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, types, values, -1, phi);
    orc_convection_weno(ndim, nx, ny, nz, dx, dy, dz, phi, u, v, w, fx, fy, fz, s2);
    for (long i = 0; i < tot; i++) phi[i] += dt * (s2[i] - 3.0 * s1[i]) / 4.0;
    orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, types, values, -1, phi);
    orc_convection_weno(ndim, nx, ny, nz, dx, dy, dz, phi, u, v, w, fx, fy, fz, s3);
    for (long i = 0; i < tot; i++)
      phi[i] += dt * (-s1[i] - s2[i] + 8.0 * s3[i]) / 12.0;
    orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, types, values, -1, phi);
  }
  free(s1);
  free(s2);
  free(s3);
}

/* ---------------- density (lsm.cpp:8-18, wrapper_func.cpp:387-402) -- */
static double orc_heaviside(double x, double w) {
  const double pi = acos(-1.0); /* global.h:13 */
  if (x < -w) return 0;
  if (x > w) return 1;
  return 0.5 * (1 + x / w + sin(pi * x / w) / pi);
}

void orc_density(long n, const double *phi, double ls_width, double ratio,
                 int positive_ref, double *rho) {
  for (long i = 0; i < n; i++) {
    double h = orc_heaviside(phi[i], ls_width);
    rho[i] = positive_ref ? h + (1.0 - h) * ratio : (1.0 - h) + h * ratio;
  }
}

/* ---------------- projection pieces -------------------------------- */
void orc_ppe_coefficients(int ndim, int nx, int ny, int nz, double dx,
                          double dy, double dz, const double *rho, double *CC,
                          double *CR, double *CL, double *CU, double *CD,
                          double *CF, double *CB) {
  /* projection_method.cpp:404-465 */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < (ndim > 2 ? nz : 1); ++K) {
        long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
        double R = 0.5 * (rho[c] + rho[c + p.s[0]]);
        double L = 0.5 * (rho[c] + rho[c - p.s[0]]);
        double Uu = 0.5 * (rho[c] + rho[c + p.s[1]]);
        double D = 0.5 * (rho[c] + rho[c - p.s[1]]);
        CR[c] = 1.0 / R / dx / dx;
        CL[c] = 1.0 / L / dx / dx;
        CU[c] = 1.0 / Uu / dy / dy;
        CD[c] = 1.0 / D / dy / dy;
        if (I == 0) CL[c] = 0.0;
        if (I == nx - 1) CR[c] = 0.0;
        if (J == 0) CD[c] = 0.0;
        if (J == ny - 1) CU[c] = 0.0;
        if (ndim > 2) {
          double Fz = 0.5 * (rho[c] + rho[c + p.s[2]]);
          double B = 0.5 * (rho[c] + rho[c - p.s[2]]);
          CF[c] = 1.0 / Fz / dz / dz;
          CB[c] = 1.0 / B / dz / dz;
          if (K == 0) CB[c] = 0.0;
          if (K == nz - 1) CF[c] = 0.0;
        }
        CC[c] = -(CR[c] + CL[c] + CU[c] + CD[c]);
        if (ndim > 2) CC[c] -= (CF[c] + CB[c]);
      }
}

void orc_projection_rhs(int ndim, int nx, int ny, int nz, double dx,
                        double dy, double dz, double dt, const double *u,
                        const double *v, const double *w, double *rhs) {
  /* projection_method.cpp:546-568 */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < (ndim > 2 ? nz : 1); ++K) {
        long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
        rhs[c] = (u[c] - u[c - p.s[0]]) / dx + (v[c] - v[c - p.s[1]]) / dy;
        if (ndim > 2) rhs[c] += (w[c] - w[c - p.s[2]]) / dz;
        rhs[c] /= dt;
      }
}

void orc_projection_velocity(int ndim, int nx, int ny, int nz, double dx,
                             double dy, double dz, double dt,
                             const double *rho, const double *pr, double *u,
                             double *v, double *w) {
  /* projection_method.cpp:630-654 (loops over the padded range) */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  for (int i = 0; i < p.N[0] - 1; ++i)
    for (int j = 0; j < p.N[1] - 1; ++j)
      for (int k = 0; k < p.N[2]; ++k) {
        long c = PIDX(p, i, j, k);
        double density = 0.5 * (rho[c] + rho[c + p.s[0]]);
        u[c] -= dt / dx / density * (pr[c + p.s[0]] - pr[c]);
        density = 0.5 * (rho[c] + rho[c + p.s[1]]);
        v[c] -= dt / dy / density * (pr[c + p.s[1]] - pr[c]);
        if (ndim > 2 && k < p.N[2] - 1) {
          density = 0.5 * (rho[c] + rho[c + p.s[2]]);
          w[c] -= dt / dz / density * (pr[c + p.s[2]] - pr[c]);
        }
      }
}

int orc_projection_sor(int ndim, int nx, int ny, int nz, double omega,
                       int max_iter, double tol, double tol2,
                       const double *CC, const double *CR, const double *CL,
                       const double *CU, const double *CD, const double *CF,
                       const double *CB, const double *rhs, double *pr,
                       double *tmp) {
  /* projection_method.cpp:572-626, one thread */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  long tot = (long)p.N[0] * p.N[1] * p.N[2];
  int types[6] = {0, 0, 0, 0, 0, 0};
  double vals[6] = {0, 0, 0, 0, 0, 0};
  int iter = 0;
  double error, sump, linf;
  do {
    memcpy(tmp, pr, sizeof(double) * tot);
    sump = 0.0;
    error = 0.0;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < (ndim > 2 ? nz : 1); ++k) {
          long c = PIDX(p, i + p.off[0], j + p.off[1], k + p.off[2]);
          double pnew = (rhs[c] - (CL[c] * pr[c - p.s[0]] + CR[c] * pr[c + p.s[0]] +
                                   CD[c] * pr[c - p.s[1]] + CU[c] * pr[c + p.s[1]]));
          if (ndim > 2)
            pnew -= (CB[c] * pr[c - p.s[2]] + CF[c] * pr[c + p.s[2]]);
          pnew /= CC[c];
          pr[c] = omega * pnew + (1.0 - omega) * pr[c];
          sump += pr[c];
          error += fabs(pnew * CC[c] - CC[c] * pr[c]);
        }
    sump = sump / (nx * ny);
    if (ndim > 2) sump = sump / nz;
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < (ndim > 2 ? nz : 1); ++k)
          pr[PIDX(p, i + p.off[0], j + p.off[1], k + p.off[2])] -= sump;
    orc_bc_apply(ndim, nx, ny, nz, 1.0, 1.0, 1.0, types, vals, -1, pr);
    ++iter;
    /* linfity (wrapper_func.cpp:420-432) */
    linf = 0.0;
    for (long i = 0; i < tot; i++) linf = fmax(linf, fabs(tmp[i] - pr[i]));
    linf = sqrt(linf / (double)tot);
  } while (iter < max_iter && error > tol && linf > tol2);
  return iter;
}

void orc_fpc_coefficients(int ndim, int nx, int ny, int nz, double dx,
                          double dy, double dz, double rho12,
                          const double *rho, const double *f, double *CC,
                          double *CR, double *CL, double *CU, double *CD,
                          double *CF, double *CB, double *RHS) {
  /* projection_method.cpp:467-541 */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < (ndim > 2 ? nz : 1); ++K) {
        long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
        double R = 0.5 * (rho[c] + rho[c + p.s[0]]);
        double L = 0.5 * (rho[c] + rho[c - p.s[0]]);
        double Uu = 0.5 * (rho[c] + rho[c + p.s[1]]);
        double D = 0.5 * (rho[c] + rho[c - p.s[1]]);
        CR[c] = (1.0 - rho12 / R) / dx / dx;
        CL[c] = (1.0 - rho12 / L) / dx / dx;
        CU[c] = (1.0 - rho12 / Uu) / dy / dy;
        CD[c] = (1.0 - rho12 / D) / dy / dy;
        if (I == 0) CL[c] = 0.0;
        if (I == nx - 1) CR[c] = 0.0;
        if (J == 0) CD[c] = 0.0;
        if (J == ny - 1) CU[c] = 0.0;
        if (ndim > 2) {
          double Fz = 0.5 * (rho[c] + rho[c + p.s[2]]);
          double B = 0.5 * (rho[c] + rho[c - p.s[2]]);
          CF[c] = (1.0 - rho12 / Fz) / dz / dz;
          CB[c] = (1.0 - rho12 / B) / dz / dz;
          if (K == 0) CB[c] = 0.0;
          if (K == nz - 1) CF[c] = 0.0;
        }
        CC[c] = -(CR[c] + CL[c] + CU[c] + CD[c]);
        if (ndim > 2) CC[c] -= (CF[c] + CB[c]);
        RHS[c] = CC[c] * f[c] + CR[c] * f[c + p.s[0]] + CL[c] * f[c - p.s[0]] +
                 CU[c] * f[c + p.s[1]] + CD[c] * f[c - p.s[1]];
        if (ndim > 2) RHS[c] += CF[c] * f[c + p.s[2]] + CB[c] * f[c - p.s[2]];
      }
}

void orc_fpc_rhs(int ndim, int nx, int ny, int nz, double dx, double dy,
                 double dz, double dt, double rho12, const double *u,
                 const double *v, const double *w, const double *RHS,
                 double *mg_rhs) {
  /* projection_method.cpp:667-692 */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  const int nzz = (ndim > 2 ? nz : 1);
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < nzz; ++K) {
        long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
        long pos = I + (long)nx * (J + (long)ny * K);
        mg_rhs[pos] = (u[c] - u[c - p.s[0]]) / dx + (v[c] - v[c - p.s[1]]) / dy;
        if (ndim > 2) mg_rhs[pos] += (w[c] - w[c - p.s[2]]) / dz;
        mg_rhs[pos] *= rho12 / dt;
        mg_rhs[pos] += RHS[c];
      }
}

void orc_mg_to_padded(int ndim, int nx, int ny, int nz, const double *sol,
                      double *pr) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  const int nzz = (ndim > 2 ? nz : 1);
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < nzz; ++K)
        pr[PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2])] =
            sol[I + (long)nx * (J + (long)ny * K)];
}

void orc_padded_to_mg(int ndim, int nx, int ny, int nz, const double *pr,
                      double *sol) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  const int nzz = (ndim > 2 ? nz : 1);
  for (int I = 0; I < nx; ++I)
    for (int J = 0; J < ny; ++J)
      for (int K = 0; K < nzz; ++K)
        sol[I + (long)nx * (J + (long)ny * K)] =
            pr[PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2])];
}

void orc_fpc_velocity(int ndim, int nx, int ny, int nz, double dx, double dy,
                      double dz, double dt, double rho12, const double *rho,
                      const double *pr, const double *ph, double *u,
                      double *v, double *w) {
  /* projection_method.cpp:723-750 */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  for (int i = 0; i < p.N[0] - 1; ++i)
    for (int j = 0; j < p.N[1] - 1; ++j)
      for (int k = 0; k < p.N[2]; ++k) {
        long c = PIDX(p, i, j, k);
        double px = (pr[c + p.s[0]] - pr[c]) / dx;
        double px2 = (ph[c + p.s[0]] - ph[c]) / dx;
        double density = 0.5 * (rho[c] + rho[c + p.s[0]]);
        u[c] -= dt * ((px - px2) / rho12 + px2 / density);
        double py = (pr[c + p.s[1]] - pr[c]) / dy;
        double py2 = (ph[c + p.s[1]] - ph[c]) / dy;
        density = 0.5 * (rho[c] + rho[c + p.s[1]]);
        v[c] -= dt * ((py - py2) / rho12 + py2 / density);
        if (ndim > 2 && k < p.N[2] - 1) {
          double pz = (pr[c + p.s[2]] - pr[c]) / dz;
          double pz2 = (ph[c + p.s[2]] - ph[c]) / dz;
          density = 0.5 * (rho[c] + rho[c + p.s[2]]);
          w[c] -= dt * ((pz - pz2) / rho12 + pz2 / density);
        }
      }
}

void orc_pressure_guess(long n, double dt_old, double dt_old2,
                        const double *old, const double *old2, double *phat) {
  /* projection_method.cpp:803-815 */
  for (long i = 0; i < n; i++) {
    double dpdt = (old[i] - old2[i]) / dt_old2;
    phat[i] = old[i] + dt_old * dpdt;
  }
}

int orc_fpc_step(int ndim, int nx, int ny, int nz, double dx, double dy,
                 double dz, double dt, double dt_old, double dt_old2,
                 double rho12, double ppe_tol, int max_cycles,
                 const int *vel_types, const double *rho, const double *p_old,
                 const double *p_old2, double *u, double *v, double *w,
                 double *pr, double *phat) {
  /* projection_method.cpp:803-815 (guess), :467-541 (coefficients),
   * :659-753 (rhs, full-cycle loop, copy, pressure BCs, velocity, velocity
   * BCs = wrapper::apply_vel_bc, wrapper.cpp:21-26) */
  orc_pf p = pf_make(ndim, nx, ny, nz);
  const long tot = (long)p.N[0] * p.N[1] * p.N[2];
  orc_pressure_guess(tot, dt_old, dt_old2, p_old, p_old2, phat);
  double *C = (double *)calloc((size_t)8 * tot, sizeof(double));
  orc_fpc_coefficients(ndim, nx, ny, nz, dx, dy, dz, rho12, rho, phat, C, C + tot, C + 2 * tot, C + 3 * tot,
                       C + 4 * tot, C + 5 * tot, C + 6 * tot, C + 7 * tot);
  void *m = orc_mg_create(ndim, nx, ny, ndim > 2 ? nz : 1, dx, dy, dz, 0);
  orc_fpc_rhs(ndim, nx, ny, nz, dx, dy, dz, dt, rho12, u, v, w, C + 7 * tot, orc_mg_vec(m, 0, 0));
  double err = orc_mg_residual(m, 0);
  int iter = 0;
  while (err > ppe_tol && iter < max_cycles) {
    ++iter;
    orc_mg_fullcycle_ref(m);
    err = orc_mg_residual(m, 0);
  }
  orc_mg_to_padded(ndim, nx, ny, nz, orc_mg_vec(m, 0, 1), pr);
  const int nt[6] = {0, 0, 0, 0, 0, 0};
  const double zero[6] = {0, 0, 0, 0, 0, 0};
  orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, nt, zero, -1, pr);
  orc_fpc_velocity(ndim, nx, ny, nz, dx, dy, dz, dt, rho12, rho, pr, phat, u, v, w);
  /* the velocity wrapper turns SLIP into NO_SLIP for the wall-normal
   * component (wrapper.cpp:111-125 and the 3-D constructor) */
  double *comp[3] = {u, v, w};
  for (int a = 0; a < 3; ++a) {
    int t[6];
    for (int q = 0; q < 6; ++q) t[q] = vel_types[q];
    for (int q = 2 * a; q < 2 * a + 2; ++q)
      if (t[q] == 4) t[q] = 3;
    orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, t, zero, a, comp[a]);
  }
  orc_mg_destroy(m);
  free(C);
  return iter;
}

/* ================================================================== */
/* legacy VTK snapshot: vtkWriter::create / add_scalar / add_vector /   */
/* close (vtkWriter.cpp:18-111), header doubles as ostream's default   */
/* "%g", values big-endian (SwapEnd, vtkWriter.cpp:7-11), k-j-i order.  */
#include <stdio.h>

static void vtk_put_be(FILE *f, double v) {
  unsigned char b[8], r[8];
  memcpy(b, &v, 8);
  for (int q = 0; q < 8; q++) r[q] = b[7 - q];
  fwrite(r, 1, 8, f);
}

int orc_vtk_write(const char *path, int ndim, int nx, int ny, int nz,
                  const double *origin, const double *spacing, int nscalar,
                  const char *const *snames, const double *const *scalars,
                  int nvector, const char *const *vnames,
                  const double *const *vectors) {
  orc_pf p = pf_make(ndim, nx, ny, nz);
  FILE *f = fopen(path, "wb");
  if (!f) return -1;
  int dim[3];
  double org[3], sp[3];
  for (int a = 0; a < 3; a++) {
    dim[a] = a < ndim ? p.n[a] + 1 : 1;       /* structured_grid.cpp:30-36 */
    org[a] = a < ndim ? origin[a] : 0.0;
    sp[a] = a < ndim ? spacing[a] : 0.0;
  }
  fprintf(f, "# vtk DataFile Version 3.0\nvtk output\nBINARY\n");
  fprintf(f, "DATASET STRUCTURED_POINTS\n");
  fprintf(f, "DIMENSIONS %d %d %d\n", dim[0], dim[1], dim[2]);
  fprintf(f, "ORIGIN %g %g %g\n", org[0], org[1], org[2]);
  fprintf(f, "SPACING %g %g %g\n", sp[0], sp[1], sp[2]);
  long size = 1;
  for (int a = 0; a < 3; a++)
    if (dim[a] > 1) size *= dim[a] - 1;
  fprintf(f, "CELL_DATA %ld\n", size);
  for (int q = 0; q < nscalar; q++) {
    fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", snames[q]);
    for (int k = 0; k < p.n[2]; k++)
      for (int j = 0; j < p.n[1]; j++)
        for (int i = 0; i < p.n[0]; i++)
          vtk_put_be(f, scalars[q][PIDX(p, i + p.off[0], j + p.off[1], k + p.off[2])]);
  }
  for (int q = 0; q < nvector; q++) {
    fprintf(f, "VECTORS %s double\n", vnames[q]);
    for (int k = 0; k < p.n[2]; k++)
      for (int j = 0; j < p.n[1]; j++)
        for (int i = 0; i < p.n[0]; i++) {
          const long id = PIDX(p, i + p.off[0], j + p.off[1], k + p.off[2]);
          for (int c = 0; c < 3; c++) vtk_put_be(f, vectors[3 * q + c][id]);
        }
  }
  return fclose(f) == 0 ? 0 : -1;
}

/* ================================================================== */
/* momentum predictor: projection_method::find_source_sec (:314-382)   */
/* with add_stress_x/y/z (:31-181), add_body_force (:183-203),         */
/* add_surface_force (:205-240), then find_intermediate_velocity       */
/* (:384-402).  Level-set helpers: Heaviside/Delta (lsm.cpp:8-30),     */
/* find_delta / find_density / find_viscosity / find_curvature         */
/* (wrapper_func.cpp:20-29, 115-172, 387-416), identity_with_          */
/* extrapolation[_face_x/y/z] (flux.cpp:98-375), second_order_solver   */
/* (second_order_solver.cpp:13-242), convection_sec (source.cpp:42-73),*/
/* all_to_face_x/y/z and node_from_face (wrapper_func.cpp:177-293).    */
/* Every array is a padded field; derivative arrays that the reference */
/* only writes on part of the padded range are zero elsewhere (they are*/
/* zero-initialised, init_array, and nothing else writes them).        */

static const double orc_pi = 3.14159265358979323846; /* global.h:13 */

static double orc_delta(double x, double w) {
  if (x < -w) return 0;
  if (x > w) return 0;
  return 0.5 * (1 + cos(orc_pi * x / w)) / w;
}

static double orc_heav(double x, double w) {
  if (x < -w) return 0;
  if (x > w) return 1;
  return 0.5 * (1 + x / w + sin(orc_pi * x / w) / orc_pi);
}

/* identity_with_extrapolation[_face_a]: flux = data, then ghosts copied
 * axis by axis (x, y, z) over every index of the other axes; on the face
 * axis `fa` the low side copies the wall face (index off-1) into the
 * ghosts below it and keeps the face itself (flux.cpp:166-375). fa = -1:
 * plain (flux.cpp:98-149). */
static void orc_extrap(const orc_pf *p, const double *data, double *flux, int fa) {
  const long tot = (long)p->N[0] * p->N[1] * p->N[2];
  memcpy(flux, data, sizeof(double) * tot);
  for (int a = 0; a < p->ndim; a++) {
    const int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
    const int lo = p->off[a], hi = p->off[a] + p->n[a] - 1;
    for (int u = 0; u < p->N[t1]; u++)
      for (int v = 0; v < p->N[t2]; v++) {
        const long base = (long)u * p->s[t1] + (long)v * p->s[t2];
        for (int q = 1; q <= ORC_GHC; q++) flux[base + (hi + q) * p->s[a]] = flux[base + hi * p->s[a]];
        if (a == fa) {
          for (int q = 2; q <= ORC_GHC; q++) flux[base + (lo - q) * p->s[a]] = flux[base + (lo - 1) * p->s[a]];
        } else {
          for (int q = 1; q <= ORC_GHC; q++) flux[base + (lo - q) * p->s[a]] = flux[base + lo * p->s[a]];
        }
      }
  }
}

/* second_order_solver::find_fx/fy/fz(f) on the whole padded range, one-
 * sided at the array ends (second_order_solver.cpp:13-88) */
static void orc_sec_full(const orc_pf *p, int a, double h, const double *f, double *d1, double *d2) {
  const long st = p->s[a];
  for (int I = 0; I < p->N[0]; I++)
    for (int J = 0; J < p->N[1]; J++)
      for (int K = 0; K < p->N[2]; K++) {
        const int idx[3] = {I, J, K};
        const long c = PIDX(*p, I, J, K);
        const int q = idx[a];
        if (q == 0) {
          d1[c] = (-1.5 * f[c] + 2 * f[c + st] - 0.5 * f[c + 2 * st]) / h;
          d2[c] = (2 * f[c] - 5 * f[c + st] + 4 * f[c + 2 * st] - f[c + 3 * st]) / (h * h);
        } else if (q == p->N[a] - 1) {
          d1[c] = (1.5 * f[c] - 2 * f[c - st] + 0.5 * f[c - 2 * st]) / h;
          d2[c] = (2 * f[c] - 5 * f[c - st] + 4 * f[c - 2 * st] - f[c - 3 * st]) / (h * h);
        } else {
          d1[c] = (f[c + st] - f[c - st]) / (2 * h);
          d2[c] = (f[c + st] - 2 * f[c] + f[c - st]) / (h * h);
        }
      }
}

/* second_order_solver::mixed_xy / mixed_yz / mixed_zx: d/d(axis a) of g,
 * one-sided at the array ends (second_order_solver.cpp:173-230) */
static void orc_sec_mixed(const orc_pf *p, int a, double h, const double *g, double *out) {
  const long st = p->s[a];
  for (int I = 0; I < p->N[0]; I++)
    for (int J = 0; J < p->N[1]; J++)
      for (int K = 0; K < p->N[2]; K++) {
        const int idx[3] = {I, J, K};
        const long c = PIDX(*p, I, J, K);
        const int q = idx[a];
        if (q == 0) out[c] = (-1.5 * g[c] + 2 * g[c + st] - 0.5 * g[c + 2 * st]) / h;
        else if (q == p->N[a] - 1) out[c] = (1.5 * g[c] - 2 * g[c - st] + 0.5 * g[c - 2 * st]) / h;
        else out[c] = (g[c + st] - g[c - st]) / (2 * h);
      }
}

/* second_order_solver::find_fx/fy/fz(f, vel): upwind by vel_a > 0 on the
 * interior of axis a, every index of the other axes
 * (second_order_solver.cpp:90-148) */
static void orc_sec_upwind(const orc_pf *p, int a, double h, const double *f, const double *va,
                           double *d1, double *d2) {
  const long st = p->s[a];
  for (int I = 0; I < p->N[0]; I++)
    for (int J = 0; J < p->N[1]; J++)
      for (int K = 0; K < p->N[2]; K++) {
        const int idx[3] = {I, J, K};
        const int q = idx[a];
        if (q < p->off[a] || q >= p->off[a] + p->n[a]) continue;
        const long c = PIDX(*p, I, J, K);
        if (va[c] > 0.0) d1[c] = (1.5 * f[c] - 2 * f[c - st] + 0.5 * f[c - 2 * st]) / h;
        else d1[c] = (-1.5 * f[c] + 2 * f[c + st] - 0.5 * f[c + 2 * st]) / h;
        d2[c] = (f[c + st] - 2 * f[c] + f[c - st]) / (h * h);
      }
}

typedef struct {
  double dt, Re, Fr, We, ls_width, density_ratio, viscosity_ratio;
  int positive_ref;
} orc_flow;

static void orc_vel_bc(int ndim, const int *n, const double *h, const int *types, int comp, int stag, double *f) {
  /* the component's BC factory: wall-normal SLIP made NO_SLIP (wrapper.cpp:111-125) */
  int t[6];
  for (int q = 0; q < 6; q++) t[q] = (q / 2 == comp && types[q] == 4) ? 3 : types[q];
  double z[6] = {0, 0, 0, 0, 0, 0};
  orc_bc_apply(ndim, n[0], n[1], n[2], h[0], h[1], h[2], t, z, stag, f);
}

int orc_momentum_step(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                      const double *prm8, const int *vel_types, const double *phi, double *vel0,
                      double *vel1, double *vel2, double *nv0, double *nv1, double *nv2,
                      double *src0, double *src1, double *src2, double *rho_out) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const int n[3] = {nx, ny, nz};
  const double hh[3] = {dx, dy, dz};
  orc_flow F = {prm8[0], prm8[1], prm8[2], prm8[3], prm8[4], prm8[5], prm8[6], (int)prm8[7]};
  const long T = (long)p.N[0] * p.N[1] * p.N[2];
  double *vel[3] = {vel0, vel1, vel2}, *nv[3] = {nv0, nv1, nv2}, *src[3] = {src0, src1, src2};
  /* scratch: 3 old sources, delta, density, viscosity, curvature, level-set
   * flux and 9 derivatives, flux + 6 derivatives of one velocity component,
   * flux + 2 derivatives of one nvel component */
  const int NS = 3 + 4 + 10 + 7 + 3;
  double *mem = calloc((size_t)NS * T, sizeof(double));
  if (!mem) return -1;
  double *old[3] = {mem, mem + T, mem + 2 * T};
  double *dl = mem + 3 * T, *rho = mem + 4 * T, *mu = mem + 5 * T, *curv = mem + 6 * T;
  double *lfl = mem + 7 * T, *lf[3] = {mem + 8 * T, mem + 9 * T, mem + 10 * T};
  double *lff[3] = {mem + 11 * T, mem + 12 * T, mem + 13 * T};
  double *lfxy = mem + 14 * T, *lfyz = mem + 15 * T, *lfzx = mem + 16 * T;
  double *vfl = mem + 17 * T, *vd1[3] = {mem + 18 * T, mem + 19 * T, mem + 20 * T};
  double *vd2[3] = {mem + 21 * T, mem + 22 * T, mem + 23 * T};
  double *nfl = mem + 24 * T, *nd1 = mem + 25 * T, *nd2 = mem + 26 * T;
  const double ref = F.positive_ref ? 1.0 : -1.0;

  for (int a = 0; a < 3; a++) memcpy(old[a], src[a], sizeof(double) * T);
  /* find_delta, find_density, find_viscosity (heaviside itself is unused) */
  for (long c = 0; c < T; c++) {
    dl[c] = orc_delta(phi[c], F.ls_width);
    const double H = orc_heav(phi[c], F.ls_width);
    rho[c] = F.positive_ref ? H + (1.0 - H) * F.density_ratio : (1.0 - H) + H * F.density_ratio;
    mu[c] = F.positive_ref ? H + (1.0 - H) * F.viscosity_ratio : (1.0 - H) + H * F.viscosity_ratio;
  }
  /* identity_with_extrapolation + secSol->find_derivatives_all (lsf) */
  orc_extrap(&p, phi, lfl, -1);
  for (int a = 0; a < ndim; a++) orc_sec_full(&p, a, hh[a], lfl, lf[a], lff[a]);
  if (ndim > 1) orc_sec_mixed(&p, 1, dy, lf[0], lfxy);
  if (ndim > 2) {
    orc_sec_mixed(&p, 2, dz, lf[1], lfyz);
    orc_sec_mixed(&p, 0, dx, lf[2], lfzx);
  }

  for (int a = 0; a < ndim; a++) {
    const long sa = p.s[a];
    /* all_to_face_a (wrapper_func.cpp:177-262): nvel_a = vel_a, the other
     * components averaged onto face a from their own faces */
    const int lo2 = ndim == 2 ? 0 : 1, hi2 = ndim == 2 ? p.N[2] : p.N[2] - 1;
    for (int I = 1; I < p.N[0] - 1; I++)
      for (int J = 1; J < p.N[1] - 1; J++)
        for (int K = lo2; K < hi2; K++) {
          const long c = PIDX(p, I, J, K);
          nv[a][c] = vel[a][c];
          for (int b = 0; b < ndim; b++) {
            if (b == a) continue;
            const long sb = p.s[b];
            if (b < a)  /* components below a: shifted +a, -b (x onto y: u[i][j], u[i][j+1], u[i-1][j], u[i-1][j+1]) */
              nv[b][c] = 0.25 * (vel[b][c] + vel[b][c + sa] + vel[b][c - sb] + vel[b][c - sb + sa]);
            else        /* components above a: vel_b[c], [c+a], [c-b], [c+a-b] */
              nv[b][c] = 0.25 * (vel[b][c] + vel[b][c + sa] + vel[b][c - sb] + vel[b][c + sa - sb]);
          }
        }
    /* nvel->apply_vel_a_bc (wrapper.cpp:40-59) */
    for (int b = 0; b < 3; b++) orc_vel_bc(ndim, n, hh, vel_types, b, a, nv[b]);
    /* convection_sec(vel_a, nvel, src_a, identity_with_extrapolation_face_a) */
    orc_extrap(&p, vel[a], vfl, a);
    for (int q = 0; q < 3; q++) {
      memset(vd1[q], 0, sizeof(double) * T);
      memset(vd2[q], 0, sizeof(double) * T);
    }
    for (int q = 0; q < ndim; q++) orc_sec_upwind(&p, q, hh[q], vfl, nv[q], vd1[q], vd2[q]);
    for (long c = 0; c < T; c++) src[a][c] = -nv[0][c] * vd1[0][c];
    if (ndim > 1)
      for (long c = 0; c < T; c++) src[a][c] -= nv[1][c] * vd1[1][c];
    if (ndim > 2)
      for (long c = 0; c < T; c++) src[a][c] -= nv[2][c] * vd1[2][c];
    /* add_stress_a: derivatives along a of the other nvel components, then
     * the interior stress term */
    double *cross[3] = {NULL, NULL, NULL};
    double *cross_mem = calloc((size_t)3 * T, sizeof(double));
    if (!cross_mem) {
      free(mem);
      return -1;
    }
    for (int b = 0; b < ndim; b++) {
      if (b == a) continue;
      cross[b] = cross_mem + b * T;
      orc_extrap(&p, nv[b], nfl, a);
      memset(nd2, 0, sizeof(double) * T);
      orc_sec_upwind(&p, a, hh[a], nfl, nv[a], cross[b], nd2);
    }
    (void)nd1;
    for (int I = 0; I < n[0]; I++)
      for (int J = 0; J < (ndim > 1 ? n[1] : 1); J++)
        for (int K = 0; K < (ndim > 2 ? n[2] : 1); K++) {
          const long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
          const double delta = 0.5 * (dl[c] + dl[c + sa]);
          const double m = 0.5 * (mu[c] + mu[c + sa]);
          const double r = 0.5 * (rho[c] + rho[c + sa]);
          double ph[3] = {0, 0, 0};
          for (int q = 0; q < ndim; q++) ph[q] = 0.5 * (lf[q][c] + lf[q][c + sa]);
          double p1, p2;
          if (a == 0) {
            p1 = vd2[0][c] + vd2[1][c];
            p2 = 2.0 * ph[0] * vd1[0][c] + ph[1] * (vd1[1][c] + cross[1][c]);
            if (ndim > 2) {
              p1 += vd2[2][c];
              p2 += ph[2] * (vd1[2][c] + cross[2][c]);
            }
          } else if (a == 1) {
            p1 = vd2[0][c] + vd2[1][c];
            p2 = 2.0 * ph[1] * vd1[1][c] + ph[0] * (vd1[0][c] + cross[0][c]);
            if (ndim > 2) {
              p1 += vd2[2][c];
              p2 += ph[2] * (vd1[2][c] + cross[2][c]);
            }
          } else { /* add_stress_z has no w_zz term (projection_method.cpp:171) */
            p1 = vd2[0][c] + vd2[1][c];
            p2 = 2.0 * ph[2] * vd1[2][c] + ph[0] * (vd1[0][c] + cross[0][c]) + ph[1] * (vd1[1][c] + cross[1][c]);
          }
          p1 *= m;
          p2 *= ref * delta * (1.0 - F.viscosity_ratio);
          src[a][c] += (p1 + p2) / F.Re / r;
        }
    free(cross_mem);
  }
  /* add_body_force: gravity along y (2-D) or z (3-D) on every padded cell */
  if (F.Fr > 0.0) {
    double *g = src[ndim == 2 ? 1 : 2];
    if (ndim >= 2)
      for (long c = 0; c < T; c++) g[c] += -1.0 / F.Fr;
  }
  /* add_surface_force */
  if (F.We > 0.0) {
    for (long c = 0; c < T; c++) {
      double k;
      if (ndim == 2) {
        const double fx = lf[0][c], fy = lf[1][c], fxx = lff[0][c], fyy = lff[1][c], fxy = lfxy[c];
        k = (fxx * fy * fy + fyy * fx * fx - 2.0 * fxy * fx * fy) / pow(fx * fx + fy * fy, 1.5);
        if (k > 0.0) k = 1.0 / (1.0 / k + phi[c]);
        else k = -1.0 / (-1.0 / k + phi[c]);
      } else {
        const double fx = lf[0][c], fy = lf[1][c], fz = lf[2][c];
        const double fxx = lff[0][c], fyy = lff[1][c], fzz = lff[2][c];
        const double fxy = lfxy[c], fxz = lfzx[c], fyz = lfyz[c];
        k = (fx * fx * (fyy + fzz) + fy * fy * (fxx + fzz) + fz * fz * (fxx + fyy) -
             2.0 * (fx * fy * fxy + fx * fz * fxz + fy * fz * fyz)) /
            pow(fx * fx + fy * fy + fz * fz, 1.5);
        if (k > 0.0) k = 1.0 / (2.0 / k + phi[c]);
        else k = -1.0 / (-2.0 / k + phi[c]);
      }
      curv[c] = k;
    }
    for (int I = 0; I < p.N[0] - 1; I++)
      for (int J = 0; J < p.N[1] - 1; J++)
        for (int K = 0; K < p.N[2]; K++) {
          const long c = PIDX(p, I, J, K);
          for (int a = 0; a < ndim; a++) {
            const long sa = p.s[a];
            if (a == 2 && K >= p.N[2] - 1) continue;
            const double cv = 0.5 * (curv[c] + curv[c + sa]);
            const double delta = 0.5 * (dl[c] + dl[c + sa]);
            const double pa = 0.5 * (lf[a][c] + lf[a][c + sa]);
            const double r = 0.5 * (rho[c] + rho[c + sa]);
            src[a][c] -= cv * delta * pa / (F.We * r);
          }
        }
  }
  /* find_intermediate_velocity: AB2 on every padded cell, then the BCs */
  for (int a = 0; a < 3; a++)
    for (long c = 0; c < T; c++) vel[a][c] += F.dt * (1.5 * src[a][c] - 0.5 * old[a][c]);
  for (int a = 0; a < 3; a++) orc_vel_bc(ndim, n, hh, vel_types, a, a, vel[a]);
  if (rho_out) memcpy(rho_out, rho, sizeof(double) * T);
  free(mem);
  return 0;
}

/* node_from_face + nvel->apply_nvel_bc (wrapper_func.cpp:264-293,
 * wrapper.cpp:28-33): cell-centred velocities from the face values */
void orc_node_velocity(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                       const int *vel_types, const double *u, const double *v, const double *w,
                       double *nu, double *nv, double *nw) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const int n[3] = {nx, ny, nz};
  const double hh[3] = {dx, dy, dz};
  const int klo = ndim == 2 ? 0 : 1;
  for (int I = 1; I < p.N[0] - 1; I++)
    for (int J = 1; J < p.N[1] - 1; J++)
      for (int K = klo; K < p.N[2]; K++) {
        const long c = PIDX(p, I, J, K);
        nu[c] = 0.5 * (u[c] + u[c - p.s[0]]);
        nv[c] = 0.5 * (v[c] + v[c - p.s[1]]);
        if (ndim == 3) nw[c] = 0.5 * (w[c] + w[c - p.s[2]]);
      }
  double *f[3] = {nu, nv, nw};
  for (int a = 0; a < 3; a++) orc_vel_bc(ndim, n, hh, vel_types, a, -1, f[a]);
}

/* ================================================================== */
/* compact schemes: twin (2x2 block) tridiagonal decomposition and back */
/* substitution (matrix_solver.cpp:21-83), coefficients of the CCD     */
/* (ccd_base.cpp:43-104) and UCCD (uccd_base.cpp:16-117) line systems, */
/* and the derivatives along padded lines (ccd_solver.cpp:14-142,      */
/* uccd_solver.cpp:14-206).  A coefficient set is 12 arrays of n:      */
/* a0 a1 a2 b0 b1 b2 aa0 aa1 aa2 bb0 bb1 bb2.                          */

static void orc_twin_dec(double *c, int n) {
  double *a[3] = {c, c + n, c + 2 * n}, *b[3] = {c + 3 * n, c + 4 * n, c + 5 * n};
  double *aa[3] = {c + 6 * n, c + 7 * n, c + 8 * n}, *bb[3] = {c + 9 * n, c + 10 * n, c + 11 * n};
  a[0][0] = 0.0;
  aa[0][0] = 0.0;
  b[0][0] = 1.0;
  bb[0][0] = 1.0;
  a[2][n - 1] = 0.0;
  aa[2][n - 1] = 0.0;
  b[2][n - 1] = 1.0;
  bb[2][n - 1] = 1.0;
  for (int i = 1; i < n; i++) {
    const double den = a[1][i - 1] * bb[1][i - 1] - aa[1][i - 1] * b[1][i - 1];
    double sc1 = -(aa[1][i - 1] * b[0][i] - a[0][i] * bb[1][i - 1]) / den;
    double sc2 = (b[0][i] * a[1][i - 1] - b[1][i - 1] * a[0][i]) / den;
    a[0][i] = sc1;
    a[1][i] = a[1][i] - (sc1 * a[2][i - 1] + sc2 * aa[2][i - 1]);
    b[0][i] = sc2;
    b[1][i] = b[1][i] - (sc1 * b[2][i - 1] + sc2 * bb[2][i - 1]);
    sc1 = -(aa[1][i - 1] * bb[0][i] - aa[0][i] * bb[1][i - 1]) / den;
    sc2 = (bb[0][i] * a[1][i - 1] - b[1][i - 1] * aa[0][i]) / den;
    aa[0][i] = sc1;
    aa[1][i] = aa[1][i] - (sc1 * a[2][i - 1] + sc2 * aa[2][i - 1]);
    bb[0][i] = sc2;
    bb[1][i] = bb[1][i] - (sc1 * b[2][i - 1] + sc2 * bb[2][i - 1]);
  }
}

/* s, ss: right-hand sides along the line (stride st), overwritten */
static void orc_twin_bks(const double *c, int n, double *s, double *ss, long st) {
  const double *a[3] = {c, c + n, c + 2 * n}, *b[3] = {c + 3 * n, c + 4 * n, c + 5 * n};
  const double *aa[3] = {c + 6 * n, c + 7 * n, c + 8 * n}, *bb[3] = {c + 9 * n, c + 10 * n, c + 11 * n};
#define S(i) s[(long)(i) * st]
#define SS(i) ss[(long)(i) * st]
  for (int i = 1; i < n; i++) {
    S(i) = S(i) - (a[0][i] * S(i - 1) + b[0][i] * SS(i - 1));
    SS(i) = SS(i) - (aa[0][i] * S(i - 1) + bb[0][i] * SS(i - 1));
  }
  double den = a[1][n - 1] * bb[1][n - 1] - aa[1][n - 1] * b[1][n - 1];
  double sols = -(b[1][n - 1] * SS(n - 1) - bb[1][n - 1] * S(n - 1)) / den;
  double solss = (a[1][n - 1] * SS(n - 1) - aa[1][n - 1] * S(n - 1)) / den;
  S(n - 1) = sols;
  SS(n - 1) = solss;
  for (int i = n - 2; i > -1; --i) {
    S(i) = S(i) - (a[2][i] * S(i + 1) + b[2][i] * SS(i + 1));
    SS(i) = SS(i) - (aa[2][i] * S(i + 1) + bb[2][i] * SS(i + 1));
    den = a[1][i] * bb[1][i] - aa[1][i] * b[1][i];
    sols = -(b[1][i] * SS(i) - bb[1][i] * S(i)) / den;
    solss = (a[1][i] * SS(i) - aa[1][i] * S(i)) / den;
    S(i) = sols;
    SS(i) = solss;
  }
#undef S
#undef SS
}

/* orc_twin_bks carried out in long double (x86-64: 64-bit significand) on
 * the same double right-hand sides and factorised coefficients, rounded to
 * double at the end: the sweep's result with its own rounding error made
 * about 2^11 times smaller, the yardstick for line solvers that round
 * differently from the reference's sequential order. */
static void orc_twin_bks_ld(const double *c, int n, double *s, double *ss) {
  const double *a[3] = {c, c + n, c + 2 * n}, *b[3] = {c + 3 * n, c + 4 * n, c + 5 * n};
  const double *aa[3] = {c + 6 * n, c + 7 * n, c + 8 * n}, *bb[3] = {c + 9 * n, c + 10 * n, c + 11 * n};
  long double *S = malloc(sizeof(long double) * 2 * n), *SS = S + n;
  for (int i = 0; i < n; i++) {
    S[i] = s[i];
    SS[i] = ss[i];
  }
  for (int i = 1; i < n; i++) {
    S[i] = S[i] - ((long double)a[0][i] * S[i - 1] + (long double)b[0][i] * SS[i - 1]);
    SS[i] = SS[i] - ((long double)aa[0][i] * S[i - 1] + (long double)bb[0][i] * SS[i - 1]);
  }
  long double den = (long double)a[1][n - 1] * bb[1][n - 1] - (long double)aa[1][n - 1] * b[1][n - 1];
  long double sols = -((long double)b[1][n - 1] * SS[n - 1] - (long double)bb[1][n - 1] * S[n - 1]) / den;
  long double solss = ((long double)a[1][n - 1] * SS[n - 1] - (long double)aa[1][n - 1] * S[n - 1]) / den;
  S[n - 1] = sols;
  SS[n - 1] = solss;
  for (int i = n - 2; i > -1; --i) {
    const long double si = S[i] - ((long double)a[2][i] * S[i + 1] + (long double)b[2][i] * SS[i + 1]);
    const long double ssi = SS[i] - ((long double)aa[2][i] * S[i + 1] + (long double)bb[2][i] * SS[i + 1]);
    den = (long double)a[1][i] * bb[1][i] - (long double)aa[1][i] * b[1][i];
    S[i] = -((long double)b[1][i] * ssi - (long double)bb[1][i] * si) / den;
    SS[i] = ((long double)a[1][i] * ssi - (long double)aa[1][i] * si) / den;
  }
  for (int i = 0; i < n; i++) {
    s[i] = (double)S[i];
    ss[i] = (double)SS[i];
  }
  free(S);
}

/* the boundary closure shared by CCD and UCCD (ccd_base.cpp:63-101) */
static void orc_compact_ends(double *c, int n, double h) {
  double *a[3] = {c, c + n, c + 2 * n}, *b[3] = {c + 3 * n, c + 4 * n, c + 5 * n};
  double *aa[3] = {c + 6 * n, c + 7 * n, c + 8 * n}, *bb[3] = {c + 9 * n, c + 10 * n, c + 11 * n};
  const double alpha_1 = 2.0, beta_1 = -1.0, alpha_2 = 5.0, beta_2 = -6.0;
  a[0][0] = 0.0; a[1][0] = 1.0; a[2][0] = alpha_1;
  b[0][0] = 0.0; b[1][0] = 0.0; b[2][0] = beta_1 * h;
  aa[0][0] = 0.0; aa[1][0] = 0.0; aa[2][0] = beta_2 / h;
  bb[0][0] = 0.0; bb[1][0] = 1.0; bb[2][0] = alpha_2;
  a[0][n - 1] = alpha_1; a[1][n - 1] = 1.0; a[2][n - 1] = 0.0;
  b[0][n - 1] = -beta_1 * h; b[1][n - 1] = 0.0; b[2][n - 1] = 0.0;
  aa[0][n - 1] = -beta_2 / h; aa[1][n - 1] = 0.0; aa[2][n - 1] = 0.0;
  bb[0][n - 1] = alpha_2; bb[1][n - 1] = 1.0; bb[2][n - 1] = 0.0;
}

void orc_ccd_coefficients(int n, double h, double *c) {
  for (int i = 0; i < n; i++) {
    c[i] = 7.0 / 16.0; c[n + i] = 1.0; c[2 * n + i] = 7.0 / 16.0;
    c[3 * n + i] = h / 16.0; c[4 * n + i] = 0.0; c[5 * n + i] = -h / 16.0;
    c[6 * n + i] = -9.0 / 8.0 / h; c[7 * n + i] = 0.0; c[8 * n + i] = 9.0 / 8.0 / h;
    c[9 * n + i] = -1.0 / 8.0; c[10 * n + i] = 1.0; c[11 * n + i] = -1.0 / 8.0;
  }
  orc_compact_ends(c, n, h);
  orc_twin_dec(c, n);
}

void orc_uccd_coefficients(int n, double h, double *up, double *dn) {
  const double a1 = 0.875, b1 = 0.1251282341599089, b2 = -0.2487176584009104, b3 = 0.0001282341599089;
  for (int i = 0; i < n; i++) {
    up[i] = a1; up[n + i] = 1.0; up[2 * n + i] = 0.0;
    dn[i] = 0.0; dn[n + i] = 1.0; dn[2 * n + i] = a1;
    up[3 * n + i] = b1 * h; up[4 * n + i] = b2 * h; up[5 * n + i] = b3 * h;
    dn[3 * n + i] = -b3 * h; dn[4 * n + i] = -b2 * h; dn[5 * n + i] = -b1 * h;
    up[6 * n + i] = -9.0 / 8.0 / h; up[7 * n + i] = 0.0; up[8 * n + i] = 9.0 / 8.0 / h;
    dn[6 * n + i] = -9.0 / 8.0 / h; dn[7 * n + i] = 0.0; dn[8 * n + i] = 9.0 / 8.0 / h;
    up[9 * n + i] = -1.0 / 8.0; up[10 * n + i] = 1.0; up[11 * n + i] = -1.0 / 8.0;
    dn[9 * n + i] = -1.0 / 8.0; dn[10 * n + i] = 1.0; dn[11 * n + i] = -1.0 / 8.0;
  }
  orc_compact_ends(up, n, h);
  orc_compact_ends(dn, n, h);
  orc_twin_dec(up, n);
  orc_twin_dec(dn, n);
}

/* right-hand sides s, ss (contiguous) of one padded line f (stride st,
 * length n): the one-sided ends and the interior stencil of the scheme;
 * kind 0 CCD, 1 UCCD upwind, 2 UCCD downwind */
static void orc_compact_rhs(const double *f, long st, int n, double h, int kind, double *s, double *ss) {
  const double bda1 = -3.5, bdb1 = 4.0, bdc1 = -0.5, bda2 = 9.0, bdb2 = -12.0, bdc2 = 3.0;
  const double c3 = -0.06096119008109, c2 = 1.99692238016218, c1 = -1.93596119008109;
#define F(i) f[(long)(i) * st]
  s[0] = (bda1 * F(0) + bdb1 * F(1) + bdc1 * F(2)) / h;
  ss[0] = (bda2 * F(0) + bdb2 * F(1) + bdc2 * F(2)) / h / h;
  s[n - 1] = -(bda1 * F(n - 1) + bdb1 * F(n - 2) + bdc1 * F(n - 3)) / h;
  ss[n - 1] = (bda2 * F(n - 1) + bdb2 * F(n - 2) + bdc2 * F(n - 3)) / h / h;
  for (int i = 1; i < n - 1; i++) {
    double v;
    if (kind == 0) v = 15.0 / 16.0 * (F(i + 1) - F(i - 1)) / h;
    else if (kind == 1) v = (c1 * F(i - 1) + c2 * F(i) + c3 * F(i + 1)) / h;
    else v = -(c3 * F(i - 1) + c2 * F(i) + c1 * F(i + 1)) / h;
    s[i] = v;
    ss[i] = (3.0 * F(i - 1) - 6.0 * F(i) + 3.0 * F(i + 1)) / h / h;
  }
#undef F
}

/* every padded line along axis a: CCD (vel = NULL) -> d1 = s, d2 = ss;
 * UCCD -> d1 = upwind or downwind by vel > 0, d2 = mean of the two ss */
static void orc_compact_axis(const orc_pf *p, int a, double h, const double *f, const double *vel,
                             double *d1, double *d2, int extended) {
  const int n = p->N[a];
  const long st = p->s[a];
  double *cu = malloc(sizeof(double) * 12 * n), *cd = malloc(sizeof(double) * 12 * n);
  double *su = malloc(sizeof(double) * 4 * n);
  double *ssu = su + n, *sd = su + 2 * n, *ssd = su + 3 * n;
  if (vel) orc_uccd_coefficients(n, h, cu, cd);
  else orc_ccd_coefficients(n, h, cu);
  const int t1 = (a + 1) % 3, t2 = (a + 2) % 3;
  for (int u = 0; u < p->N[t1]; u++)
    for (int v = 0; v < p->N[t2]; v++) {
      const long base = (long)u * p->s[t1] + (long)v * p->s[t2];
      orc_compact_rhs(f + base, st, n, h, vel ? 1 : 0, su, ssu);
      if (vel) {
        orc_compact_rhs(f + base, st, n, h, 2, sd, ssd);
        /* the downwind ends equal the upwind ones (uccd_solver.cpp:47-50) */
        sd[0] = su[0]; ssd[0] = ssu[0]; sd[n - 1] = su[n - 1]; ssd[n - 1] = ssu[n - 1];
      }
      if (extended) {
        orc_twin_bks_ld(cu, n, su, ssu);
        if (vel) orc_twin_bks_ld(cd, n, sd, ssd);
      } else {
        orc_twin_bks(cu, n, su, ssu, 1);
        if (vel) orc_twin_bks(cd, n, sd, ssd, 1);
      }
      for (int i = 0; i < n; i++) {
        const long c = base + (long)i * st;
        if (vel) {
          d1[c] = vel[c] > 0 ? su[i] : sd[i];
          d2[c] = 0.5 * (ssu[i] + ssd[i]);
        } else {
          d1[c] = su[i];
          d2[c] = ssu[i];
        }
      }
    }
  free(cu);
  free(cd);
  free(su);
}

/* uccd_solver::find_derivatives(f, vel) and ccd_solver::find_derivatives(f)
 * on the padded flux (vel_x/y/z = NULL for CCD) */
void orc_compact_derivatives(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, const double *flux,
                             const double *u, const double *v, const double *w, int upwind, double *fx,
                             double *fy, double *fz, double *fxx, double *fyy, double *fzz) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const double hh[3] = {dx, dy, dz};
  const double *vel[3] = {u, v, w};
  double *d1[3] = {fx, fy, fz}, *d2[3] = {fxx, fyy, fzz};
  for (int a = 0; a < ndim; a++) orc_compact_axis(&p, a, hh[a], flux, upwind ? vel[a] : NULL, d1[a], d2[a], 0);
}

/* orc_compact_derivatives with the twin sweeps in long double
 * (orc_twin_bks_ld): the derivatives to within about 2^-11 of the double
 * sweep's rounding error */
void orc_compact_derivatives_ld(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                                const double *flux, const double *u, const double *v, const double *w, int upwind,
                                double *fx, double *fy, double *fz, double *fxx, double *fyy, double *fzz) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const double hh[3] = {dx, dy, dz};
  const double *vel[3] = {u, v, w};
  double *d1[3] = {fx, fy, fz}, *d2[3] = {fxx, fyy, fzz};
  for (int a = 0; a < ndim; a++) orc_compact_axis(&p, a, hh[a], flux, upwind ? vel[a] : NULL, d1[a], d2[a], 1);
}

/* runge_kutta::tvd_rk3(phi, nvel, identity_flux, convection) with the UCCD
 * convection term (source.cpp:8-40, uccd_solver.cpp:14-215) and the
 * scalar BCs after each stage (runge_kutta.cpp:13-49)                  */
static void orc_tvd_rk3_uccd_impl(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                                  const int *types, const double *values, double *phi, const double *u,
                                  const double *v, const double *w, int nsteps, int extended) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const long T = (long)p.N[0] * p.N[1] * p.N[2];
  double *m = calloc((size_t)9 * T, sizeof(double));
  double *s[3] = {m, m + T, m + 2 * T}, *fx = m + 3 * T, *fy = m + 4 * T, *fz = m + 5 * T;
  double *fxx = m + 6 * T, *fyy = m + 7 * T, *fzz = m + 8 * T;
  for (int step = 0; step < nsteps; step++)
    for (int stage = 0; stage < 3; stage++) {
      if (extended)
        orc_compact_derivatives_ld(ndim, nx, ny, nz, dx, dy, dz, phi, u, v, w, 1, fx, fy, fz, fxx, fyy, fzz);
      else
        orc_compact_derivatives(ndim, nx, ny, nz, dx, dy, dz, phi, u, v, w, 1, fx, fy, fz, fxx, fyy, fzz);
      double *r = s[stage];
      for (long c = 0; c < T; c++) r[c] = -u[c] * fx[c];
      if (ndim > 1)
        for (long c = 0; c < T; c++) r[c] -= v[c] * fy[c];
      if (ndim > 2)
        for (long c = 0; c < T; c++) r[c] -= w[c] * fz[c];
      for (long c = 0; c < T; c++) {
        if (stage == 0) phi[c] += dt * s[0][c];
        else if (stage == 1) phi[c] += dt * (s[1][c] - 3.0 * s[0][c]) / 4.0;
        else phi[c] += dt * (-s[0][c] - s[1][c] + 8.0 * s[2][c]) / 12.0;
      }
      orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, types, values, -1, phi);
    }
  free(m);
}

void orc_tvd_rk3_uccd(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                      const int *types, const double *values, double *phi, const double *u, const double *v,
                      const double *w, int nsteps) {
  orc_tvd_rk3_uccd_impl(ndim, nx, ny, nz, dx, dy, dz, dt, types, values, phi, u, v, w, nsteps, 0);
}

/* orc_tvd_rk3_uccd on orc_compact_derivatives_ld's derivatives */
void orc_tvd_rk3_uccd_ld(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt,
                         const int *types, const double *values, double *phi, const double *u, const double *v,
                         const double *w, int nsteps) {
  orc_tvd_rk3_uccd_impl(ndim, nx, ny, nz, dx, dy, dz, dt, types, values, phi, u, v, w, nsteps, 1);
}

/* lsf_mass (wrapper_func.cpp:450-511): density-weighted Heaviside mass,
 * 24-weight (2-D) / 77-weight (3-D) stencil on interior cells away from
 * the walls; the 3-D branch multiplies by dv twice, as the reference does */
double orc_lsf_mass(int ndim, int nx, int ny, int nz, double dv, double ls_width, double ratio, int positive_ref,
                    const double *phi) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const long T = (long)p.N[0] * p.N[1] * p.N[2];
  double *dn = malloc(sizeof(double) * 2 * T), *hv = dn + T;
  for (long c = 0; c < T; c++) {
    const double H = orc_heav(phi[c], ls_width);
    hv[c] = H;
    dn[c] = positive_ref ? H + (1.0 - H) * ratio : (1.0 - H) + H * ratio;
  }
  double mass = 0.0;
  for (int I = 0; I < p.n[0]; I++)
    for (int J = 0; J < p.n[1]; J++)
      for (int K = 0; K < p.n[2]; K++) {
        const long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
        if (ndim == 2) {
          if (I != 0 && I != p.n[0] - 1 && J != 0 && J != p.n[1] - 1) {
            double lm = 15.0 * dn[c] * hv[c];
            for (int ii = -1; ii < 2; ii++)
              for (int jj = -1; jj < 2; jj++) {
                const long q = c + ii * p.s[0] + jj * p.s[1];
                lm += dn[q] * hv[q];
              }
            mass += lm / 24.0;
          } else {
            mass += dn[c] * hv[c];
          }
        } else {
          if (I != 0 && I != p.n[0] - 1 && J != 0 && J != p.n[1] - 1 && K != 0 && K != p.n[2] - 1) {
            double lm = 50.0 * dn[c] * hv[c];
            for (int ii = -1; ii < 2; ii++)
              for (int jj = -1; jj < 2; jj++)
                for (int kk = -1; kk < 2; kk++) {
                  const long q = c + ii * p.s[0] + jj * p.s[1] + kk * p.s[2];
                  lm += dn[q] * hv[q];
                }
            mass += lm * dv / 77.0;
          } else {
            mass += dn[c] * hv[c] * dv;
          }
        }
      }
  free(dn);
  return mass * dv;
}

/* gradient |grad phi| of the CCD derivatives (find_gradient,
 * wrapper_func.cpp:42-57) */
static double orc_grad_mag(int ndim, double fx, double fy, double fz) {
  double g = pow(fx, 2);
  if (ndim > 1) g += pow(fy, 2);
  if (ndim > 2) g += pow(fz, 2);
  return sqrt(g);
}

/* `count` runge_kutta::euler steps with the mpls right-hand side
 * (source.cpp:109-146): identity_with_extrapolation + CCD derivatives of
 * phi, lsf_mass, Heaviside, delta and gradient, the global factor
 * eta = (mass0 - mass) / (sum g delta |grad| dv dt), s = eta delta |grad|;
 * phi += dt s and the scalar BCs.  delta_out (nullable): the last delta
 * (dummy->delta after the call).                                       */
void orc_mpls_euler(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt, double ls_width,
                    double ratio, int positive_ref, double mass0, const int *types, const double *values,
                    double *phi, int count, double *delta_out) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const long T = (long)p.N[0] * p.N[1] * p.N[2];
  const double dv = ndim == 2 ? dx * dy : dx * dy * dz;
  double *m = calloc((size_t)9 * T, sizeof(double));
  double *fl = m, *d1[3] = {m + T, m + 2 * T, m + 3 * T}, *d2[3] = {m + 4 * T, m + 5 * T, m + 6 * T};
  double *dl = m + 7 * T, *gr = m + 8 * T;
  for (int it = 0; it < count; it++) {
    orc_extrap(&p, phi, fl, -1);
    orc_compact_derivatives(ndim, nx, ny, nz, dx, dy, dz, fl, NULL, NULL, NULL, 0, d1[0], d1[1], d1[2], d2[0],
                            d2[1], d2[2]);
    const double mass = orc_lsf_mass(ndim, nx, ny, nz, dv, ls_width, ratio, positive_ref, phi);
    for (long c = 0; c < T; c++) {
      dl[c] = orc_delta(phi[c], ls_width);
      gr[c] = orc_grad_mag(ndim, d1[0][c], d1[1][c], d1[2][c]);
    }
    double eta = 0.0;
    for (int I = 0; I < p.n[0]; I++)
      for (int J = 0; J < p.n[1]; J++)
        for (int K = 0; K < p.n[2]; K++) {
          const long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
          const double H = orc_heav(phi[c], ls_width), d = dl[c];
          double g = d * (2.0 * (1.0 - ratio) * H + ratio);
          if (!positive_ref) g = d * (2.0 * (ratio - 1.0) * H + 1.0);
          eta += g * d * gr[c];
        }
    eta = (mass0 - mass) / (eta * dv * dt);
    for (long c = 0; c < T; c++) phi[c] += dt * (eta * dl[c] * gr[c]);
    orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, types, values, -1, phi);
  }
  if (delta_out) memcpy(delta_out, dl, sizeof(double) * T);
  free(m);
}

/* weno5_for_godunov (godunov_gradient.cpp:184-201) */
static double orc_weno_godunov(double a, double b, double c, double d) {
  const double eps = orc_eps;
  const double is0 = 13.0 * pow(a - b, 2.0) + 3.0 * pow(a - 3.0 * b, 2.0);
  const double is1 = 13.0 * pow(b - c, 2.0) + 3.0 * pow(b + c, 2.0);
  const double is2 = 13.0 * pow(c - d, 2.0) + 3.0 * pow(3.0 * c - d, 2.0);
  const double alp0 = 1.0 / pow(eps + is0, 2.0);
  const double alp1 = 6.0 / pow(eps + is1, 2.0);
  const double alp2 = 3.0 / pow(eps + is2, 2.0);
  const double w0 = alp0 / (alp0 + alp1 + alp2);
  const double w2 = alp2 / (alp0 + alp1 + alp2);
  return w0 / 3.0 * (a - 2.0 * b + c) + (w2 - 0.5) / 6.0 * (b - 2.0 * c + d);
}

/* godunov_gradient (godunov_gradient.cpp:7-152): Godunov upwind |grad phi|
 * on the interior from WENO5 one-sided differences, chosen by sign; 0 on
 * the ghosts */
static void orc_godunov(const orc_pf *p, const double *hh, const double *phi, const double *sign, double *grad) {
  const long T = (long)p->N[0] * p->N[1] * p->N[2];
  for (long c = 0; c < T; c++) grad[c] = 0.0;
  for (int a = 0; a < p->ndim; a++) {
    const long s = p->s[a];
    const double h = hh[a];
    for (int I = 0; I < p->n[0]; I++)
      for (int J = 0; J < p->n[1]; J++)
        for (int K = 0; K < p->n[2]; K++) {
          const long c = PIDX(*p, I + p->off[0], J + p->off[1], K + p->off[2]);
#define P(q) phi[c + (q) * s]
          const double v = 1.0 / (12.0 * h) *
                           (-(P(-1) - P(-2)) + 7.0 * (P(0) - P(-1)) + 7.0 * (P(1) - P(0)) - (P(2) - P(1)));
          const double up = v + 1.0 / h *
                                    orc_weno_godunov(P(3) - 2.0 * P(2) + P(1), P(2) - 2.0 * P(1) + P(0),
                                                     P(1) - 2.0 * P(0) + P(-1), P(0) - 2.0 * P(-1) + P(-2));
          const double um = v - 1.0 / h *
                                    orc_weno_godunov(P(-3) - 2.0 * P(-2) + P(-1), P(-2) - 2.0 * P(-1) + P(0),
                                                     P(-1) - 2.0 * P(0) + P(1), P(0) - 2.0 * P(1) + P(2));
#undef P
          double g;
          if (sign[c] > 0.0) {
            const double upm = -fmin(up, 0.0), ump = fmax(um, 0.0);
            g = pow(fmax(upm, ump), 2);
          } else {
            const double upp = fmax(up, 0.0), upm = -fmin(um, 0.0);
            g = pow(fmax(upp, upm), 2);
          }
          if (a == 0) grad[c] = g;
          else grad[c] += g;
        }
  }
  for (long c = 0; c < T; c++) grad[c] = sqrt(grad[c]);
}

/* zero_order_extrapolation (simple_bc.cpp:7-43) over the padded array */
static void orc_zero_extrap(const orc_pf *p, double *f) {
  const int g = ORC_GHC;
  for (int a = 0; a < 3; a++) {
    if (a > 0 && p->N[a] <= 1) continue;
    const int t1 = (a + 1) % 3, t2 = (a + 2) % 3, N = p->N[a];
    for (int u = 0; u < p->N[t1]; u++)
      for (int v = 0; v < p->N[t2]; v++) {
        const long base = (long)u * p->s[t1] + (long)v * p->s[t2];
        for (int i = 0; i < g; i++) {
          f[base + i * p->s[a]] = f[base + g * p->s[a]];
          f[base + (N - 1 - i) * p->s[a]] = f[base + (N - 1 - g) * p->s[a]];
        }
      }
  }
}

/* Sign (lsm.cpp:33-43) of every padded cell (find_sign) */
void orc_find_sign(long n, const double *phi, double ls_width, double *sign) {
  for (long c = 0; c < n; c++) {
    const double x = phi[c];
    if (x < -ls_width) sign[c] = -1;
    else if (x > ls_width) sign[c] = 1;
    else sign[c] = x / sqrt(x * x + ls_width * ls_width);
  }
}

/* `nsteps` runge_kutta::tvd_rk3 steps with lsf_redistance_lambda
 * (source.cpp:148-240): s = -sign (|grad| - 1) rdt/dt with the Godunov
 * gradient, plus the interior mass-constraint term a/b delta |grad|
 * rdt/dt from the zero-order-extrapolated a = sign delta (|grad| - 1),
 * b = |grad| delta^2 (16-weight centre + 4 diagonal neighbours in 2-D,
 * 51 + 8 in 3-D), |b| > 1e-8.  sign and delta are the dummy arrays the
 * reference reads (find_sign / find_delta of earlier calls).            */
/* stabilized_upon_gradient (godunov_gradient.cpp:159-181): the Godunov
 * gradient with the stored sign, its maximum over every padded cell (a
 * running fmax from 0.0) and phi /= that maximum on every cell.        */
double orc_stabilize_gradient(int ndim, int nx, int ny, int nz, double dx, double dy, double dz,
                              const double *sign, double *phi, double *grad) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const long T = (long)p.N[0] * p.N[1] * p.N[2];
  const double hh[3] = {dx, dy, dz};
  orc_godunov(&p, hh, phi, sign, grad);
  double mx = 0.0;
  for (long c = 0; c < T; c++) mx = fmax(mx, grad[c]);
  for (long c = 0; c < T; c++) phi[c] /= mx;
  return mx;
}

void orc_redistance_rk3(int ndim, int nx, int ny, int nz, double dx, double dy, double dz, double dt, double rdt,
                        const double *sign, const double *delta, const int *types, const double *values,
                        double *phi, int nsteps) {
  const orc_pf p = pf_make(ndim, nx, ny, nz);
  const long T = (long)p.N[0] * p.N[1] * p.N[2];
  const double hh[3] = {dx, dy, dz};
  double *m = calloc((size_t)6 * T, sizeof(double));
  double *s[3] = {m, m + T, m + 2 * T}, *gr = m + 3 * T, *A = m + 4 * T, *B = m + 5 * T;
  for (int step = 0; step < nsteps; step++)
    for (int stage = 0; stage < 3; stage++) {
      double *r = s[stage];
      orc_godunov(&p, hh, phi, sign, gr);
      for (long c = 0; c < T; c++) r[c] = -sign[c] * (gr[c] - 1.0) * rdt / dt;
      for (long c = 0; c < T; c++) {
        A[c] = sign[c] * delta[c] * (gr[c] - 1.0);
        B[c] = gr[c] * delta[c] * delta[c];
      }
      orc_zero_extrap(&p, A);
      orc_zero_extrap(&p, B);
      for (int I = 0; I < p.n[0]; I++)
        for (int J = 0; J < p.n[1]; J++)
          for (int K = 0; K < p.n[2]; K++) {
            const long c = PIDX(p, I + p.off[0], J + p.off[1], K + p.off[2]);
            double a, b;
            if (ndim == 2) {
              a = 16.0 * A[c];
              b = 16.0 * B[c];
              for (int ii = -1; ii < 2; ii++)
                for (int jj = -1; jj < 2; jj++)
                  if (ii != 0 && jj != 0) {
                    a += A[c + ii * p.s[0] + jj * p.s[1]];
                    b += B[c + ii * p.s[0] + jj * p.s[1]];
                  }
            } else {
              a = 51.0 * A[c];
              b = 51.0 * B[c];
              for (int ii = -1; ii < 2; ii++)
                for (int jj = -1; jj < 2; jj++)
                  for (int kk = -1; kk < 2; kk++)
                    if (ii != 0 && jj != 0 && kk != 0) {
                      a += A[c + ii * p.s[0] + jj * p.s[1] + kk * p.s[2]];
                      b += B[c + ii * p.s[0] + jj * p.s[1] + kk * p.s[2]];
                    }
            }
            if (fabs(b) > 1e-8) r[c] += a / b * delta[c] * gr[c] * rdt / dt;
          }
      for (long c = 0; c < T; c++) {
        if (stage == 0) phi[c] += dt * s[0][c];
        else if (stage == 1) phi[c] += dt * (s[1][c] - 3.0 * s[0][c]) / 4.0;
        else phi[c] += dt * (-s[0][c] - s[1][c] + 8.0 * s[2][c]) / 12.0;
      }
      orc_bc_apply(ndim, nx, ny, nz, dx, dy, dz, types, values, -1, phi);
    }
  free(m);
}
