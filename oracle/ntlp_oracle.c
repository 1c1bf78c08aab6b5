/*
 * ntlp_oracle.c -- CPU FP64 restatement of NTLP's Lagrangian droplet step.
 *
 * TEST INFRASTRUCTURE ONLY (see ntlp_oracle.h).  Single-threaded, droplet by
 * droplet in array order, like one MPI rank of the reference walking its
 * linked list.  Built with -O2 -ffp-contract=off so every operation rounds
 * where the Fortran expression rounds.
 *
 * Cited files are those of the Fortran reference: particles.f90, les.F,
 * defs.F.
 */
#include "ntlp_oracle.h"
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#define NR OR_NREAL

/* ------------------------------------------------------------------ helpers */

/* les.F:10764-10774 (Alduchov & Eskridge Magnus form) */
double or_mod_magnus(double T) {
  return 610.94 * exp((17.6257 * (T - 273.15)) / (243.04 + (T - 273.15)));
}

/* les.F:10775-10784 */
double or_exner(double p0, double p, const or_consts *c) {
  return pow(p0 / p, -c->Rd / c->Cpa);
}

/* les.F:10787-10795 */
double or_func_p_base(double p_surf, double T_surf, double z, const or_consts *c) {
  return p_surf * pow(1.0 - z * c->grav / c->Cpa / T_surf, c->Cpa / c->Rd);
}

/* --------------------------------------------------------------- cell index */

/* Fortran minloc(a, 1, mask=(a .gt. x)) over a(0:nnz+1): 1-based position of
 * the first element greater than x, or 0 when none is (the arrays are
 * monotone, so the first masked element is the minimum). */
static int minloc_gt(const double *a, int len, double x) {
  for (int p = 0; p < len; ++p)
    if (a[p] > x) return p + 1;
  return 0;
}

/* particles.f90:780-788 (interp), 1180-1183 (coupling), 3332-3334 (tauc) */
void or_cell_index(const or_grid *g, double x, double y, double z, int32_t out[5]) {
  int len = g->nnz + 2;
  out[0] = (int32_t)floor(x / g->dx) + 1;
  out[1] = (int32_t)floor(y / g->dy) + 1;
  int mzz = minloc_gt(g->zz, len, z);
  int mz = minloc_gt(g->z, len, z);
  out[2] = mzz - 2;
  out[3] = mz - 2;
  out[4] = mz - 1;
}

void or_cell_index_many(const or_grid *g, const double *x, const double *y, const double *z,
                        int64_t n, int32_t *out) {
  for (int64_t i = 0; i < n; ++i) or_cell_index(g, x[i], y[i], z[i], out + 5 * i);
}

/* --------------------------------------------------------- counting sort */

int64_t or_num_ucells(const or_grid *g) {
  return (int64_t)(g->nnx + 1) * (g->nny + 1) * (g->nnz + 1);
}

static int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* u-cell key of one droplet.  ipt in [1,nnx+1], jpt in [1,nny+1] and kpt in
 * [0,nnz] cover every in-domain position (x==xl is kept by the strict `>`
 * of particle_bcs_periodic, particles.f90:2826).  Out-of-domain indices
 * (kpt = -1/-2 from an empty minloc) are clamped into that box. */
static int32_t ukey(const or_grid *g, double x, double y, double z) {
  int32_t ci[5];
  or_cell_index(g, x, y, z, ci);
  int i0 = clampi(ci[0] - 1, 0, g->nnx);
  int j0 = clampi(ci[1] - 1, 0, g->nny);
  int k0 = clampi(ci[2], 0, g->nnz);
  return k0 + (g->nnz + 1) * (j0 + (g->nny + 1) * i0);
}

/* Stable counting sort (histogram -> exclusive scan -> scatter in index
 * order).  The reference keeps droplets in a linked list and has no sort;
 * the sort's contract is fixed by the key above and stability. */
void or_bin(const or_grid *g, const double *x, const double *y, const double *z, int64_t n,
            int32_t *keys, int32_t *counts, int32_t *offsets, int32_t *perm) {
  int64_t nc = or_num_ucells(g);
  memset(counts, 0, sizeof(int32_t) * nc);
  for (int64_t i = 0; i < n; ++i) {
    keys[i] = ukey(g, x[i], y[i], z[i]);
    counts[keys[i]]++;
  }
  offsets[0] = 0;
  for (int64_t c = 0; c < nc; ++c) offsets[c + 1] = offsets[c] + counts[c];
  int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * nc);
  memcpy(cur, offsets, sizeof(int32_t) * nc);
  for (int64_t i = 0; i < n; ++i) perm[cur[keys[i]]++] = (int32_t)i;
  free(cur);
}

/* calc_tauc_min's num_array (particles.f90:3320-3339): w-cell kc=minloc(z)-1 */
void or_wcell_counts(const or_grid *g, const double *x, const double *y, const double *z,
                     int64_t n, int32_t *num) {
  memset(num, 0, sizeof(int32_t) * (size_t)g->nnx * g->nny * g->nnz);
  for (int64_t i = 0; i < n; ++i) {
    int32_t ci[5];
    or_cell_index(g, x[i], y[i], z[i], ci);
    int ix = ci[0], iy = ci[1], kc = ci[4];
    if (ix < 1 || ix > g->nnx || iy < 1 || iy > g->nny || kc < 1 || kc > g->nnz) continue;
    num[(ix - 1) + (size_t)g->nnx * ((iy - 1) + (size_t)g->nny * (kc - 1))]++;
  }
}

/* ------------------------------------------------------------- ext arrays */

static inline int64_t ext_off(const or_grid *g, int k, int j, int i) {
  /* (0:nnz+1, -1:nny+3, -1:nnx+3), les.F:8441 with iys=mxs=1 */
  return (int64_t)k + (int64_t)(g->nnz + 2) * ((j + 1) + (int64_t)(g->nny + 5) * (i + 1));
}

/* fill_ext (particles.f90:91-328) on one periodic domain: halo columns
 * -1,0 and n+1..n+3 are periodic images of the core 1..n. */
void or_fill_ext(const or_grid *g, const double *core, double *ext) {
  int nx = g->nnx, ny = g->nny, nz2 = g->nnz + 2;
  for (int i = -1; i <= nx + 3; ++i) {
    int ic = ((i - 1) % nx + nx) % nx + 1;
    for (int j = -1; j <= ny + 3; ++j) {
      int jc = ((j - 1) % ny + ny) % ny + 1;
      const double *src = core + (int64_t)nz2 * ((jc - 1) + (int64_t)ny * (ic - 1));
      double *dst = ext + ext_off(g, 0, j, i);
      memcpy(dst, src, sizeof(double) * nz2);
    }
  }
}

/* ----------------------------------------------------------- interpolation */

/* Accessors into the record of one droplet (p[OR_*]). */

/* uf_interp_lin (particles.f90:767-847), including the zeroth-order wall
 * override inside the (i,j) loop (831-842).  k indices are clamped into
 * 0..nnz+1 only where the reference would read outside its array. */
static void interp_lin(const or_grid *g, const or_fields *fl, double *p) {
  const double dx = g->dx, dy = g->dy;
  int32_t ci[5];
  or_cell_index(g, p[OR_X], p[OR_Y], p[OR_Z], ci);
  int ipt = ci[0], jpt = ci[1], kpt = ci[2], kwpt = ci[3];
  int kmax = g->nnz + 1;
  double uf1 = 0.0, uf2 = 0.0, uf3 = 0.0, Tf = 0.0, qf = 0.0;
  int nnz = g->nnz;
  int kp1 = clampi(kpt + 1, 0, g->nnz + 2), kwp1 = clampi(kwpt + 1, 0, g->nnz + 2);
  for (int i = 0; i <= 1; ++i) {
    int ix = 0, iy = 0;
    for (int j = 0; j <= 1; ++j) {
      for (int k = 0; k <= 1; ++k) {
        double xv = dx * (double)(i + ipt - 1);
        double yv = dy * (double)(j + jpt - 1);
        int izuv = clampi(k + kpt, 0, kmax);
        int izw = clampi(k + kwpt, 0, kmax);
        double zv = g->zz[izuv];
        double zwv = g->z[izw];
        double wtx = (1.0 - fabs(p[OR_X] - xv) / dx);
        double wty = (1.0 - fabs(p[OR_Y] - yv) / dy);
        double wtz = (1.0 - fabs(p[OR_Z] - zv) / g->dzu[kp1]);
        double wtzw = (1.0 - fabs(p[OR_Z] - zwv) / g->dzw[kwp1]);
        ix = ipt + i;
        iy = jpt + j;
        int64_t o = ext_off(g, izuv, iy, ix), ow = ext_off(g, izw, iy, ix);
        uf1 = uf1 + fl->u[o] * wtx * wty * wtz;
        uf2 = uf2 + fl->v[o] * wtx * wty * wtz;
        uf3 = uf3 + fl->w[ow] * wtx * wty * wtzw;
        Tf = Tf + fl->T[o] * wtx * wty * wtz;
        qf = qf + fl->q[o] * wtx * wty * wtz;
      }
      if (kpt == nnz) {
        int64_t o = ext_off(g, kpt, iy, ix);
        uf1 = fl->u[o]; uf2 = fl->v[o]; Tf = fl->T[o]; qf = fl->q[o];
      }
      if (kpt == 0) {
        int64_t o = ext_off(g, 1, iy, ix);
        uf1 = fl->u[o]; uf2 = fl->v[o]; Tf = fl->T[o]; qf = fl->q[o];
      }
    }
  }
  p[OR_UFX] = uf1; p[OR_UFY] = uf2; p[OR_UFZ] = uf3; p[OR_TF] = Tf; p[OR_QINF] = qf;
}

/* uf_interp (particles.f90:554-765): 6th-order Lagrange on 6x6x6 nodes with
 * the reference's one-sided truncation tables near the walls.  Node indices
 * that the reference reads with a zero weight are clamped into the array. */
static void interp6(const or_grid *g, const or_fields *fl, double *p) {
  int ijpts[2][6], kuv[6], kw[6];
  double wt[4][6];
  int nnz = g->nnz, first, last;
  int32_t ci[5];
  or_cell_index(g, p[OR_X], p[OR_Y], p[OR_Z], ci);
  ijpts[0][2] = ci[0];
  ijpts[1][2] = ci[1];
  for (int d = 0; d < 2; ++d) {
    ijpts[d][1] = ijpts[d][2] - 1;
    ijpts[d][0] = ijpts[d][1] - 1;
    ijpts[d][3] = ijpts[d][2] + 1;
    ijpts[d][4] = ijpts[d][3] + 1;
    ijpts[d][5] = ijpts[d][4] + 1;
  }
  kuv[2] = ci[2];
  kuv[3] = kuv[2] + 1; kuv[4] = kuv[3] + 1; kuv[5] = kuv[4] + 1;
  kuv[1] = kuv[2] - 1; kuv[0] = kuv[1] - 1;
  kw[2] = ci[3];
  kw[3] = kw[2] + 1; kw[4] = kw[3] + 1; kw[5] = kw[4] + 1;
  kw[1] = kw[2] - 1; kw[0] = kw[1] - 1;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 6; ++b) wt[a][b] = 0.0;
  double dxvec[2] = {g->dx, g->dy};
  double xpv[2] = {p[OR_X], p[OR_Y]};
  for (int iz = 0; iz < 2; ++iz) {
    for (int j = 0; j < 6; ++j) {
      double xjval = dxvec[iz] * (double)(ijpts[iz][j] - 1);
      double pj = 1.0;
      for (int k = 0; k < 6; ++k) {
        double xkval = dxvec[iz] * (double)(ijpts[iz][k] - 1);
        if (j != k) pj = pj * (xpv[iz] - xkval) / (xjval - xkval);
      }
      wt[iz][j] = pj;
    }
  }
  /* u,v levels: particles.f90:623-676 (first/last are 1-based there) */
  if (kuv[2] == 1) {
    first = 3; last = 4; kuv[0] = 1; kuv[1] = 1;
  } else if (kuv[2] == 0) {
    first = 4; last = 5; kuv[0] = 1; kuv[1] = 1; kuv[2] = 1;
  } else if (kuv[2] < 0) {
    first = 0; last = 0;
  } else if (kuv[2] == 2) {
    first = 2; last = 5;
  } else if (kuv[2] == nnz) {
    first = 2; last = 3; kuv[3] = nnz; kuv[4] = nnz; kuv[5] = nnz;
  } else if (kuv[2] > nnz) {
    first = 0; last = 0;
  } else if (kuv[2] == nnz - 1) {
    first = 3; last = 4; kuv[4] = nnz; kuv[5] = nnz;
  } else if (kuv[2] == nnz - 2) {
    first = 2; last = 5;
  } else {
    first = 1; last = 6;
  }
  for (int j = first; j <= last && first > 0; ++j) {
    double xjval = g->zz[clampi(kuv[j - 1], 0, nnz + 1)];
    double pj = 1.0;
    for (int k = first; k <= last; ++k) {
      double xkval = g->zz[clampi(kuv[k - 1], 0, nnz + 1)];
      if (j != k) pj = pj * (p[OR_Z] - xkval) / (xjval - xkval);
    }
    wt[2][j - 1] = pj;
  }
  /* w levels: particles.f90:690-728 */
  if (kw[2] == 0) {
    first = 3; last = 4; kw[0] = 1; kw[1] = 1;
  } else if (kw[2] < 0) {
    first = 0; last = 0; kw[0] = 1; kw[1] = 1; kw[2] = 1;
  } else if (kw[2] == 1) {
    first = 2; last = 5; kw[0] = 1;
  } else if (kw[2] == nnz - 1) {
    first = 3; last = 4; kw[4] = nnz; kw[5] = nnz;
  } else if (kw[2] >= nnz) {
    first = 0; last = 0; kw[2] = nnz; kw[3] = nnz; kw[4] = nnz; kw[5] = nnz;
  } else if (kw[2] == nnz - 2) {
    first = 2; last = 5; kw[5] = nnz;
  } else {
    first = 1; last = 6;
  }
  for (int j = first; j <= last && first > 0; ++j) {
    double xjval = g->z[clampi(kw[j - 1], 0, nnz + 1)];
    double pj = 1.0;
    for (int k = first; k <= last; ++k) {
      double xkval = g->z[clampi(kw[k - 1], 0, nnz + 1)];
      if (j != k) pj = pj * (p[OR_Z] - xkval) / (xjval - xkval);
    }
    wt[3][j - 1] = pj;
  }
  double uf1 = 0, uf2 = 0, uf3 = 0, Tf = 0, qf = 0;
  for (int k = 0; k < 6; ++k)
    for (int j = 0; j < 6; ++j)
      for (int i = 0; i < 6; ++i) {
        int ix = ijpts[0][i], iy = ijpts[1][j];
        int izuv = clampi(kuv[k], 0, nnz + 1), izw = clampi(kw[k], 0, nnz + 1);
        int64_t o = ext_off(g, izuv, iy, ix), ow = ext_off(g, izw, iy, ix);
        uf1 = uf1 + fl->u[o] * wt[0][i] * wt[1][j] * wt[2][k];
        uf2 = uf2 + fl->v[o] * wt[0][i] * wt[1][j] * wt[2][k];
        uf3 = uf3 + fl->w[ow] * wt[0][i] * wt[1][j] * wt[3][k];
        Tf = Tf + fl->T[o] * wt[0][i] * wt[1][j] * wt[2][k];
        qf = qf + fl->q[o] * wt[0][i] * wt[1][j] * wt[2][k];
      }
  p[OR_UFX] = uf1; p[OR_UFY] = uf2; p[OR_UFZ] = uf3; p[OR_TF] = Tf; p[OR_QINF] = qf;
}

static void interp_one(const or_grid *g, const or_flags *f, const or_fields *fl, double *p) {
  if (f->ilin == 1) interp_lin(g, fl, p);
  else interp6(g, fl, p);
}

/* SoA <-> record */
static void load_rec(const double *soa, int64_t n, int64_t i, double *p) {
  for (int c = 0; c < NR; ++c) p[c] = soa[(int64_t)c * n + i];
}
static void store_rec(double *soa, int64_t n, int64_t i, const double *p) {
  for (int c = 0; c < NR; ++c) soa[(int64_t)c * n + i] = p[c];
}

void or_interp(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
               double *soa, int64_t n) {
  (void)c;
  double p[NR];
  for (int64_t i = 0; i < n; ++i) {
    load_rec(soa, n, i, p);
    interp_one(g, f, fl, p);
    store_rec(soa, n, i, p);
  }
}

/* ---------------------------------------------------------- crit radius */

/* setup_radval + crit_radius (particles.f90:5294-5327) */
double or_crit_radius(const or_consts *c, double m_s, double kappa_s, double Tf) {
  const double radstart = -8.0, radend = -3.0;
  const double dr = (radstart - radend) / 1024.0;
  double best = 0.0, rbest = 0.0;
  for (int i = 1; i <= 1024; ++i) {
    double rv = pow(10.0, radstart - (double)(i - 1) * dr);
    double ss = 2.0 * c->Mw * c->Gam / c->Ru / c->rhow / Tf / rv -
                kappa_s * m_s / (c->rhos * c->pi2 * 2.0 / 3.0 * (rv * rv * rv));
    if (i == 1 || ss > best) { best = ss; rbest = rv; }
  }
  return rbest;
}

void or_crit_radius_many(const or_consts *c, const double *m_s, const double *kappa_s,
                         const double *Tp, int64_t n, double *rc) {
  for (int64_t i = 0; i < n; ++i) rc[i] = or_crit_radius(c, m_s[i], kappa_s[i], Tp[i]);
}

/* ------------------------------------------------------------ BCs */

/* particle_bcs_nonperiodic, icase=0 reflect branch (2683-2812) followed by
 * particle_bcs_periodic (2814-2843). */
static void bcs(const or_grid *g, const or_flags *f, double *p) {
  double top = (f->icase == 4) ? g->z[g->nnz - 1] - p[OR_RADIUS] : g->z[g->nnz] - p[OR_RADIUS];
  double bot = 0.0 + p[OR_RADIUS];
  if (p[OR_Z] > top) {
    p[OR_Z] = top - (p[OR_Z] - top);
    p[OR_VZ] = -p[OR_VZ];
  } else if (p[OR_Z] < bot) {
    p[OR_Z] = bot + (bot - p[OR_Z]);
    p[OR_VZ] = -p[OR_VZ];
  }
  if (p[OR_X] > g->xl) p[OR_X] = p[OR_X] - g->xl;
  else if (p[OR_X] < 0) p[OR_X] = g->xl + p[OR_X];
  if (p[OR_Y] > g->yl) p[OR_Y] = p[OR_Y] - g->yl;
  else if (p[OR_Y] < 0) p[OR_Y] = g->yl + p[OR_Y];
}

/* air density and exner conversion of Tf (2895-2903, 3064-3072) */
static double air_rho(const or_consts *c, const or_flags *f, double *p) {
  if (f->iexner == 1) {
    double pb = or_func_p_base(c->surf_p, c->tsfcc1, p[OR_Z], c);
    p[OR_TF] = p[OR_TF] * or_exner(c->surf_p, pb, c);
    return or_func_p_base(c->surf_p, c->tsfcc1, p[OR_Z], c) / c->Rd / p[OR_TF];
  }
  return c->surf_rho;
}

/* ------------------------------------------------------------ RK3 */

/* Body of the droplet loop of particle_update_rk3 (2879-2973) after the
 * interpolation and the air-density set-up. */
static void rk3_body(const or_consts *c, const or_flags *f, double *p, double rhoa, int istage,
                     int it) {
  const double dt = c->dt;
  const double zs = c->zetas[istage - 1], gs = c->gama[istage - 1];
  if (it <= 1) {
    p[OR_VX] = p[OR_UFX]; p[OR_VY] = p[OR_UFY]; p[OR_VZ] = p[OR_UFZ];
    p[OR_TP] = p[OR_TF];
  }
  double r = p[OR_RADIUS];
  double diff0 = p[OR_VX] - p[OR_UFX], diff1 = p[OR_VY] - p[OR_UFY], diff2 = p[OR_VZ] - p[OR_UFZ];
  double diffnorm = sqrt(diff0 * diff0 + diff1 * diff1 + diff2 * diff2);
  double Rep = 2.0 * r * diffnorm / c->nuf;
  double Volp = c->pi2 * 2.0 / 3.0 * (r * r * r);
  double rhop = (p[OR_M_S] + Volp * c->rhow) / Volp;
  double taup_i = 18.0 * rhoa * c->nuf / rhop / ((2.0 * r) * (2.0 * r));
  double corrfac = (1.0 + 0.15 * pow(Rep, 0.687));
  double Nup = 2.0 + 0.6 * pow(Rep, 1.0 / 2.0) * pow(c->Pra, 1.0 / 3.0);
  double Shp = 2.0 + 0.6 * pow(Rep, 1.0 / 2.0) * pow(c->Sc, 1.0 / 3.0);
  double einf = or_mod_magnus(p[OR_TF]);
  double Eff_C = 2.0 * c->Mw * c->Gam / (c->Ru * c->rhow * r * p[OR_TP]);
  double Eff_S = p[OR_KAPPA_S] * p[OR_M_S] * c->rhow / c->rhos / (Volp * rhop - p[OR_M_S]);
  double estar = einf * exp(c->Mw * c->Lv / c->Ru * (1.0 / p[OR_TF] - 1.0 / p[OR_TP]) + Eff_C - Eff_S);
  double qstar = c->Mw / c->Ru * estar / p[OR_TP] / rhoa;
  p[OR_QSTAR] = qstar;

  double xtmp[3], vtmp[3];
  for (int d = 0; d < 3; ++d) {
    xtmp[d] = p[OR_X + d] + dt * zs * p[OR_XRHSX + d];
    vtmp[d] = p[OR_VX + d] + dt * zs * p[OR_VRHSX + d];
  }
  double Tptmp = p[OR_TP] + dt * zs * p[OR_TPRHS_S];
  Tptmp = Tptmp + dt * zs * p[OR_TPRHS_L];
  double radiustmp = r + dt * zs * p[OR_RADRHS];

  for (int d = 0; d < 3; ++d) {
    p[OR_XRHSX + d] = p[OR_VX + d];
    p[OR_VRHSX + d] = corrfac * taup_i * (p[OR_UFX + d] - p[OR_VX + d]) + c->part_grav[d];
  }
  if (f->ievap == 1)
    p[OR_RADRHS] = Shp / 9.0 / c->Sc * rhop / c->rhow * r * taup_i * (p[OR_QINF] - qstar);
  else
    p[OR_RADRHS] = 0.0;
  p[OR_TPRHS_S] = -(Nup / 3.0 / c->Pra * c->CpaCpp * rhop / c->rhow * taup_i * (p[OR_TP] - p[OR_TF]));
  p[OR_TPRHS_L] = 3.0 * c->Lv / c->Cpp / r * p[OR_RADRHS];
  for (int d = 0; d < 3; ++d) {
    p[OR_X + d] = xtmp[d] + dt * gs * p[OR_XRHSX + d];
    p[OR_VX + d] = vtmp[d] + dt * gs * p[OR_VRHSX + d];
  }
  p[OR_TP] = Tptmp + dt * gs * p[OR_TPRHS_S];
  p[OR_TP] = p[OR_TP] + dt * gs * p[OR_TPRHS_L];
  p[OR_RADIUS] = radiustmp + dt * gs * p[OR_RADRHS];
  if (istage == 3) p[OR_RES] = p[OR_RES] + dt;
}

void or_update_rk3(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                   double *soa, int64_t n, int istage, int it) {
  double p[NR];
  for (int64_t i = 0; i < n; ++i) {
    load_rec(soa, n, i, p);
    if (istage == 1) {
      p[OR_VXOLD] = p[OR_VX]; p[OR_VYOLD] = p[OR_VY]; p[OR_VZOLD] = p[OR_VZ];
      p[OR_TPOLD] = p[OR_TP];
      p[OR_RADIUSOLD] = p[OR_RADIUS];
    }
    interp_one(g, f, fl, p);
    double rhoa = air_rho(c, f, p);
    rk3_body(c, f, p, rhoa, istage, it);
    bcs(g, f, p);
    store_rec(soa, n, i, p);
  }
}

/* -------------------------------------------------------- BE microphysics */

/* ie_vrt_nd (particles.f90:4263-4327).  The velocity residual v_output is
 * not used by any caller and is not formed. */
void or_ie_vrt_nd(const or_consts *c, const double *p, double rhoa, const double vnext[3],
                  double tempr, double tempt, double h, double F[2]) {
  const double ri = c->radius_init;
  double taup0 = (((p[OR_M_S]) / ((2. / 3.) * c->pi2 * (ri * ri * ri)) + c->rhow) *
                  ((ri * 2) * (ri * 2))) / (18 * rhoa * c->nuf);
  double rnext = tempr * p[OR_RADIUS];
  double Tnext = tempt * p[OR_TP];
  double dnext = rnext * 2.;
  double esa = or_mod_magnus(p[OR_TF]);
  double VolP = (2. / 3.) * c->pi2 * (rnext * rnext * rnext);
  double rhop = (p[OR_M_S] + VolP * c->rhow) / VolP;
  double d0 = p[OR_UFX] - vnext[0], d1 = p[OR_UFY] - vnext[1], d2 = p[OR_UFZ] - vnext[2];
  double diffnorm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
  double Rep = dnext * diffnorm / c->nuf;
  double taup = (rhop * (dnext * dnext)) / (18.0 * rhoa * c->nuf);
  double qstr = (c->Mw / (c->Ru * Tnext * rhoa)) * esa *
                exp(((c->Lv * c->Mw / c->Ru) * ((1. / p[OR_TF]) - (1. / Tnext))) +
                    ((2. * c->Mw * c->Gam) / (c->Ru * c->rhow * rnext * Tnext)) -
                    ((p[OR_KAPPA_S] * p[OR_M_S] * c->rhow / c->rhos) / (VolP * rhop - p[OR_M_S])));
  double Shp = 2. + 0.6 * pow(Rep, 1. / 2.) * pow(c->Sc, 1. / 3.);
  double rprime = (1. / 9.) * (Shp / c->Sc) * (rhop / c->rhow) * (rnext / taup) * (p[OR_QINF] - qstr);
  rprime = rprime * (taup0 / p[OR_RADIUS]);
  double Nup = 2. + 0.6 * pow(Rep, 1. / 2.) * pow(c->Pra, 1. / 3.);
  double Tprime = -((1. / 3.) * (Nup / c->Pra) * c->CpaCpp * (rhop / c->rhow) * (1. / taup) *
                    (Tnext - p[OR_TF])) +
                  3. * c->Lv * (1. / (rnext * c->Cpp)) * rprime * (p[OR_RADIUS] / taup0);
  Tprime = Tprime * (taup0 / p[OR_TP]);
  F[0] = rnext / p[OR_RADIUS] - 1.0 - h * rprime;
  F[1] = Tnext / p[OR_TP] - 1.0 - h * Tprime;
}

/* jacob_approx_2d (4224-4252): forward difference, step 1e-12; the backward
 * evaluations of the reference are never used and are not formed. */
static void jacob_approx_2d(const or_consts *c, const double *p, double rhoa,
                            const double vnext[3], double rnext, double tnext, double h,
                            double J[2][2] /* J[row][col] */) {
  const double diff = 1e-12;
  double y[2] = {rnext, tnext}, f0[2], fx[2], xper[2];
  or_ie_vrt_nd(c, p, rhoa, vnext, rnext, tnext, h, f0);
  xper[0] = y[0]; xper[1] = y[1];
  for (int n = 0; n < 2; ++n) {
    xper[n] = xper[n] + diff;
    or_ie_vrt_nd(c, p, rhoa, vnext, xper[0], xper[1], h, fx);
    J[0][n] = (fx[0] - f0[0]) / diff;
    J[1][n] = (fx[1] - f0[1]) / diff;
    xper[n] = y[n];
  }
}

/* inverse_finder_2d (4253-4262) */
static void inverse_2d(double C[2][2], double det, double inv[2][2]) {
  double s = 1. / det;
  inv[0][0] = s * C[1][1];
  inv[1][0] = s * (-C[1][0]);
  inv[0][1] = s * (-C[0][1]);
  inv[1][1] = s * C[0][0];
}

/* gauss_newton_2d (4071-4142) */
void or_gauss_newton_2d(const or_consts *c, const double *p, const double vnext[3], double h,
                        double rhoa, const double vec1[2], double vec2[2], int *flag, int *iters) {
  int iterations = 0;
  const double error = 1.0e-8;
  double v1[2] = {vec1[0], vec1[1]};
  double correct[2] = {0.0, 0.0};
  const int iteration_max = 50;
  *flag = 0;
  vec2[0] = v1[0]; vec2[1] = v1[1];
  while (iterations < iteration_max) {
    iterations++;
    double fv1[2], J[2][2], fancy[2][2], inv[2][2], fJ[2][2];
    or_ie_vrt_nd(c, p, rhoa, vnext, v1[0], v1[1], h, fv1);
    jacob_approx_2d(c, p, rhoa, vnext, v1[0], v1[1], h, J);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) fancy[i][j] = J[0][i] * J[0][j] + J[1][i] * J[1][j];
    double det = fancy[0][0] * fancy[1][1] - fancy[0][1] * fancy[1][0];
    if (fabs(det) < 1.0e-10) { *flag = 1; break; }
    inverse_2d(fancy, det, inv);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) fJ[i][j] = inv[i][0] * J[j][0] + inv[i][1] * J[j][1];
    correct[0] = fJ[0][0] * fv1[0] + fJ[0][1] * fv1[1];
    correct[1] = fJ[1][0] * fv1[0] + fJ[1][1] * fv1[1];
    vec2[0] = v1[0] - correct[0];
    vec2[1] = v1[1] - correct[1];
    int counts = 0;
    double relax = 1.0;
    while ((vec2[0] < 0) || (vec2[1] < 0) || isnan(vec2[0])) {
      counts++;
      relax = relax * 0.5;
      vec2[0] = v1[0] - correct[0] * relax;
      vec2[1] = v1[1] - correct[1] * relax;
      if (counts > 10) break;
    }
    if (vec2[0] < 0.01 || vec2[1] < 0.01 || isnan(vec2[0]) || isnan(vec2[1])) { *flag = 1; break; }
    if (sqrt(correct[0] * correct[0] + correct[1] * correct[1]) < error) break;
    v1[0] = vec2[0]; v1[1] = vec2[1];
  }
  if (iterations == iteration_max) *flag = 1;
  if (isnan(vec2[0]) || vec2[0] < 0 || isnan(vec2[1]) || vec2[1] < 0) *flag = 1;
  if (iters) *iters = iterations;
}

/* LV_solver (4143-4223) */
void or_lv_solver(const or_consts *c, const double *p, const double vnext[3], double h,
                  double rhoa, const double vec1[2], double vec2[2], int *flag, int *iters) {
  int iterations = 0;
  const int iterations_max = 1000;
  const double error = 1.0e-8;
  double v1[2] = {vec1[0], vec1[1]};
  double lambda = 0.001, lup = 2.0, ldown = 2.0;
  *flag = 0;
  while (iterations < iterations_max) {
    iterations++;
    double J[2][2], fv1[2], gm[2][2], invg[2][2], gradC[2], C[2], correct[2], fv2[2], newC[2];
    jacob_approx_2d(c, p, rhoa, vnext, v1[0], v1[1], h, J);
    or_ie_vrt_nd(c, p, rhoa, vnext, v1[0], v1[1], h, fv1);
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j)
        gm[i][j] = (J[0][i] * J[0][j] + J[1][i] * J[1][j]) + lambda * (i == j ? 1.0 : 0.0);
    gradC[0] = J[0][0] * fv1[0] + J[1][0] * fv1[1];
    gradC[1] = J[0][1] * fv1[0] + J[1][1] * fv1[1];
    C[0] = 0.5 * fv1[0] * fv1[0];
    C[1] = 0.5 * fv1[1] * fv1[1];
    double det = gm[0][0] * gm[1][1] - gm[0][1] * gm[1][0];
    if (fabs(det) < 1.0e-10) { *flag = 1; break; }
    inverse_2d(gm, det, invg);
    correct[0] = invg[0][0] * gradC[0] + invg[0][1] * gradC[1];
    correct[1] = invg[1][0] * gradC[0] + invg[1][1] * gradC[1];
    vec2[0] = v1[0] - correct[0];
    vec2[1] = v1[1] - correct[1];
    if (sqrt(correct[0] * correct[0] + correct[1] * correct[1]) < error) break;
    if (vec2[0] < 0.0 || vec2[1] < 0.0 || isnan(vec2[0]) || isnan(vec2[1])) { *flag = 1; break; }
    or_ie_vrt_nd(c, p, rhoa, vnext, vec2[0], vec2[1], h, fv2);
    newC[0] = 0.5 * fv2[0] * fv2[0];
    newC[1] = 0.5 * fv2[1] * fv2[1];
    if (sqrt(newC[0] * newC[0] + newC[1] * newC[1]) < sqrt(C[0] * C[0] + C[1] * C[1])) {
      v1[0] = vec2[0]; v1[1] = vec2[1];
      lambda = lambda / ldown;
    } else {
      lambda = lambda * lup;
    }
  }
  if (iterations == iterations_max) *flag = 1;
  if (vec2[0] < 0 || vec2[1] < 0) *flag = 1;
  if (iters) *iters = iterations;
}

/* rad_solver2 (4328-4371): Koehler equilibrium radius as initial guess */
double or_rad_solver2(const or_consts *c, const double *p, double rhoa, int *mflag) {
  *mflag = 0;
  double Tf = p[OR_TF];
  double esa = or_mod_magnus(Tf);
  double L = log((c->Ru * Tf * rhoa * p[OR_QINF]) / (c->Mw * esa));
  double a = -((2 * c->Mw * c->Gam) / (c->Ru * c->rhow * Tf) / L);
  double cc = (p[OR_KAPPA_S] * p[OR_M_S]) / ((2.0 / 3.0) * c->pi2 * c->rhos) / L;
  double Q = pow(a, 2.0) / 9.0;
  double R = (2.0 * pow(a, 3.0) + 27.0 * cc) / 54.0;
  double M = pow(R, 2.0) - pow(Q, 3.0);
  double guess;
  if (M < 0) {
    double theta = acos(R / sqrt(pow(Q, 3.0)));
    guess = -(2 * sqrt(Q) * cos((theta - c->pi2) / 3.0)) - a / 3.0;
    if (guess < 0) guess = -(2 * sqrt(Q) * cos((theta + c->pi2) / 3.0)) - a / 3.0;
  } else {
    double S = -(R / fabs(R)) * pow(fabs(R) + sqrt(M), 1.0 / 3.0);
    double T = Q / S;
    guess = S + T - a / 3.0;
    if (guess < 0) {
      guess = p[OR_RADIUS];
      *mflag = 1;
    }
  }
  return guess;
}

/* The radius/temperature block of particle_update_BE (3103-3201), from
 * taup0 to qstar; p already holds the new velocity, the gathered fluid
 * state and the pre-solve radius and temperature. */
static void be_microphysics(const or_consts *c, const or_flags *f, double *p, double rhoa,
                            double Rep, double rhop, double taup_i, or_counters *cnt) {
  const double dt = c->dt;
  const double ri = c->radius_init;
  double taup0 = (((p[OR_M_S]) / ((2. / 3.) * c->pi2 * (ri * ri * ri)) + c->rhow) *
                  ((ri * 2) * (ri * 2))) / (18 * rhoa * c->nuf);
  double dt_taup0 = dt / taup0;
  double vnext[3] = {p[OR_VX], p[OR_VY], p[OR_VZ]};
  if (f->ievap == 1 && p[OR_QINF] > 0.0) {
    int mflag, flag, it1 = 0, it2 = 0;
    double rt_start[2], rt_zeroes[2];
    double guess = or_rad_solver2(c, p, rhoa, &mflag);
    if (mflag == 0) {
      rt_start[0] = guess / p[OR_RADIUS];
      rt_start[1] = p[OR_TF] / p[OR_TP];
    } else {
      rt_start[0] = 1.0;
      rt_start[1] = 1.0;
    }
    or_gauss_newton_2d(c, p, vnext, dt_taup0, rhoa, rt_start, rt_zeroes, &flag, &it1);
    if (cnt) { cnt->gn_iters += it1; cnt->solves++; }
    if (flag == 1) {
      if (cnt) cnt->num100++;
      or_lv_solver(c, p, vnext, dt_taup0, rhoa, rt_start, rt_zeroes, &flag, &it2);
      if (cnt) cnt->lm_iters += it2;
    }
    if ((flag == 1) || isnan(rt_zeroes[0]) || (rt_zeroes[0] * p[OR_RADIUS] < 0) ||
        isnan(rt_zeroes[1]) || (rt_zeroes[1] < 0) || (rt_zeroes[0] * p[OR_RADIUS] > 1.0e-2)) {
      if (cnt) cnt->numimpos++;
      rt_zeroes[0] = 1.0;
      rt_zeroes[1] = p[OR_TF] / p[OR_TP];
    }
    if (p[OR_RADIUS] > p[OR_RC] && p[OR_RADIUS] * rt_zeroes[0] < p[OR_RC]) {
      if (cnt) cnt->denum++;
    } else if (p[OR_RADIUS] < p[OR_RC] && p[OR_RADIUS] * rt_zeroes[0] > p[OR_RC]) {
      if (cnt) cnt->actnum++;
      p[OR_NUMACT] = p[OR_NUMACT] + 1.0;
      p[OR_ACTRES] = 0.0;
    }
    p[OR_RADIUS] = rt_zeroes[0] * p[OR_RADIUS];
    p[OR_TP] = rt_zeroes[1] * p[OR_TP];
  } else {
    double Nup = 2.0 + 0.6 * pow(Rep, 1.0 / 2.0) * pow(c->Pra, 1.0 / 3.0);
    double tmp_coeff = Nup / 3.0 / c->Pra * c->CpaCpp * rhop / c->rhow * taup_i;
    p[OR_TP] = (p[OR_TP] + tmp_coeff * dt * p[OR_TF]) / (1 + dt * tmp_coeff);
  }
  double r = p[OR_RADIUS];
  double Volp = c->pi2 * 2.0 / 3.0 * (r * r * r);
  rhop = (p[OR_M_S] + Volp * c->rhow) / Volp;
  double einf = or_mod_magnus(p[OR_TF]);
  p[OR_QSTAR] = (c->Mw / (c->Ru * p[OR_TP] * rhoa)) * einf *
                exp(((c->Lv * c->Mw / c->Ru) * ((1. / p[OR_TF]) - (1. / p[OR_TP]))) +
                    ((2. * c->Mw * c->Gam) / (c->Ru * c->rhow * r * p[OR_TP])) -
                    ((p[OR_KAPPA_S] * p[OR_M_S] * c->rhow / c->rhos) / (Volp * rhop - p[OR_M_S])));
  p[OR_RES] = p[OR_RES] + dt;
  p[OR_ACTRES] = p[OR_ACTRES] + dt;
}

/* The flux statistics of particle_update_BE (3208-3238): after the advance
 * and the radius/temperature update, before the walls, a droplet that moved
 * up from below w level z(iz) to above it adds mult, its mass rhop*Volp*mult
 * and its enthalpy rhop*Volp*Cpp*Tp*mult to level iz, for iz = fluxloci ..
 * fluxloc-1; moving down it subtracts them for iz = fluxloc .. fluxloci-1.
 * fluxloc(i) = minloc(z, z > xp(3)) - 1 at the new (initial) height, with
 * fluxloc = nnz+1 above zl and 0 below the ground; zl is the domain height
 * z(nnz).  flux = pflux(0:nnz) | pmassflux(0:nnz) | penegflux(0:nnz); levels
 * outside 0..nnz (the reference's pflux(-1) for a droplet at or above
 * z(nnz+1) before the step) are not counted. */
static void be_flux(const or_grid *g, const or_consts *c, const double *p, double xp3i, int64_t mult,
                    double *flux) {
  const int len = g->nnz + 2, nlev = g->nnz + 1;
  const double zl = g->z[g->nnz];
  int fluxloc, fluxloci;
  if (p[OR_Z] > zl) {
    fluxloc = g->nnz + 1;
    fluxloci = minloc_gt(g->z, len, xp3i) - 1;
  } else if (p[OR_Z] < 0.0) {
    fluxloci = minloc_gt(g->z, len, xp3i) - 1;
    fluxloc = 0;
  } else {
    fluxloc = minloc_gt(g->z, len, p[OR_Z]) - 1;
    fluxloci = minloc_gt(g->z, len, xp3i) - 1;
  }
  const double r = p[OR_RADIUS];
  const double Volp = c->pi2 * 2.0 / 3.0 * (r * r * r);
  const double rhop = (p[OR_M_S] + Volp * c->rhow) / Volp;
  const double m = (double)mult;
  if (xp3i < p[OR_Z]) {
    for (int iz = fluxloci; iz <= fluxloc - 1; ++iz) {
      if (iz < 0 || iz >= nlev) continue;
      flux[iz] = flux[iz] + m;
      flux[nlev + iz] = flux[nlev + iz] + rhop * Volp * m;
      flux[2 * nlev + iz] = flux[2 * nlev + iz] + rhop * Volp * c->Cpp * p[OR_TP] * m;
    }
  } else if (xp3i > p[OR_Z]) {
    for (int iz = fluxloc; iz <= fluxloci - 1; ++iz) {
      if (iz < 0 || iz >= nlev) continue;
      flux[iz] = flux[iz] - m;
      flux[nlev + iz] = flux[nlev + iz] - rhop * Volp * m;
      flux[2 * nlev + iz] = flux[2 * nlev + iz] - rhop * Volp * c->Cpp * p[OR_TP] * m;
    }
  }
}

/* particle_update_BE (3000-3280) on droplets i0 <= i < i1 of a SoA array
 * with n droplets per column; the flux statistics (3208-3238) are added to
 * flux when it is not NULL (mult is then the droplets' multiplicity) */
static void update_be_range(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                            double *soa, int64_t n, int64_t i0, int64_t i1, int it, or_counters *cnt,
                            const int64_t *mult, double *flux) {
  const double dt = c->dt;
  double p[NR];
  for (int64_t i = i0; i < i1; ++i) {
    load_rec(soa, n, i, p);
    const double xp3i = p[OR_Z];
    p[OR_VXOLD] = p[OR_VX]; p[OR_VYOLD] = p[OR_VY]; p[OR_VZOLD] = p[OR_VZ];
    p[OR_TPOLD] = p[OR_TP];
    p[OR_RADIUSOLD] = p[OR_RADIUS];
    interp_one(g, f, fl, p);
    if (it <= 1) { p[OR_VX] = p[OR_UFX]; p[OR_VY] = p[OR_UFY]; p[OR_VZ] = p[OR_UFZ]; }
    double rhoa = air_rho(c, f, p);
    double r = p[OR_RADIUS];
    double d0 = p[OR_VX] - p[OR_UFX], d1 = p[OR_VY] - p[OR_UFY], d2 = p[OR_VZ] - p[OR_UFZ];
    double diffnorm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    double Volp = c->pi2 * 2.0 / 3.0 * (r * r * r);
    double rhop = (p[OR_M_S] + Volp * c->rhow) / Volp;
    double taup_i = 18.0 * rhoa * c->nuf / rhop / ((2.0 * r) * (2.0 * r));
    double Rep = 2.0 * r * diffnorm / c->nuf;
    double corrfac = 1.0;
    for (int d = 0; d < 3; ++d) p[OR_X + d] = p[OR_X + d] + dt * p[OR_VX + d];
    for (int d = 0; d < 3; ++d)
      p[OR_VX + d] = (p[OR_VX + d] + taup_i * dt * corrfac * p[OR_UFX + d] + dt * c->part_grav[d]) /
                     (1 + dt * corrfac * taup_i);
    be_microphysics(c, f, p, rhoa, Rep, rhop, taup_i, cnt);
    if (flux) be_flux(g, c, p, xp3i, mult[i], flux);
    bcs(g, f, p);
    store_rec(soa, n, i, p);
  }
}

void or_update_be(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                  double *soa, int64_t n, int it, or_counters *cnt) {
  update_be_range(g, c, f, fl, soa, n, 0, n, it, cnt, NULL, NULL);
}

void or_update_be_flux(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                       double *soa, const int64_t *mult, int64_t n, int it, or_counters *cnt, double *flux) {
  const int nlev = g->nnz + 1;
  for (int k = 0; k < 3 * nlev; ++k) flux[k] = 0.0; /* 3030-3032 */
  update_be_range(g, c, f, fl, soa, n, 0, n, it, cnt, mult, flux);
}

/* The same over the host's cores (OpenMP; a CPU timing baseline, not a
 * parity oracle): droplets are independent, the counters are summed per
 * thread. */
void or_update_be_par(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                      double *soa, int64_t n, int it, or_counters *cnt) {
#pragma omp parallel
  {
    or_counters local;
    memset(&local, 0, sizeof(local));
    const int64_t nt = omp_get_num_threads(), t = omp_get_thread_num();
    update_be_range(g, c, f, fl, soa, n, n * t / nt, n * (t + 1) / nt, it, &local, NULL, NULL);
#pragma omp critical
    {
      int64_t *a = (int64_t *)cnt;
      const int64_t *b = (const int64_t *)&local;
      for (size_t k = 0; k < sizeof(or_counters) / sizeof(int64_t); ++k) a[k] += b[k];
    }
  }
}

/* ------------------------------------------------------------ coupling */

static inline int64_t src_off(const or_grid *g, int k, int j, int i) {
  /* (0:nnz+1, 1:nny+1, 1:nnx+1), les.F:8467-8479 with iys=mxs=1 */
  return (int64_t)k + (int64_t)(g->nnz + 2) * ((j - 1) + (int64_t)(g->nny + 1) * (i - 1));
}

/* particle_coupling_update (1151-1276).  Node indices are wrapped/clamped
 * into the array where the reference would write outside it (its ERROR1-4
 * branches, 1228-1236). */
/* src[o] = src[o] + v, as one atomic update when several threads deposit */
#define OR_ADD(dst, v)        \
  do {                        \
    double *a_ = &(dst);      \
    const double v_ = (v);    \
    if (atomic) {             \
      _Pragma("omp atomic")   \
      *a_ += v_;              \
    } else {                  \
      *a_ = *a_ + v_;         \
    }                         \
  } while (0)

static void coupling_range(const or_grid *g, const or_consts *c, const or_flags *f, const double *soa,
                           const int64_t *mult, int64_t n, int64_t i0, int64_t i1, double *src, int atomic) {
  const double dt = c->dt, dx = g->dx, dy = g->dy;
  int64_t nf = (int64_t)(g->nnz + 2) * (g->nny + 1) * (g->nnx + 1);
  double p[NR];
  for (int64_t ip = i0; ip < i1; ++ip) {
    load_rec(soa, n, ip, p);
    int32_t ci[5];
    or_cell_index(g, p[OR_X], p[OR_Y], p[OR_Z], ci);
    /* the deposit hangs off the droplet's u cell clamped into
       [1,nnx+1] x [1,nny+1] x [0,nnz], which only changes positions outside
       the domain (where the reference writes outside its arrays) */
    int ipt = clampi(ci[0], 1, g->nnx + 1), jpt = clampi(ci[1], 1, g->nny + 1), kpt = clampi(ci[2], 0, g->nnz);
    int ipn = ipt, jpn = jpt;
    double vrhs[3];
    for (int d = 0; d < 3; ++d) vrhs[d] = (p[OR_VX + d] - p[OR_VXOLD + d]) / dt;
    double r = p[OR_RADIUS];
    double radrhs = (r - p[OR_RADIUSOLD]) / dt;
    double Tprhs_L = 3.0 * c->Lv / c->Cpp / r * radrhs;
    double Tprhs_s = (p[OR_TP] - p[OR_TPOLD]) / dt - Tprhs_L;
    double rm = (double)mult[ip];
    for (int i = 0; i <= 1; ++i)
      for (int j = 0; j <= 1; ++j)
        for (int k = 0; k <= 1; ++k) {
          double xv = dx * (double)(i + ipt - 1);
          double yv = dy * (double)(j + jpt - 1);
          double zv = g->zz[k + kpt];
          double dV = dx * dy * g->dzu[kpt + 1];
          double wtx = (1.0 - fabs(p[OR_X] - xv) / dx);
          double wty = (1.0 - fabs(p[OR_Y] - yv) / dy);
          double wtz = (1.0 - fabs(p[OR_Z] - zv) / g->dzu[kpt + 1]);
          double wtt = wtx * wty * wtz;
          double rhoa;
          if (f->iexner == 1)
            rhoa = or_func_p_base(c->surf_p, c->tsfcc1, p[OR_Z], c) / c->Rd / p[OR_TF];
          else
            rhoa = c->surf_rho;
          double rhop = (p[OR_M_S] + c->pi2 * 2.0 / 3.0 * (r * r * r) * c->rhow) /
                        (c->pi2 * 2.0 / 3.0 * (r * r * r));
          double partmass = rhop * 2.0 / 3.0 * c->pi2 * (r * r * r);
          double taup_i = 18.0 * rhoa * c->nuf / rhop / ((2.0 * r) * (2.0 * r));
          int ix = ipn + i, iy = jpn + j, iz = kpt + k;
          if (ix > g->nnx + 1) ix -= g->nnx;
          if (iy > g->nny + 1) iy -= g->nny;
          int64_t o = src_off(g, iz, iy, ix);
          /* the reference's src = src - a - b as two additions of -a, -b (a + (-b) == a - b) */
          if (f->icouple == 1) {
            for (int d = 0; d < 3; ++d) {
              OR_ADD(src[d * nf + o], -(partmass / rhoa * (vrhs[d] - c->part_grav[d]) * wtt / dV * rm));
              OR_ADD(src[d * nf + o], -(c->rhow / rhoa * c->pi2 * 2 * (r * r) * radrhs *
                                         p[OR_VX + d] * wtt / dV * rm));
            }
          }
          if (f->iTcouple == 1) {
            OR_ADD(src[3 * nf + o], -((Tprhs_s * 6.0 * c->rhow / rhop / c->CpaCpp / taup_i * (c->pi2 / 2.0) * r *
                                       c->nuf) * wtt / dV * rm));
          }
          if (f->iHcouple == 1) {
            OR_ADD(src[4 * nf + o], -(c->rhow / rhoa * c->pi2 * 2 * (r * r) * radrhs * wtt / dV * rm));
            OR_ADD(src[5 * nf + o], -(c->rhow / rhoa * c->pi2 * 2 * (r * r) * radrhs * c->Cpv / c->Cpa * p[OR_TP] *
                                      wtt / dV * rm));
            OR_ADD(src[5 * nf + o], c->rhow / rhoa * c->pi2 * 2 * (r * r) * radrhs * c->Cpv / c->Cpa * p[OR_TF] *
                                        wtt / dV * rm);
          }
        }
  }
}

void or_coupling(const or_grid *g, const or_consts *c, const or_flags *f, const double *soa,
                 const int64_t *mult, int64_t n, double *src) {
  int64_t nf = (int64_t)(g->nnz + 2) * (g->nny + 1) * (g->nnx + 1);
  memset(src, 0, sizeof(double) * 6 * nf);
  coupling_range(g, c, f, soa, mult, n, 0, n, src, 0);
}

/* The same over the host's cores (OpenMP, atomic deposits; a CPU timing
 * baseline, not a parity oracle: the order of the additions varies). */
void or_coupling_par(const or_grid *g, const or_consts *c, const or_flags *f, const double *soa,
                     const int64_t *mult, int64_t n, double *src) {
  int64_t nf = (int64_t)(g->nnz + 2) * (g->nny + 1) * (g->nnx + 1);
  memset(src, 0, sizeof(double) * 6 * nf);
#pragma omp parallel
  {
    const int64_t nt = omp_get_num_threads(), t = omp_get_thread_num();
    coupling_range(g, c, f, soa, mult, n, n * t / nt, n * (t + 1) / nt, src, 1);
  }
}

int or_max_threads(void) { return omp_get_max_threads(); }

/* ------------------------------------------------ coupling exchange/apply */

/* particle_coupling_exchange (1066-1149) with tproc = bproc = rproc = lproc =
 * this rank: the sendrecv pairs become in-place adds, in the same order. */
void or_coupling_exchange(const or_grid *g, double *src, double *xsrc) {
  const int nx = g->nnx, ny = g->nny, nz2 = g->nnz + 2;
  const int64_t nf = (int64_t)nz2 * (ny + 1) * (nx + 1);
  for (int f = 0; f < 6; ++f) {
    double *s = src + f * nf;
    /* top send (1084-1097): plane mxe+1 onto plane mxs, rows iys..iye+1 */
    for (int iy = 1; iy <= ny + 1; ++iy)
      for (int k = 0; k < nz2; ++k) s[src_off(g, k, iy, 1)] = s[src_off(g, k, iy, 1)] + s[src_off(g, k, iy, nx + 1)];
    /* right send (1099-1113): row iye+1 onto row iys, columns mxs..mxe */
    for (int ix = 1; ix <= nx; ++ix)
      for (int k = 0; k < nz2; ++k) s[src_off(g, k, 1, ix)] = s[src_off(g, k, 1, ix)] + s[src_off(g, k, ny + 1, ix)];
    /* ztox_trans (1117-1146) */
    double *o = xsrc + (int64_t)f * nz2 * ny * nx;
    for (int k = 0; k < nz2; ++k)
      for (int iy = 1; iy <= ny; ++iy)
        for (int ix = 1; ix <= nx; ++ix) o[(ix - 1) + (int64_t)nx * ((iy - 1) + (int64_t)ny * k)] = s[src_off(g, k, iy, ix)];
  }
}

/* apply_particle_coupling (1278-1313) over iz = izs..ize = 1..nnz */
void or_apply_coupling(const or_grid *g, const or_consts *c, const or_flags *f, const double *xsrc,
                       double *T, double *q) {
  const int nx = g->nnx, ny = g->nny, nz = g->nnz;
  const int64_t nf = (int64_t)(nz + 2) * ny * nx;
  const double *sT = xsrc + 3 * nf, *sH = xsrc + 4 * nf, *sTE = xsrc + 5 * nf;
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        const int64_t xs = (ix - 1) + (int64_t)nx * ((iy - 1) + (int64_t)ny * iz);
        const int64_t t = (ix - 1) + (int64_t)(nx + 2) * ((iy - 1) + (int64_t)ny * iz);
        if (f->iTcouple == 1) {
          if (f->iexner)
            T[t] = T[t] + c->dt * (sT[xs] / or_exner(c->surf_p, or_func_p_base(c->surf_p, c->tsfcc1, g->zz[iz], c), c));
          else
            T[t] = T[t] + c->dt * sT[xs];
        }
        if (f->iHcouple == 1) {
          T[t] = T[t] + c->dt * sTE[xs];
          q[t] = q[t] + c->dt * sH[xs];
        }
      }
}

/* ------------------------------------------------------------ statistics */

/* calc_tauc_min (3310-3368).  A droplet whose cell lies outside
 * (1:nnx, 1:nny, 1:nnz) is not counted (the reference would index outside
 * its arrays). */
double or_calc_tauc_min(const or_grid *g, const or_consts *c, const double *soa, const int64_t *mult, int64_t n) {
  const int nx = g->nnx, ny = g->nny, nz = g->nnz;
  const int64_t nc = (int64_t)nx * ny * nz;
  double *rad = calloc((size_t)nc, sizeof(double));
  double *ncs = calloc((size_t)nc, sizeof(double));
  int32_t *num = calloc((size_t)nc, sizeof(int32_t));
  for (int64_t i = 0; i < n; ++i) {
    int32_t ci[5];
    or_cell_index(g, soa[OR_X * n + i], soa[OR_Y * n + i], soa[OR_Z * n + i], ci);
    const int ipt = ci[0], jpt = ci[1], kpt = ci[4];
    if (ipt < 1 || ipt > nx || jpt < 1 || jpt > ny || kpt < 1 || kpt > nz) continue;
    const int64_t e = (ipt - 1) + (int64_t)nx * ((jpt - 1) + (int64_t)ny * (kpt - 1));
    rad[e] = rad[e] + soa[OR_RADIUS * n + i];
    num[e] = num[e] + 1;
    ncs[e] = ncs[e] + (double)mult[i];
  }
  double tauc_min = 1.0e10;
  for (int iz = 1; iz <= nz; ++iz)
    for (int iy = 1; iy <= ny; ++iy)
      for (int ix = 1; ix <= nx; ++ix) {
        const int64_t e = (ix - 1) + (int64_t)nx * ((iy - 1) + (int64_t)ny * (iz - 1));
        if (num[e] != 0) {
          const double radavg = rad[e] / (double)num[e];
          const double Nc = ncs[e] / (g->dx * g->dy * g->dzw[iz]);
          const double denom = c->pi2 * (c->nuf / c->Sc) * 2.0 * radavg * Nc;
          const double t = 1.0 / (denom + 1.0e-10);
          if (t < tauc_min) tauc_min = t;
        }
      }
  free(rad);
  free(ncs);
  free(num);
  return tauc_min * 5.0;
}

/* particle_xy_stats (3370-3770): the loops over the droplet list. */
void or_xy_stats(const or_grid *g, const or_consts *c, const double *soa, const int64_t *mult, int64_t n,
                 double *s0, double *s1, double *lwp) {
  const int nx = g->nnx, ny = g->nny, nz2 = g->nnz + 2;
  double cnt[3] = {0, 0, 0}, sum[6] = {0, 0, 0, 0, 0, 0};
  double mx[6] = {0.0, 0.0, 0.0, -1000.0, -1000.0, -1000.0};
  double mn[6] = {1000.0, 1000.0, 1000.0, 1000.0, 1000.0, 1000.0};
  memset(s1, 0, sizeof(double) * 24 * nz2);
  memset(lwp, 0, sizeof(double) * nx * ny);
  const double pi = 4.0 * atan(1.0);
  for (int64_t i = 0; i < n; ++i) {
    double p[OR_NREAL];
    for (int k = 0; k < OR_NREAL; ++k) p[k] = soa[(int64_t)k * n + i];
    /* 0th order (3425-3468) */
    cnt[0] += 1.0;
    if (p[OR_RADIUS] > p[OR_RC]) cnt[1] += 1.0;
    else cnt[2] += 1.0;
    sum[4] = sum[4] + p[OR_RADIUS];
    sum[5] = sum[5] + pow(p[OR_RADIUS], 2);
    const double v0[6] = {p[OR_RADIUS], p[OR_TP], p[OR_QINF], p[OR_VX], p[OR_VY], p[OR_VZ]};
    for (int k = 0; k < 6; ++k) {
      mx[k] = v0[k] > mx[k] ? v0[k] : mx[k];
      mn[k] = v0[k] < mn[k] ? v0[k] : mn[k];
    }
    const double Volp = c->pi2 * 2.0 / 3.0 * pow(p[OR_RADIUS], 3);
    const double rhop0 = (p[OR_M_S] + Volp * c->rhow) / Volp;
    const double d0 = p[OR_VX] - p[OR_UFX], d1 = p[OR_VY] - p[OR_UFY], d2 = p[OR_VZ] - p[OR_UFZ];
    const double diffnorm = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
    const double Rep = 2.0 * p[OR_RADIUS] * diffnorm / c->nuf;
    sum[0] = sum[0] + Rep;
    sum[1] = sum[1] + Volp * c->rhow;
    sum[2] = sum[2] + Volp * rhop0;
    sum[3] = sum[3] + Volp;
    /* 1st order (3580-3636) */
    int32_t ci[5];
    or_cell_index(g, p[OR_X], p[OR_Y], p[OR_Z], ci);
    const int ipt = ci[0], jpt = ci[1], kpt = ci[4];
    const double m = (double)mult[i];
    const double mw = c->rhow * 4.0 / 3.0 * pi * pow(p[OR_RADIUS], 3);
    if (kpt >= 0 && kpt < nz2) {
      const double q[24] = {1.0,
                            p[OR_VX], p[OR_VY], p[OR_VZ],
                            pow(p[OR_VX], 2), pow(p[OR_VY], 2), pow(p[OR_VZ], 2),
                            p[OR_VX] * p[OR_VZ],
                            p[OR_UFX], p[OR_UFY], p[OR_UFZ],
                            pow(p[OR_UFX], 2), pow(p[OR_UFY], 2), pow(p[OR_UFZ], 2),
                            p[OR_TP], pow(p[OR_TP], 2), p[OR_TF],
                            p[OR_RADIUS], pow(p[OR_RADIUS], 2), p[OR_QINF],
                            p[OR_TP] * p[OR_VZ], m, m * mw, p[OR_QSTAR]};
      for (int k = 0; k < 24; ++k) s1[k * nz2 + kpt] = s1[k * nz2 + kpt] + q[k];
    }
    if (ipt >= 1 && ipt <= nx && jpt >= 1 && jpt <= ny) {
      const int64_t e = (ipt - 1) + (int64_t)nx * (jpt - 1);
      lwp[e] = lwp[e] + m * mw;
    }
  }
  for (int k = 0; k < 3; ++k) s0[k] = cnt[k];
  for (int k = 0; k < 6; ++k) s0[3 + k] = sum[k];
  for (int k = 0; k < 6; ++k) s0[9 + k] = mx[k];
  for (int k = 0; k < 6; ++k) s0[15 + k] = mn[k];
}

