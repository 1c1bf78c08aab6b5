/*
 * ntlp_oracle.h -- CPU FP64 restatement of NTLP's Lagrangian droplet step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the HIP path is compared
 * against; it is never linked into, loaded by, or called from the product
 * library (ntlp_amd/).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.
 *
 * Every routine follows one routine of the Fortran reference (particles.f90,
 * les.F, defs.F of RichterLab/NTLP) in its order of operations, and cites the
 * lines it restates.  The reference compiles with -r8, so every `real` there
 * is an IEEE binary64 `double` here.
 *
 * Each function is pinned by a check that does not re-type its formula:
 * known-answer vectors, closed forms, mpmath roots, conservation budgets or
 * hand cases, in tests/test_oracle_pins.py and tests/test_oracle_budgets.py.
 */
#ifndef NTLP_ORACLE_H
#define NTLP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Column order of the droplet real-valued SoA block: column c of droplet i
 * is at soa[c*n + i].  Mirrors the FP64 members of `type particle`
 * (particles.f90:76-85) minus the SFS-only members (dist, u_sub, sigm_s). */
enum {
  OR_X = 0, OR_Y, OR_Z,              /* xp(1:3)   */
  OR_VX, OR_VY, OR_VZ,               /* vp(1:3)   */
  OR_UFX, OR_UFY, OR_UFZ,            /* uf(1:3)   */
  OR_XRHSX, OR_XRHSY, OR_XRHSZ,      /* xrhs(1:3) */
  OR_VRHSX, OR_VRHSY, OR_VRHSZ,      /* vrhs(1:3) */
  OR_TP, OR_TPRHS_S, OR_TPRHS_L, OR_TF, OR_RADIUS, OR_RADRHS, OR_QINF, OR_QSTAR,
  OR_RES, OR_M_S, OR_KAPPA_S, OR_RC, OR_ACTRES, OR_NUMACT,
  OR_VXOLD, OR_VYOLD, OR_VZOLD, OR_TPOLD, OR_RADIUSOLD,
  OR_NREAL
};

typedef struct {
  int nnx, nny, nnz;
  double dx, dy, xl, yl;
  const double *z;    /* z(0:nnz+1)   w levels            (defs.F:255)  */
  const double *zz;   /* zz(0:nnz+1)  u levels            (defs.F:255)  */
  const double *dzw;  /* dzw(0:nnz+2)                     (defs.F:82)   */
  const double *dzu;  /* dzu(0:nnz+2)                     (defs.F:82)   */
} or_grid;

typedef struct {
  double dt;
  double nuf, Cpa, Cpp, CpaCpp, Pra, Sc, Rd, rhow, rhos, Mw, Ru, Gam, Lv, Cpv;
  double surf_rho, surf_p, tsfcc1, grav, radius_init, pi2;
  double part_grav[3];
  double zetas[3], gama[3];
} or_consts;

typedef struct {
  int ilin, icouple, iTcouple, iHcouple, ievap, iexner, icase;
} or_flags;

/* five ext arrays (0:nnz+1, -1:nny+3, -1:nnx+3), z fastest */
typedef struct {
  const double *u, *v, *w, *T, *q;
} or_fields;

typedef struct {
  int64_t num100, numimpos, actnum, denum;
  int64_t gn_iters, lm_iters, solves;
} or_counters;

/* les.F:10764-10795 */
double or_mod_magnus(double T);
double or_exner(double p0, double p, const or_consts *c);
double or_func_p_base(double p_surf, double T_surf, double z, const or_consts *c);

/* Reference cell indices (particles.f90:780-788, 1180-1183, 3332-3334):
 * out[0]=ipt, out[1]=jpt, out[2]=kpt (u level), out[3]=kwpt (w level, -2 form),
 * out[4]=kc (w cell, minloc(z)-1 form). Linear masked-minloc scans. */
void or_cell_index(const or_grid *g, double x, double y, double z, int32_t out[5]);
void or_cell_index_many(const or_grid *g, const double *x, const double *y,
                        const double *z, int64_t n, int32_t *out /* [n][5] */);

/* Stable counting sort of droplets by u cell (ipt,jpt,kpt), key
 * kc + (nnz+1)*((jpt-1) + (nny+1)*(ipt-1)) with the key clamp described in
 * the implementation.  counts[ncell], offsets[ncell+1], perm[n]. */
int64_t or_num_ucells(const or_grid *g);
void or_bin(const or_grid *g, const double *x, const double *y, const double *z,
            int64_t n, int32_t *keys, int32_t *counts, int32_t *offsets, int32_t *perm);
/* Per w-cell droplet count num_array(ix,iy,kc), calc_tauc_min
 * (particles.f90:3320-3339); array (nnx, nny, nnz) Fortran order. */
void or_wcell_counts(const or_grid *g, const double *x, const double *y, const double *z,
                     int64_t n, int32_t *num_array);

/* periodic ext fill of one field from the core (0:nnz+1, 1:nny, 1:nnx) */
void or_fill_ext(const or_grid *g, const double *core, double *ext);

/* interpolation of all droplets (uf, Tf, qinf columns) */
void or_interp(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
               double *soa, int64_t n);

double or_crit_radius(const or_consts *c, double m_s, double kappa_s, double Tf);
void or_crit_radius_many(const or_consts *c, const double *m_s, const double *kappa_s,
                         const double *Tp, int64_t n, double *rc);

/* one RK3 stage over all droplets incl. BCs (particle_update_rk3, 2845-2998) */
void or_update_rk3(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                   double *soa, int64_t n, int istage, int it);
/* one BE step incl. BCs (particle_update_BE, 3000-3280) */
void or_update_be(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                  double *soa, int64_t n, int it, or_counters *cnt);
/* the same step with the flux statistics of 3030-3032, 3208-3238:
 * flux[3][nnz+1] = pflux(0:nnz), pmassflux(0:nnz), penegflux(0:nnz), zeroed
 * first; mult[n] the droplets' multiplicities */
void or_update_be_flux(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                       double *soa, const int64_t *mult, int64_t n, int it, or_counters *cnt, double *flux);
/* particle_coupling_update (1151-1276): src[6][(nnz+2)*(nny+1)*(nnx+1)] */
void or_coupling(const or_grid *g, const or_consts *c, const or_flags *f, const double *soa,
                 const int64_t *mult, int64_t n, double *src);
/* OpenMP versions over the host's cores, for the CPU timing baseline only
 * (the deposit's additions are atomic, so their order varies run to run) */
void or_update_be_par(const or_grid *g, const or_consts *c, const or_flags *f, const or_fields *fl,
                      double *soa, int64_t n, int it, or_counters *cnt);
void or_coupling_par(const or_grid *g, const or_consts *c, const or_flags *f, const double *soa,
                     const int64_t *mult, int64_t n, double *src);
int or_max_threads(void);

/* particle_coupling_exchange (1066-1149) on one periodic domain: the +1
 * ghost plane x = nnx+1 is added onto x = 1 (all y incl. nny+1), then the
 * ghost row y = nny+1 onto y = 1 (x = 1..nnx), then ztox_trans writes the
 * x-pencil arrays xsrc[6][(nnz+2)][nny][nnx] (partsrc(1:nnx,1:nny,0:nnz+1,1:3),
 * partTsrc, partHsrc, partTEsrc).  src is modified (folded) in place. */
void or_coupling_exchange(const or_grid *g, double *src, double *xsrc);
/* apply_particle_coupling (1278-1313): t(1:nnx,1:nny,1,1:nnz) and
 * t(...,2,...) += dt * sources, on arrays in x-pencil layout
 * (1:nnx+2, 1:nny, 0:nnz+1). */
void or_apply_coupling(const or_grid *g, const or_consts *c, const or_flags *f, const double *xsrc,
                       double *T, double *q);

/* calc_tauc_min (3310-3368) on one domain: 5 * the smallest condensation
 * time scale over the w cells (ipt, jpt, kc) holding droplets. */
double or_calc_tauc_min(const or_grid *g, const or_consts *c, const double *soa, const int64_t *mult, int64_t n);
/* particle_xy_stats (3370-3770) on one domain, before the per-level
 * divisions: s0[21] = numpart, numdrop, numaerosol, then the sums of Rep,
 * water mass, droplet mass, volume, radius, radius^2, then the maxima of
 * radius, Tp, qinf, vp(1:3) and the minima of the same (with the reference's
 * starting values); s1[24][nnz+2] = statsvec1(iz, 1..24) for iz = 0..nnz+1;
 * lwp[nny][nnx] = LWP(ipt, jpt). */
void or_xy_stats(const or_grid *g, const or_consts *c, const double *soa, const int64_t *mult, int64_t n,
                 double *s0, double *s1, double *lwp);

/* single-droplet pieces exposed for pinning tests */
void or_ie_vrt_nd(const or_consts *c, const double *p /* one droplet record, OR_* order */,
                  double rhoa, const double vnext[3], double tempr, double tempt, double h,
                  double F[2]);
double or_rad_solver2(const or_consts *c, const double *p, double rhoa, int *mflag);
void or_gauss_newton_2d(const or_consts *c, const double *p, const double vnext[3], double h,
                        double rhoa, const double vec1[2], double vec2[2], int *flag, int *iters);
void or_lv_solver(const or_consts *c, const double *p, const double vnext[3], double h,
                  double rhoa, const double vec1[2], double vec2[2], int *flag, int *iters);

#ifdef __cplusplus
}
#endif
#endif
