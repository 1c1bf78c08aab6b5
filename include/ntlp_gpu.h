/*
 * ntlp_gpu.h -- C ABI of the MI355X droplet step (libntlp_gpu.so).
 *
 * Drop-in for the Lagrangian droplet path of NTLP (particles.f90).  The
 * Fortran host binds these with ISO_C_BINDING (fortran/ntlp_gpu_mod.f90); the
 * Python test and bench harness binds them with ctypes (ntlp_amd/__init__.py).
 *
 * Conventions
 *  - Every function returns NTLP_OK (0) or a negative NTLP_E_* code; no
 *    exception or abort crosses the ABI.  ntlp_gpu_last_error() describes the
 *    last failure.
 *  - Arrays use Fortran column-major index math and the reference's own index
 *    ranges (documented per argument).  Host pointers are borrowed for the
 *    duration of the call.
 *  - One process-wide context on the current HIP device, like the module
 *    globals of particles.f90.  Calls from several host threads must be
 *    serialised by the caller, except ntlp_serve_submit/poll/wait/stats,
 *    which any thread may call while the serving loop is open.
 *  - Droplets keep the order in which they were put: ntlp_gpu_particles_get
 *    returns them in that order, whatever order the device keeps internally.
 *  - A population of no droplets (put with n = 0, or nothing put yet) is
 *    valid, like a rank whose particle list is empty: the updates return at
 *    once, the coupling sources, counters and flux statistics are zero, and
 *    calc_tauc_min and the minima/maxima of particle_xy_stats keep the
 *    reference's start values.
 *  - ntlp_gpu_init checks all its arguments before it touches the device, so
 *    a rejected call leaves an existing context as it was.
 */
#ifndef NTLP_GPU_H
#define NTLP_GPU_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NTLP_OK 0
#define NTLP_E_INVALID (-1)     /* bad argument or inconsistent sizes       */
#define NTLP_E_NOT_INIT (-2)    /* ntlp_gpu_init has not been called         */
#define NTLP_E_HIP (-3)         /* a HIP runtime call failed                 */
#define NTLP_E_NOMEM (-4)       /* device allocation failed                  */
#define NTLP_E_RANGE (-5)       /* size beyond a documented maximum          */
#define NTLP_E_UNSUPPORTED (-6) /* flag combination the library does not run */
#define NTLP_E_STATE (-7)       /* call out of order (no fields, a failed step) */

/* Column order of the droplet real SoA block: column c of droplet i lives at
 * soa[c*n + i] (a Fortran array real(8) :: soa(n, NTLP_NREAL)).  These are
 * the FP64 members of `type particle` (particles.f90:76-85) except the
 * stochastic-subgrid members dist, u_sub and sigm_s. */
enum ntlp_column {
  NTLP_X = 0, NTLP_Y, NTLP_Z,
  NTLP_VX, NTLP_VY, NTLP_VZ,
  NTLP_UFX, NTLP_UFY, NTLP_UFZ,
  NTLP_XRHSX, NTLP_XRHSY, NTLP_XRHSZ,
  NTLP_VRHSX, NTLP_VRHSY, NTLP_VRHSZ,
  NTLP_TP, NTLP_TPRHS_S, NTLP_TPRHS_L, NTLP_TF, NTLP_RADIUS, NTLP_RADRHS, NTLP_QINF, NTLP_QSTAR,
  NTLP_RES, NTLP_M_S, NTLP_KAPPA_S, NTLP_RC, NTLP_ACTRES, NTLP_NUMACT,
  NTLP_VXOLD, NTLP_VYOLD, NTLP_VZOLD, NTLP_TPOLD, NTLP_RADIUSOLD,
  NTLP_NREAL
};

/* Grid metadata: nnx, nny, nnz (pars, defs.F:51-53); dx=xl/nnx, dy=yl/nny
 * (les.F:5488-5489); z, zz are z(0:nnz+1), zz(0:nnz+1) and dzw, dzu are
 * dzw(0:nnz+2), dzu(0:nnz+2) as get_dz builds them (les.F:5836-5874). */
typedef struct ntlp_grid {
  int32_t nnx, nny, nnz, reserved;
  double dx, dy, xl, yl;
  const double *z, *zz, *dzw, *dzu;
} ntlp_grid;

/* Physical constants (namelist &constants, defs.F:43-46, particles.f90:42)
 * and the low-storage RK3 coefficients zetas/gama (les.F:856-861). */
typedef struct ntlp_consts {
  double dt;
  double nuf, Cpa, Cpp, CpaCpp, Pra, Sc, Rd, rhow, rhos, Mw, Ru, Gam, Lv, Cpv;
  double surf_rho, surf_p, tsfcc1, grav, radius_init, pi2;
  double part_grav[3];
  double zetas[3], gama[3];
} ntlp_consts;

/* Hot-path flags of namelist &flags (defs.F:8-14).  ilin=1 trilinear,
 * otherwise 6th-order Lagrange; icase must be 0 (reflecting walls). */
typedef struct ntlp_flags {
  int32_t ilin, icouple, iTcouple, iHcouple, ievap, iexner, icase, reserved;
} ntlp_flags;

/* Newton bookkeeping of particle_update_BE (particles.f90:27-30, 3035-3039)
 * plus iteration totals.  gn_hist[k] counts solves whose Gauss-Newton loop
 * ran k iterations (k = 1..50); gn_hist[0] counts solves that fell back to
 * Levenberg-Marquardt. */
#define NTLP_GN_HIST 64
typedef struct ntlp_counters {
  int64_t num100, numimpos, actnum, denum;
  int64_t gn_iters, lm_iters, solves, steps;
  int64_t gn_hist[NTLP_GN_HIST];
} ntlp_counters;

/* Field layouts accepted by ntlp_gpu_set_fields. */
#define NTLP_LAYOUT_EXT 0  /* uext(0:nnz+1, -1:nny+3, -1:nnx+3): fill_ext's output */
#define NTLP_LAYOUT_CORE 1 /* (0:nnz+1, 1:nny, 1:nnx) z fastest; the device fills
                              the periodic x-y halo (fill_ext on one domain)     */
#define NTLP_LAYOUT_XPENCIL 2 /* u(1:nnx+2, 1:nny, 0:nnz+1) x fastest, the flow
                                 solver's own arrays (les.F:8394-8397); the device
                                 transposes to z fastest and fills the halo      */

/* Step modes and precisions of ntlp_gpu_step. */
#define NTLP_MODE_RK3 1        /* ipart_method=1: three particle_update_rk3 stages */
#define NTLP_MODE_BE 2         /* ipart_method=2: particle_update_BE               */
#define NTLP_PREC_FP64 0       /* FP64 throughout, reference Newton (FD Jacobian)   */
#define NTLP_PREC_MIXED 1      /* FP32 gather weights, drag and explicit stages; FP64
                                  positions, Newton solve (analytic Jacobian) and
                                  coupling accumulation                            */
#define NTLP_PREC_FP64_AJ 2    /* FP64 throughout with the analytic Jacobian         */

/* ---------------------------------------------------------------- context */

/* Replaces the grid/constant set-up of particle_setup (particles.f90:1829-1986)
 * on the current HIP device.  Copies grid metadata; may be called again to
 * re-initialise (all device state is released first). */
int ntlp_gpu_init(const ntlp_grid *grid, const ntlp_consts *consts, const ntlp_flags *flags);
/* Releases every device buffer of the context. */
int ntlp_gpu_finalize(void);
/* Text of the last error (thread-local, never NULL). */
const char *ntlp_gpu_last_error(void);
/* Changes dt (the reference recomputes dt every step in get_dt, les.F:621-709). */
int ntlp_gpu_set_dt(double dt);

/* ----------------------------------------------------------------- fields */

/* Replaces fill_ext (particles.f90:91-328): u, v, w, T (t(:,:,1,:)) and
 * q (t(:,:,2,:)) as host FP64 arrays in `layout`.  Copied to the device and
 * made resident; the ext halo is filled on the device. */
int ntlp_gpu_set_fields(const double *u, const double *v, const double *w, const double *T,
                        const double *q, int32_t layout);

/* ---------------------------------------------------------------- droplets */

/* Packs droplets in (the linked list of `type particle`, particles.f90:76-87).
 * real_soa is real(8) :: soa(n, NTLP_NREAL); mult, pidx, procidx per droplet.
 * Replaces any droplets already on the device. */
int ntlp_gpu_particles_put(const double *real_soa, const int64_t *mult, const int32_t *pidx,
                           const int32_t *procidx, int64_t n);
/* Unpacks droplets in put order.  Any output pointer may be NULL. */
int ntlp_gpu_particles_get(double *real_soa, int64_t *mult, int32_t *pidx, int32_t *procidx,
                           int64_t n);
/* Restart records (save_particles / read_part_res, particles.f90:1988-2146):
 * 352-byte records of 4 int32 (pidx, procidx, nbr_pidx, nbr_procidx), 39 real(8)
 * in `type particle` order, and 3 int64 (mult + two pointer slots). */
#define NTLP_RECORD_BYTES 352
int ntlp_gpu_particles_put_records(const void *records, int64_t n);
int ntlp_gpu_particles_get_records(void *records, int64_t n);
int64_t ntlp_gpu_particles_count(void);
/* One column (enum ntlp_column) of every droplet, in put order. */
int ntlp_gpu_particles_get_column(int32_t column, double *out, int64_t n);

/* A droplet ensemble drawn on the device.  It has no counterpart in the
 * reference, whose particle_init/create_particle (particles.f90:2226-2300,
 * 2440-2520) draw droplets one by one on the host with ran2/gasdev.  Droplet i
 * takes its random numbers from the counter-based generator Philox4x32-10
 * keyed by `seed` at counter first + i, so an ensemble is the same however it
 * is split.  Position: uniform over [0,xl) x [0,yl), height r + u (z(nnz) - 2r).
 * Radius: radius_kind with radius_p, or with probability frac2 radius2_kind
 * with radius2_p; NTLP_DIST_LOGNORMAL: p = (median, sigma_ln);
 * NTLP_DIST_LOGUNIFORM: p = (lo, hi).  Salt: NTLP_SALT_FIXED m_s = salt_p[0];
 * NTLP_SALT_DRY_LOGNORMAL: dry radius lognormal (median salt_p[0], sigma_ln
 * salt_p[1]), m_s = rhos (4/3) pi r_d^3 (particles.f90:2498-2499);
 * NTLP_SALT_CHOICE3: one of salt_p[0..2] with equal probability.  Then, as
 * create_particle does: vp = uf(xp), Tp = Tf(xp), rc = crit_radius; res,
 * actres, numact and the other record members zero; pidx = first + i + 1. */
#define NTLP_DIST_LOGNORMAL 0
#define NTLP_DIST_LOGUNIFORM 1
#define NTLP_SALT_FIXED 0
#define NTLP_SALT_DRY_LOGNORMAL 1
#define NTLP_SALT_CHOICE3 2
typedef struct ntlp_ensemble {
  uint64_t seed;
  int32_t radius_kind, radius2_kind;
  double radius_p[2], radius2_p[2], frac2;
  int32_t salt_kind, reserved;
  double salt_p[3];
  double kappa_s;
  int64_t mult;
  int64_t first;
} ntlp_ensemble;
/* Replaces the droplets on the device by counts[0] droplets of ens[0], then
 * counts[1] of ens[1], ... (nens <= NTLP_MAX_ENSEMBLES; the fields must be
 * set). */
#define NTLP_MAX_ENSEMBLES 65536
int ntlp_gpu_particles_generate(const ntlp_ensemble *ens, const int64_t *counts, int32_t nens);
/* save_particles / read_part_res (particles.f90:1988-2146) for one rank: the
 * file is an int32 droplet count followed by one 352-byte record per droplet.
 * Reading keeps the records with 0 < x < xl and 0 < y < yl (the reference's
 * strict bounds test) and replaces the droplets on the device; *kept (may be
 * NULL) receives their number. */
int ntlp_gpu_save_particles(const char *path);
int ntlp_gpu_read_part_res(const char *path, int64_t *kept);

/* ------------------------------------------------- reference entry points */

/* uf_interp_lin / uf_interp (particles.f90:554-847) for every droplet: fills
 * uf, Tf and qinf.  set_vp / set_tp then copy uf to vp and Tf to Tp, the
 * it<=1 initialisation of particle_update_rk3 (2907-2910). */
int ntlp_gpu_interp(int32_t set_vp, int32_t set_tp);
/* create_particle's critical radius for every droplet: rc = crit_radius(m_s,
 * kappa_s, Tp) over setup_radval's 1024 radii (particles.f90:2264-2266,
 * 5294-5327); rc = 0 when ievap /= 1. */
int ntlp_gpu_crit_radius(void);
/* particle_update_rk3(istage) (particles.f90:2845-2998) including its wall
 * and periodic boundary conditions; `it` is the time-step counter. */
int ntlp_gpu_particle_update_rk3(int32_t istage, int32_t it);
/* particle_update_BE (particles.f90:3000-3280) including boundary conditions. */
int ntlp_gpu_particle_update_be(int32_t it);
/* The same two entry points with an explicit NTLP_PREC_* arithmetic. */
int ntlp_gpu_particle_update_rk3_prec(int32_t istage, int32_t it, int32_t precision);
int ntlp_gpu_particle_update_be_prec(int32_t it, int32_t precision);
/* particle_coupling_update (particles.f90:1151-1276).  Sources stay resident;
 * when src_out is not NULL it receives partsrc_t(0:nnz+1,1:nny+1,1:nnx+1,1:3),
 * partTsrc_t, partHsrc_t, partTEsrc_t as 6 consecutive arrays.  The method:
 * deterministic=0 sums each 8x8x8 node box in LDS and stores it (the default,
 * used by ntlp_gpu_step); 1 sums each node's droplets in a fixed order (same
 * bits on every run); 2 sums runs of equal cells inside each wavefront, then
 * adds the run sums with FP64 global atomics.  Other values: NTLP_E_INVALID. */
int ntlp_gpu_particle_coupling_update(double *src_out, int32_t deterministic);

/* Copies the resident coupling sources (same layout as src_out above) as the
 * last coupling update or fused step left them. */
int ntlp_gpu_get_sources(double *src_out);
/* particle_coupling_exchange (particles.f90:1066-1149) on one periodic domain:
 * folds the +1 ghost plane of the resident sources (x = nnx+1 onto x = 1, then
 * y = nny+1 onto y = 1, the reference's order of additions) and transposes
 * them to the flow solver's layout (ztox_trans).  Each non-NULL pointer
 * receives partsrc(1:nnx,1:nny,0:nnz+1,1:3), partTsrc, partHsrc or
 * partTEsrc(1:nnx,1:nny,0:nnz+1).  The resident unfolded sources are kept. */
int ntlp_gpu_particle_coupling_exchange(double *partsrc, double *partTsrc, double *partHsrc, double *partTEsrc);
/* apply_particle_coupling (particles.f90:1278-1313): T and q (t(:,:,1,:) and
 * t(:,:,2,:)) in NTLP_LAYOUT_XPENCIL (1:nnx+2, 1:nny, 0:nnz+1) or
 * NTLP_LAYOUT_CORE get dt times the exchanged sources added at levels
 * 1..nnz, in place; the resident T and q that the next step gathers are
 * replaced by the results. */
int ntlp_gpu_apply_particle_coupling(double *T, double *q, int32_t layout);

/* ------------------------------------------------------------- binning */

/* Cell indices of every droplet (put order) as int32 [n][4] = (ipt, jpt, kpt,
 * kwpt) with the reference definitions (particles.f90:780-788), the u-cell
 * counts (nnx+1)(nny+1)(nnz+1) and exclusive offsets of the stable counting
 * sort by u cell, and the sorted permutation (put-order indices).  Any output
 * may be NULL. */
int ntlp_gpu_bin(int32_t *cellidx, int32_t *counts, int32_t *offsets, int32_t *perm);
/* calc_tauc_min's per-w-cell count num_array(nnx,nny,nnz) (particles.f90:3320-3339). */
int ntlp_gpu_wcell_counts(int32_t *num_array);
int64_t ntlp_gpu_num_ucells(void);

/* ------------------------------------------------------------- statistics */

/* calc_tauc_min (particles.f90:3310-3368) on one domain: 5 times the smallest
 * condensation time scale 1/(2 pi (nuf/Sc) 2 r_avg N_c + 1e-10) over the w
 * cells holding droplets (1e10 when none does). */
int ntlp_gpu_calc_tauc_min(double *tauc_min);
/* particle_xy_stats (particles.f90:3370-3770) on one domain, before its
 * divisions by counts:
 *  s0[21]: numpart, numdrop, numaerosol; sums of Rep, water mass, droplet
 *          mass, volume, radius, radius^2; maxima of radius, Tp, qinf,
 *          vp(1:3); minima of the same (the reference's starting values
 *          1000/0/-1000 included);
 *  s1[24*(nnz+2)]: statsvec1(iz, 1..24), quantity-major, iz = 0..nnz+1
 *          (partcount, vp1..3, vp1..3^2, vp1*vp3, uf1..3, uf1..3^2, Tp, Tp^2,
 *          Tf, radius, radius^2, qinf, Tp*vp3, mult, mult*water mass, qstar);
 *  lwp[nnx*nny]: LWP(ipt, jpt), x fastest.
 * Any output may be NULL.  num100, numimpos, actnum and denum come from
 * ntlp_gpu_counters. */
int ntlp_gpu_particle_xy_stats(double *s0, double *s1, double *lwp);
/* ntlp_gpu_particle_xy_stats for callers that cannot allocate device memory
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  It has no counterpart in the reference.  s0[21], s1[24*(nnz+2)]
 * and lwp[nnx*nny] are DEVICE pointers, all three required; the per-workgroup
 * partial rows live in a static device workspace of 12.6 MB per GPU (held
 * until the process ends; its address is looked up by ntlp_gpu_init on each
 * device, so call that first with the device current).  Sizes the workspace
 * cannot hold return NTLP_E_RANGE; with nnz + 2 <= 1024 every grid fits.  The
 * call only enqueues work on the library stream: the outputs are ready after
 * ntlp_gpu_sync().  Every call on one GPU shares the workspace, so calls that
 * overlap on one GPU (several host threads or processes) race.  Results equal
 * ntlp_gpu_particle_xy_stats' except for the summation order of lwp (global
 * atomics). */
int ntlp_gpu_particle_xy_stats_noalloc(double *s0, double *s1, double *lwp);

/* The flux statistics of the last particle_update_BE step (particles.f90:
 * 3030-3032, 3208-3238): pflux(0:nnz), pmassflux(0:nnz), penegflux(0:nnz),
 * the droplets (times mult), droplet mass and droplet enthalpy that crossed
 * each w level z(iz) upward minus those that crossed it downward, counted
 * between the advance and the walls, with zl = z(nnz).  Each output has
 * nnz+1 values and may be NULL.  NTLP_E_STATE before the first BE step. */
int ntlp_gpu_be_flux(double *pflux, double *pmassflux, double *penegflux);

/* -------------------------------------------------------- fused stepping */

/* nsteps whole droplet steps against the resident fields (which stay fixed),
 * each: gather + advance (mode) + boundary conditions + coupling scatter +
 * counting sort for the next step.  Step k uses it = it0 + k.  The result is
 * what the same number of single-step calls would leave. */
int ntlp_gpu_step(int32_t mode, int32_t precision, int32_t nsteps, int32_t it0);

/* -------------------------------------------------- counters and timing */

/* Sums of the Newton bookkeeping over every step since the last reset. */
int ntlp_gpu_counters(ntlp_counters *out, int32_t reset);
/* Per-step rows of the same counters, oldest first, for the last steps (up to
 * max_rows, at most 4095): row r holds NTLP_STEP_ROW int64 values num100,
 * numimpos, actnum, denum, gn_iters, lm_iters, solves, steps, gn_hist[64].
 * Returns the number of rows written. */
#define NTLP_STEP_ROW 72
int ntlp_gpu_step_counters(int64_t *rows, int32_t max_rows);
/* Per-kernel timing with HIP events on the library's stream.  enable=1 starts
 * recording (and clears totals); ntlp_gpu_kernel_times fills up to `cap`
 * entries of name (32 bytes each), launches and total milliseconds. */
int ntlp_gpu_timing(int32_t enable);
int ntlp_gpu_kernel_times(char *names, int64_t *launches, double *total_ms, int32_t cap);
/* Measurement helper with no counterpart in the reference: the achievable HBM
 * bandwidth of a plain device copy of `bytes` bytes (read + write, 16-byte
 * vector accesses) on the library's stream, the median of `reps` timed copies,
 * in GB/s of bytes read plus bytes written.  Allocates and frees its two
 * buffers. */
int ntlp_gpu_copy_bandwidth(int64_t bytes, int32_t reps, double *gbps);
/* Options of the fused step and the serving loop; they have no counterpart
 * in the reference.  NTLP_OPT_STAGED_GATHER: 1 stages each workgroup's
 * trilinear field nodes (the corner columns of its droplets' cells) in LDS
 * before the gathers, 0 (the default) gathers straight through the caches;
 * both give the same bits, and the direct gather measured faster on MI355X.
 * Not while a serving loop is open. */
#define NTLP_OPT_STAGED_GATHER 1
int ntlp_gpu_set_option(int32_t option, int64_t value);
/* Waits for all work of the context. */
int ntlp_gpu_sync(void);

/* ------------------------------------------------------- serving loop */

/* A continuous-batching serving loop over the resident fields; it has no
 * counterpart in the reference (NTLP advances one droplet population per
 * run).  A request is a droplet ensemble that is created the way
 * create_particle does it (vp = uf, Tp = Tf, rc = crit_radius) and advanced
 * `steps` steps of the configured mode -- backward Euler (particle_update_BE,
 * particles.f90:3000-3280) or the three RK3 stages (particle_update_rk3,
 * particles.f90:2845-2998), without the coupling scatter -- against the
 * fields of the last ntlp_gpu_set_fields.  Requests are admitted FIFO into a device pool whose
 * capacity is hbm_budget_bytes / NTLP_SERVE_BYTES_PER_DROPLET; every
 * iteration steps all pooled droplets together (one captured hipGraph per
 * pool size class when use_graph=1) and releases the requests that finish.
 * A request's result does not depend on what else is in the pool.
 *
 * Request input  in: 6 column blocks x, y, z, radius, m_s, kappa_s (n values
 *                each, ld_in apart), in_mult[n].
 * Request output out: NTLP_SERVE_COLS column blocks x, y, z, vx, vy, vz, Tp,
 *                radius, m_s, kappa_s, rc, res, actres, numact (ld_out
 *                apart); out_mult[n].
 * The four buffers are read and written asynchronously by the serving thread
 * and must stay valid until ntlp_serve_wait returns for the ticket (pinned
 * host memory makes the copies asynchronous DMA).
 * While the loop is open, ntlp_gpu_init, ntlp_gpu_finalize and
 * ntlp_gpu_set_fields return NTLP_E_STATE. */
#define NTLP_SERVE_COLS 14
#define NTLP_SERVE_BYTES_PER_DROPLET 436
typedef struct ntlp_serve_config {
  int64_t hbm_budget_bytes; /* device memory of the pool (sets its capacity)     */
  int32_t steps;            /* steps per request                                 */
  int32_t precision;        /* NTLP_PREC_*                                       */
  int32_t use_graph;        /* 1: replay captured hipGraphs; 0: launch directly   */
  int32_t mode;             /* NTLP_MODE_BE (or 0) or NTLP_MODE_RK3               */
} ntlp_serve_config;
typedef struct ntlp_serve_totals {
  int64_t submitted, completed, queued;
  int64_t iterations, droplet_steps, max_pool, graph_launches, graphs, capacity;
  ntlp_counters newton; /* Newton bookkeeping over every served droplet-step */
  /* host time of the engine thread (ms): its whole run, inside the copy
   * calls that start uploads and downloads, and blocked on a download, an
   * upload or a control block still in use by the device */
  double ms_engine, ms_copy_calls, ms_wait_download, ms_wait_upload, ms_wait_control;
} ntlp_serve_totals;
int ntlp_serve_open(const ntlp_serve_config *cfg, int64_t *capacity);
/* Queues a request (1 <= n <= capacity, n <= 2^23) and returns at once.
 * ld_in / ld_out: distance in elements between the column blocks of in and
 * out (0 means n). */
int ntlp_serve_submit(const double *in, int64_t ld_in, const int64_t *in_mult, int64_t n, double *out,
                      int64_t ld_out, int64_t *out_mult, int64_t *ticket);
/* Queues a request whose n droplets are drawn on the device from `ens`
 * (ntlp_gpu_particles_generate's ensemble, created at admission): nothing is
 * uploaded.  out / out_mult as for ntlp_serve_submit. */
int ntlp_serve_submit_ensemble(const ntlp_ensemble *ens, int64_t n, double *out, int64_t ld_out, int64_t *out_mult,
                               int64_t *ticket);
/* 1 when the ticket's result is in place, 0 while it is queued or running. */
int ntlp_serve_poll(int64_t ticket);
/* Blocks until the result is in place; times_ms (may be NULL) receives the
 * latency from submission to result and the queueing delay before admission.
 * The ticket is released. */
int ntlp_serve_wait(int64_t ticket, double *times_ms);
/* The loop's totals so far: request counts, and the engine's iteration
 * counters and host times as of its last iteration (newton is filled only at
 * close). */
int ntlp_serve_stats(ntlp_serve_totals *out);
/* Serves every queued request to the end, stops the loop and releases the
 * pool; final_stats (may be NULL) receives the loop's totals. */
int ntlp_serve_close(ntlp_serve_totals *final_stats);
/* Page-locked host memory for request buffers (hipHostMalloc / hipHostFree). */
int ntlp_host_alloc(int64_t bytes, void **ptr);
int ntlp_host_free(void *ptr);

#ifdef __cplusplus
}
#endif
#endif
