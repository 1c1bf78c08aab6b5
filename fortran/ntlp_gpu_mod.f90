! ntlp_gpu_mod.f90 -- ISO_C_BINDING interface to libntlp_gpu.so.
!
! The derived types and interfaces below mirror include/ntlp_gpu.h one to one.
! The module procedures at the end carry the names of the reference's own
! entry points so that the time loop of les.F can call them in place of the
! CPU ones:
!
!   les.F:404  call particle_update_rk3(istage)  ->  call gpu_particle_update_rk3(istage, it)
!   les.F:422  call particle_update_BE           ->  call gpu_particle_update_BE(it)
!   les.F:431  call particle_coupling_update     ->  call gpu_particle_coupling_update(partsrc, ...)
!   les.F:435  call particle_coupling_exchange   ->  call gpu_particle_coupling_exchange(partsrc, ...)
!   les.F:439  call apply_particle_coupling      ->  call gpu_apply_particle_coupling(t1, t2)
!   particles.f90:91  fill_ext                   ->  call gpu_set_fields(u, v, w, T, q, layout)
!
! Every wrapper stops the program with the library's error text when a call
! fails, the way the reference stops on an MPI or allocation error.
module ntlp_gpu
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: NTLP_OK = 0
  integer(c_int32_t), parameter, public :: NTLP_LAYOUT_EXT = 0, NTLP_LAYOUT_CORE = 1, NTLP_LAYOUT_XPENCIL = 2
  integer(c_int32_t), parameter, public :: NTLP_MODE_RK3 = 1, NTLP_MODE_BE = 2
  integer(c_int32_t), parameter, public :: NTLP_PREC_FP64 = 0, NTLP_PREC_MIXED = 1, NTLP_PREC_FP64_AJ = 2
  integer(c_int32_t), parameter, public :: NTLP_DIST_LOGNORMAL = 0, NTLP_DIST_LOGUNIFORM = 1
  integer(c_int32_t), parameter, public :: NTLP_SALT_FIXED = 0, NTLP_SALT_DRY_LOGNORMAL = 1, NTLP_SALT_CHOICE3 = 2
  integer(c_int32_t), parameter, public :: NTLP_OPT_STAGED_GATHER = 1
  integer, parameter, public :: NTLP_NREAL = 34, NTLP_RECORD_BYTES = 352, NTLP_GN_HIST = 64

  ! column numbers (1-based) of real(8) :: soa(n, NTLP_NREAL)
  integer, parameter, public :: C_X = 1, C_Y = 2, C_Z = 3, C_VX = 4, C_VY = 5, C_VZ = 6, &
       C_UFX = 7, C_UFY = 8, C_UFZ = 9, C_XRHSX = 10, C_XRHSY = 11, C_XRHSZ = 12, &
       C_VRHSX = 13, C_VRHSY = 14, C_VRHSZ = 15, C_TP = 16, C_TPRHS_S = 17, C_TPRHS_L = 18, &
       C_TF = 19, C_RADIUS = 20, C_RADRHS = 21, C_QINF = 22, C_QSTAR = 23, C_RES = 24, &
       C_M_S = 25, C_KAPPA_S = 26, C_RC = 27, C_ACTRES = 28, C_NUMACT = 29, &
       C_VXOLD = 30, C_VYOLD = 31, C_VZOLD = 32, C_TPOLD = 33, C_RADIUSOLD = 34

  type, bind(c), public :: ntlp_grid_t
     integer(c_int32_t) :: nnx, nny, nnz, reserved
     real(c_double) :: dx, dy, xl, yl
     type(c_ptr) :: z, zz, dzw, dzu
  end type ntlp_grid_t

  type, bind(c), public :: ntlp_consts_t
     real(c_double) :: dt
     real(c_double) :: nuf, Cpa, Cpp, CpaCpp, Pra, Sc, Rd, rhow, rhos, Mw, Ru, Gam, Lv, Cpv
     real(c_double) :: surf_rho, surf_p, tsfcc1, grav, radius_init, pi2
     real(c_double) :: part_grav(3)
     real(c_double) :: zetas(3), gama(3)
  end type ntlp_consts_t

  type, bind(c), public :: ntlp_flags_t
     integer(c_int32_t) :: ilin, icouple, iTcouple, iHcouple, ievap, iexner, icase, reserved
  end type ntlp_flags_t

  ! A droplet ensemble drawn on the device (ntlp_gpu_particles_generate).
  type, bind(c), public :: ntlp_ensemble_t
     integer(c_int64_t) :: seed
     integer(c_int32_t) :: radius_kind, radius2_kind
     real(c_double) :: radius_p(2), radius2_p(2), frac2
     integer(c_int32_t) :: salt_kind, reserved
     real(c_double) :: salt_p(3)
     real(c_double) :: kappa_s
     integer(c_int64_t) :: mult
     integer(c_int64_t) :: first
  end type ntlp_ensemble_t

  type, bind(c), public :: ntlp_counters_t
     integer(c_int64_t) :: num100, numimpos, actnum, denum
     integer(c_int64_t) :: gn_iters, lm_iters, solves, steps
     integer(c_int64_t) :: gn_hist(NTLP_GN_HIST)
  end type ntlp_counters_t

  public :: ntlp_gpu_init, ntlp_gpu_finalize, ntlp_gpu_last_error, ntlp_gpu_set_dt, ntlp_gpu_set_fields, &
       ntlp_gpu_particles_put, ntlp_gpu_particles_get, ntlp_gpu_particles_put_records, &
       ntlp_gpu_particles_get_records, ntlp_gpu_particles_count, ntlp_gpu_interp, ntlp_gpu_crit_radius, &
       ntlp_gpu_particle_update_rk3, ntlp_gpu_particle_update_be, ntlp_gpu_particle_update_rk3_prec, &
       ntlp_gpu_particle_update_be_prec, ntlp_gpu_particle_coupling_update, ntlp_gpu_get_sources, &
       ntlp_gpu_particle_coupling_exchange, ntlp_gpu_apply_particle_coupling, &
       ntlp_gpu_bin, ntlp_gpu_wcell_counts, ntlp_gpu_num_ucells, ntlp_gpu_step, ntlp_gpu_counters, &
       ntlp_gpu_calc_tauc_min, ntlp_gpu_particle_xy_stats, &
       ntlp_gpu_step_counters, ntlp_gpu_sync, ntlp_gpu_be_flux, ntlp_gpu_particles_get_column, &
       ntlp_gpu_particles_generate, ntlp_gpu_save_particles, ntlp_gpu_read_part_res, &
       ntlp_gpu_particle_xy_stats_noalloc, ntlp_gpu_set_option, ntlp_gpu_timing, ntlp_gpu_kernel_times, &
       ntlp_gpu_copy_bandwidth
  public :: gpu_be_flux, gpu_save_particles, gpu_read_part_res, gpu_get_column, gpu_generate, gpu_set_option
  public :: gpu_check, gpu_setup, gpu_set_fields, gpu_put, gpu_get, gpu_interp, gpu_crit_radius, &
       gpu_particle_update_rk3, gpu_particle_update_BE, gpu_particle_coupling_update, gpu_step, gpu_counters, &
       gpu_particle_coupling_exchange, gpu_apply_particle_coupling, gpu_calc_tauc_min, gpu_particle_xy_stats

  interface
     integer(c_int32_t) function ntlp_gpu_init(grid, consts, flags) bind(c, name="ntlp_gpu_init")
       import :: c_int32_t, ntlp_grid_t, ntlp_consts_t, ntlp_flags_t
       type(ntlp_grid_t), intent(in) :: grid
       type(ntlp_consts_t), intent(in) :: consts
       type(ntlp_flags_t), intent(in) :: flags
     end function ntlp_gpu_init
     integer(c_int32_t) function ntlp_gpu_finalize() bind(c, name="ntlp_gpu_finalize")
       import :: c_int32_t
     end function ntlp_gpu_finalize
     type(c_ptr) function ntlp_gpu_last_error() bind(c, name="ntlp_gpu_last_error")
       import :: c_ptr
     end function ntlp_gpu_last_error
     integer(c_int32_t) function ntlp_gpu_set_dt(dt) bind(c, name="ntlp_gpu_set_dt")
       import :: c_int32_t, c_double
       real(c_double), value :: dt
     end function ntlp_gpu_set_dt
     integer(c_int32_t) function ntlp_gpu_set_fields(u, v, w, T, q, layout) bind(c, name="ntlp_gpu_set_fields")
       import :: c_int32_t, c_double
       real(c_double), intent(in) :: u(*), v(*), w(*), T(*), q(*)
       integer(c_int32_t), value :: layout
     end function ntlp_gpu_set_fields
     integer(c_int32_t) function ntlp_gpu_particles_put(soa, mult, pidx, procidx, n) &
          bind(c, name="ntlp_gpu_particles_put")
       import :: c_int32_t, c_int64_t, c_double
       real(c_double), intent(in) :: soa(*)
       integer(c_int64_t), intent(in) :: mult(*)
       integer(c_int32_t), intent(in) :: pidx(*), procidx(*)
       integer(c_int64_t), value :: n
     end function ntlp_gpu_particles_put
     integer(c_int32_t) function ntlp_gpu_particles_get(soa, mult, pidx, procidx, n) &
          bind(c, name="ntlp_gpu_particles_get")
       import :: c_int32_t, c_int64_t, c_double
       real(c_double), intent(out) :: soa(*)
       integer(c_int64_t), intent(out) :: mult(*)
       integer(c_int32_t), intent(out) :: pidx(*), procidx(*)
       integer(c_int64_t), value :: n
     end function ntlp_gpu_particles_get
     integer(c_int32_t) function ntlp_gpu_particles_put_records(rec, n) bind(c, name="ntlp_gpu_particles_put_records")
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: rec
       integer(c_int64_t), value :: n
     end function ntlp_gpu_particles_put_records
     integer(c_int32_t) function ntlp_gpu_particles_get_records(rec, n) bind(c, name="ntlp_gpu_particles_get_records")
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: rec
       integer(c_int64_t), value :: n
     end function ntlp_gpu_particles_get_records
     integer(c_int64_t) function ntlp_gpu_particles_count() bind(c, name="ntlp_gpu_particles_count")
       import :: c_int64_t
     end function ntlp_gpu_particles_count
     integer(c_int32_t) function ntlp_gpu_particles_get_column(column, out, n) &
          bind(c, name="ntlp_gpu_particles_get_column")
       import :: c_int32_t, c_int64_t, c_double
       integer(c_int32_t), value :: column
       real(c_double), intent(out) :: out(*)
       integer(c_int64_t), value :: n
     end function ntlp_gpu_particles_get_column
     integer(c_int32_t) function ntlp_gpu_particles_generate(ens, counts, nens) &
          bind(c, name="ntlp_gpu_particles_generate")
       import :: c_int32_t, c_int64_t, ntlp_ensemble_t
       type(ntlp_ensemble_t), intent(in) :: ens(*)
       integer(c_int64_t), intent(in) :: counts(*)
       integer(c_int32_t), value :: nens
     end function ntlp_gpu_particles_generate
     integer(c_int32_t) function ntlp_gpu_save_particles(path) bind(c, name="ntlp_gpu_save_particles")
       import :: c_int32_t, c_char
       character(kind=c_char), intent(in) :: path(*)
     end function ntlp_gpu_save_particles
     integer(c_int32_t) function ntlp_gpu_read_part_res(path, kept) bind(c, name="ntlp_gpu_read_part_res")
       import :: c_int32_t, c_int64_t, c_char
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int64_t), intent(out) :: kept
     end function ntlp_gpu_read_part_res
     integer(c_int32_t) function ntlp_gpu_interp(set_vp, set_tp) bind(c, name="ntlp_gpu_interp")
       import :: c_int32_t
       integer(c_int32_t), value :: set_vp, set_tp
     end function ntlp_gpu_interp
     integer(c_int32_t) function ntlp_gpu_crit_radius() bind(c, name="ntlp_gpu_crit_radius")
       import :: c_int32_t
     end function ntlp_gpu_crit_radius
     integer(c_int32_t) function ntlp_gpu_particle_update_rk3(istage, it) bind(c, name="ntlp_gpu_particle_update_rk3")
       import :: c_int32_t
       integer(c_int32_t), value :: istage, it
     end function ntlp_gpu_particle_update_rk3
     integer(c_int32_t) function ntlp_gpu_particle_update_be(it) bind(c, name="ntlp_gpu_particle_update_be")
       import :: c_int32_t
       integer(c_int32_t), value :: it
     end function ntlp_gpu_particle_update_be
     integer(c_int32_t) function ntlp_gpu_particle_update_rk3_prec(istage, it, prec) &
          bind(c, name="ntlp_gpu_particle_update_rk3_prec")
       import :: c_int32_t
       integer(c_int32_t), value :: istage, it, prec
     end function ntlp_gpu_particle_update_rk3_prec
     integer(c_int32_t) function ntlp_gpu_particle_update_be_prec(it, prec) bind(c, name="ntlp_gpu_particle_update_be_prec")
       import :: c_int32_t
       integer(c_int32_t), value :: it, prec
     end function ntlp_gpu_particle_update_be_prec
     integer(c_int32_t) function ntlp_gpu_particle_coupling_update(src_out, deterministic) &
          bind(c, name="ntlp_gpu_particle_coupling_update")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: src_out
       integer(c_int32_t), value :: deterministic
     end function ntlp_gpu_particle_coupling_update
     integer(c_int32_t) function ntlp_gpu_particle_coupling_exchange(partsrc, partTsrc, partHsrc, partTEsrc) &
          bind(c, name="ntlp_gpu_particle_coupling_exchange")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: partsrc, partTsrc, partHsrc, partTEsrc
     end function ntlp_gpu_particle_coupling_exchange
     integer(c_int32_t) function ntlp_gpu_apply_particle_coupling(T, q, layout) &
          bind(c, name="ntlp_gpu_apply_particle_coupling")
       import :: c_int32_t, c_double
       real(c_double), intent(inout) :: T(*), q(*)
       integer(c_int32_t), value :: layout
     end function ntlp_gpu_apply_particle_coupling
     integer(c_int32_t) function ntlp_gpu_get_sources(src_out) bind(c, name="ntlp_gpu_get_sources")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: src_out
     end function ntlp_gpu_get_sources
     integer(c_int32_t) function ntlp_gpu_bin(cellidx, counts, offsets, perm) bind(c, name="ntlp_gpu_bin")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: cellidx, counts, offsets, perm
     end function ntlp_gpu_bin
     integer(c_int32_t) function ntlp_gpu_wcell_counts(num_array) bind(c, name="ntlp_gpu_wcell_counts")
       import :: c_int32_t
       integer(c_int32_t), intent(out) :: num_array(*)
     end function ntlp_gpu_wcell_counts
     integer(c_int32_t) function ntlp_gpu_calc_tauc_min(tauc_min) bind(c, name="ntlp_gpu_calc_tauc_min")
       import :: c_int32_t, c_double
       real(c_double), intent(out) :: tauc_min
     end function ntlp_gpu_calc_tauc_min
     integer(c_int32_t) function ntlp_gpu_particle_xy_stats(s0, s1, lwp) bind(c, name="ntlp_gpu_particle_xy_stats")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: s0, s1, lwp
     end function ntlp_gpu_particle_xy_stats
     integer(c_int32_t) function ntlp_gpu_particle_xy_stats_noalloc(s0, s1, lwp) &
          bind(c, name="ntlp_gpu_particle_xy_stats_noalloc")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: s0, s1, lwp
     end function ntlp_gpu_particle_xy_stats_noalloc
     integer(c_int64_t) function ntlp_gpu_num_ucells() bind(c, name="ntlp_gpu_num_ucells")
       import :: c_int64_t
     end function ntlp_gpu_num_ucells
     integer(c_int32_t) function ntlp_gpu_step(mode, prec, nsteps, it0) bind(c, name="ntlp_gpu_step")
       import :: c_int32_t
       integer(c_int32_t), value :: mode, prec, nsteps, it0
     end function ntlp_gpu_step
     integer(c_int32_t) function ntlp_gpu_counters(out, reset) bind(c, name="ntlp_gpu_counters")
       import :: c_int32_t, ntlp_counters_t
       type(ntlp_counters_t), intent(out) :: out
       integer(c_int32_t), value :: reset
     end function ntlp_gpu_counters
     integer(c_int32_t) function ntlp_gpu_step_counters(rows, max_rows) bind(c, name="ntlp_gpu_step_counters")
       import :: c_int32_t, c_int64_t
       integer(c_int64_t), intent(out) :: rows(*)
       integer(c_int32_t), value :: max_rows
     end function ntlp_gpu_step_counters
     integer(c_int32_t) function ntlp_gpu_be_flux(pflux, pmassflux, penegflux) bind(c, name="ntlp_gpu_be_flux")
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: pflux, pmassflux, penegflux
     end function ntlp_gpu_be_flux
     integer(c_int32_t) function ntlp_gpu_timing(enable) bind(c, name="ntlp_gpu_timing")
       import :: c_int32_t
       integer(c_int32_t), value :: enable
     end function ntlp_gpu_timing
     integer(c_int32_t) function ntlp_gpu_kernel_times(names, launches, total_ms, cap) &
          bind(c, name="ntlp_gpu_kernel_times")
       import :: c_int32_t, c_int64_t, c_double, c_char
       character(kind=c_char), intent(out) :: names(*)
       integer(c_int64_t), intent(out) :: launches(*)
       real(c_double), intent(out) :: total_ms(*)
       integer(c_int32_t), value :: cap
     end function ntlp_gpu_kernel_times
     integer(c_int32_t) function ntlp_gpu_copy_bandwidth(bytes, reps, gbps) bind(c, name="ntlp_gpu_copy_bandwidth")
       import :: c_int32_t, c_int64_t, c_double
       integer(c_int64_t), value :: bytes
       integer(c_int32_t), value :: reps
       real(c_double), intent(out) :: gbps
     end function ntlp_gpu_copy_bandwidth
     integer(c_int32_t) function ntlp_gpu_set_option(option, val) bind(c, name="ntlp_gpu_set_option")
       import :: c_int32_t, c_int64_t
       integer(c_int32_t), value :: option
       integer(c_int64_t), value :: val
     end function ntlp_gpu_set_option
     integer(c_int32_t) function ntlp_gpu_sync() bind(c, name="ntlp_gpu_sync")
       import :: c_int32_t
     end function ntlp_gpu_sync
     integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! Stops with the library's error text when rc is not NTLP_OK.
  subroutine gpu_check(rc, what)
    integer(c_int32_t), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer(c_size_t) :: l
    if (rc == NTLP_OK) return
    p = ntlp_gpu_last_error()
    l = c_strlen(p)
    call c_f_pointer(p, msg, [l])
    write(*, '(a,a,i0,a,*(a))') what, ' failed (', rc, '): ', msg
    error stop 1
  end subroutine gpu_check

  ! particle_setup's grid and constants (particles.f90:1829-1986).  z, zz are
  ! z(0:nnz+1), zz(0:nnz+1); dzw, dzu are dzw(0:nnz+2), dzu(0:nnz+2).
  subroutine gpu_setup(nnx, nny, nnz, xl, yl, z, zz, dzw, dzu, consts, flags)
    integer, intent(in) :: nnx, nny, nnz
    real(c_double), intent(in) :: xl, yl
    real(c_double), intent(in), target :: z(0:nnz+1), zz(0:nnz+1), dzw(0:nnz+2), dzu(0:nnz+2)
    type(ntlp_consts_t), intent(in) :: consts
    type(ntlp_flags_t), intent(in) :: flags
    type(ntlp_grid_t) :: g
    g%nnx = nnx; g%nny = nny; g%nnz = nnz; g%reserved = 0
    g%xl = xl; g%yl = yl; g%dx = xl / nnx; g%dy = yl / nny
    g%z = c_loc(z(0)); g%zz = c_loc(zz(0)); g%dzw = c_loc(dzw(0)); g%dzu = c_loc(dzu(0))
    call gpu_check(ntlp_gpu_init(g, consts, flags), 'ntlp_gpu_init')
  end subroutine gpu_setup

  ! fill_ext (particles.f90:91-328): the resident copy of u, v, w, T, q.
  subroutine gpu_set_fields(u, v, w, T, q, layout)
    real(c_double), intent(in) :: u(*), v(*), w(*), T(*), q(*)
    integer(c_int32_t), intent(in) :: layout
    call gpu_check(ntlp_gpu_set_fields(u, v, w, T, q, layout), 'ntlp_gpu_set_fields')
  end subroutine gpu_set_fields

  subroutine gpu_put(soa, mult, pidx, procidx)
    real(c_double), intent(in) :: soa(:, :)
    integer(c_int64_t), intent(in) :: mult(:)
    integer(c_int32_t), intent(in) :: pidx(:), procidx(:)
    call gpu_check(ntlp_gpu_particles_put(soa, mult, pidx, procidx, int(size(soa, 1), c_int64_t)), &
         'ntlp_gpu_particles_put')
  end subroutine gpu_put

  subroutine gpu_get(soa, mult, pidx, procidx)
    real(c_double), intent(out) :: soa(:, :)
    integer(c_int64_t), intent(out) :: mult(:)
    integer(c_int32_t), intent(out) :: pidx(:), procidx(:)
    call gpu_check(ntlp_gpu_particles_get(soa, mult, pidx, procidx, int(size(soa, 1), c_int64_t)), &
         'ntlp_gpu_particles_get')
  end subroutine gpu_get

  subroutine gpu_interp(set_vp, set_tp)
    logical, intent(in) :: set_vp, set_tp
    call gpu_check(ntlp_gpu_interp(merge(1_c_int32_t, 0_c_int32_t, set_vp), merge(1_c_int32_t, 0_c_int32_t, set_tp)), &
         'ntlp_gpu_interp')
  end subroutine gpu_interp

  subroutine gpu_crit_radius()
    call gpu_check(ntlp_gpu_crit_radius(), 'ntlp_gpu_crit_radius')
  end subroutine gpu_crit_radius

  ! particle_update_rk3(istage) (particles.f90:2845-2998); it is les.F's step counter.
  subroutine gpu_particle_update_rk3(istage, it)
    integer, intent(in) :: istage, it
    call gpu_check(ntlp_gpu_particle_update_rk3(int(istage, c_int32_t), int(it, c_int32_t)), 'particle_update_rk3')
  end subroutine gpu_particle_update_rk3

  ! particle_update_BE (particles.f90:3000-3280); prec selects NTLP_PREC_* (default FP64).
  subroutine gpu_particle_update_BE(it, prec)
    integer, intent(in) :: it
    integer, intent(in), optional :: prec
    integer(c_int32_t) :: p
    p = NTLP_PREC_FP64
    if (present(prec)) p = int(prec, c_int32_t)
    call gpu_check(ntlp_gpu_particle_update_be_prec(int(it, c_int32_t), p), 'particle_update_BE')
  end subroutine gpu_particle_update_BE

  ! particle_coupling_update (particles.f90:1151-1276): partsrc_t(0:nnz+1,1:nny+1,1:nnx+1,1:3),
  ! partTsrc_t, partHsrc_t, partTEsrc_t, each (0:nnz+1,1:nny+1,1:nnx+1).
  subroutine gpu_particle_coupling_update(partsrc_t, partTsrc_t, partHsrc_t, partTEsrc_t, deterministic)
    real(c_double), intent(out) :: partsrc_t(:, :, :, :), partTsrc_t(:, :, :), partHsrc_t(:, :, :), partTEsrc_t(:, :, :)
    logical, intent(in), optional :: deterministic
    real(c_double), allocatable, target :: buf(:)
    integer :: m
    integer(c_int32_t) :: det
    det = 0
    if (present(deterministic)) then
       if (deterministic) det = 1
    end if
    m = size(partTsrc_t)
    allocate(buf(6 * m))
    call gpu_check(ntlp_gpu_particle_coupling_update(c_loc(buf(1)), det), 'particle_coupling_update')
    partsrc_t = reshape(buf(1:3 * m), shape(partsrc_t))
    partTsrc_t = reshape(buf(3 * m + 1:4 * m), shape(partTsrc_t))
    partHsrc_t = reshape(buf(4 * m + 1:5 * m), shape(partHsrc_t))
    partTEsrc_t = reshape(buf(5 * m + 1:6 * m), shape(partTEsrc_t))
  end subroutine gpu_particle_coupling_update

  ! particle_coupling_exchange (particles.f90:1066-1149) on one periodic domain:
  ! partsrc(1:nnx,1:nny,0:nnz+1,1:3), partTsrc, partHsrc, partTEsrc(1:nnx,1:nny,0:nnz+1).
  subroutine gpu_particle_coupling_exchange(partsrc, partTsrc, partHsrc, partTEsrc)
    real(c_double), intent(out), target :: partsrc(:, :, :, :), partTsrc(:, :, :), partHsrc(:, :, :), partTEsrc(:, :, :)
    call gpu_check(ntlp_gpu_particle_coupling_exchange(c_loc(partsrc), c_loc(partTsrc), c_loc(partHsrc), &
         c_loc(partTEsrc)), 'particle_coupling_exchange')
  end subroutine gpu_particle_coupling_exchange

  ! apply_particle_coupling (particles.f90:1278-1313) on t(:,:,1,:) and t(:,:,2,:)
  ! given as x-pencil arrays (1:nnx+2, 1:nny, 0:nnz+1).
  subroutine gpu_apply_particle_coupling(T, q)
    real(c_double), intent(inout) :: T(:, :, :), q(:, :, :)
    call gpu_check(ntlp_gpu_apply_particle_coupling(T, q, NTLP_LAYOUT_XPENCIL), 'apply_particle_coupling')
  end subroutine gpu_apply_particle_coupling

  ! calc_tauc_min (particles.f90:3310-3368) on one domain.
  subroutine gpu_calc_tauc_min(tauc_min)
    real(c_double), intent(out) :: tauc_min
    call gpu_check(ntlp_gpu_calc_tauc_min(tauc_min), 'calc_tauc_min')
  end subroutine gpu_calc_tauc_min

  ! particle_xy_stats (particles.f90:3370-3770) before its divisions:
  ! s0(21), statsvec1(0:nnz+1, 24), LWP(1:nnx, 1:nny).
  subroutine gpu_particle_xy_stats(s0, statsvec1, lwp)
    real(c_double), intent(out), target :: s0(21), statsvec1(:, :), lwp(:, :)
    call gpu_check(ntlp_gpu_particle_xy_stats(c_loc(s0), c_loc(statsvec1), c_loc(lwp)), 'particle_xy_stats')
  end subroutine gpu_particle_xy_stats

  ! nsteps fused droplet steps (gather, advance, boundary conditions, coupling, sort).
  subroutine gpu_step(mode, prec, nsteps, it0)
    integer, intent(in) :: mode, prec, nsteps, it0
    call gpu_check(ntlp_gpu_step(int(mode, c_int32_t), int(prec, c_int32_t), int(nsteps, c_int32_t), &
         int(it0, c_int32_t)), 'ntlp_gpu_step')
  end subroutine gpu_step

  ! pflux, pmassflux, penegflux(0:nnz) of the last particle_update_BE step
  ! (particles.f90:3208-3238), before les.F divides them by dt*xl*yl.
  subroutine gpu_be_flux(pflux, pmassflux, penegflux)
    real(c_double), intent(out), target :: pflux(0:), pmassflux(0:), penegflux(0:)
    call gpu_check(ntlp_gpu_be_flux(c_loc(pflux), c_loc(pmassflux), c_loc(penegflux)), 'ntlp_gpu_be_flux')
  end subroutine gpu_be_flux

  ! save_particles (particles.f90:1988-2046) for one rank: an int32 droplet
  ! count followed by one 352-byte record per droplet.
  subroutine gpu_save_particles(path)
    character(*), intent(in) :: path
    call gpu_check(ntlp_gpu_save_particles(trim(path) // c_null_char), 'save_particles')
  end subroutine gpu_save_particles

  ! read_part_res (particles.f90:2048-2146) for one rank: keeps the records with
  ! 0 < x < xl and 0 < y < yl and replaces the droplets on the device.
  subroutine gpu_read_part_res(path, kept)
    character(*), intent(in) :: path
    integer(c_int64_t), intent(out), optional :: kept
    integer(c_int64_t) :: k
    call gpu_check(ntlp_gpu_read_part_res(trim(path) // c_null_char, k), 'read_part_res')
    if (present(kept)) kept = k
  end subroutine gpu_read_part_res

  ! One column (C_X ... C_RADIUSOLD, 1-based) of every droplet, in put order.
  subroutine gpu_get_column(column, out)
    integer, intent(in) :: column
    real(c_double), intent(out) :: out(:)
    call gpu_check(ntlp_gpu_particles_get_column(int(column - 1, c_int32_t), out, int(size(out), c_int64_t)), &
         'ntlp_gpu_particles_get_column')
  end subroutine gpu_get_column

  ! particle_init / create_particle's drawing (particles.f90:2226-2300,
  ! 2440-2520) done on the device: counts(e) droplets of ens(e), in order.
  subroutine gpu_generate(ens, counts)
    type(ntlp_ensemble_t), intent(in) :: ens(:)
    integer(c_int64_t), intent(in) :: counts(:)
    if (size(counts) /= size(ens)) call gpu_check(-1_c_int32_t, 'ntlp_gpu_particles_generate (sizes)')
    call gpu_check(ntlp_gpu_particles_generate(ens, counts, int(size(ens), c_int32_t)), 'ntlp_gpu_particles_generate')
  end subroutine gpu_generate

  subroutine gpu_set_option(option, val)
    integer, intent(in) :: option
    integer, intent(in) :: val
    call gpu_check(ntlp_gpu_set_option(int(option, c_int32_t), int(val, c_int64_t)), 'ntlp_gpu_set_option')
  end subroutine gpu_set_option

  subroutine gpu_counters(c, reset)
    type(ntlp_counters_t), intent(out) :: c
    logical, intent(in) :: reset
    call gpu_check(ntlp_gpu_counters(c, merge(1_c_int32_t, 0_c_int32_t, reset)), 'ntlp_gpu_counters')
  end subroutine gpu_counters

end module ntlp_gpu
