! ntlp_dropin_driver.f90 -- the droplet part of les.F's time loop on the GPU
! library, driven from Fortran through the ntlp_gpu module.
!
! It stands for the reference's call sequence (les.F:395-431) with the GPU
! entry points in place of the CPU ones:
!
!   setup      particle_setup            -> gpu_setup
!   each step  fill_ext                  -> gpu_set_fields (fields fixed here)
!              particle_update_BE        -> gpu_particle_update_BE(it)
!              particle_coupling_update  -> gpu_particle_coupling_update(...)
!
! Usage: ntlp_dropin_driver <input file> <output file>
!
! Input (unformatted stream, native byte order):
!   int32  nnx, nny, nnz, n, nsteps, it0, precision, layout
!   real8  xl, yl
!   real8  z(0:nnz+1), zz(0:nnz+1), dzw(0:nnz+2), dzu(0:nnz+2)
!   ntlp_consts_t, ntlp_flags_t            (the C structs, field by field)
!   real8  u, v, w, T, q                   (each nx_f*ny_f*nz_f values in `layout`)
!   int32  field_count                     (values per field array)
!   real8  soa(n, 34); int64 mult(n); int32 pidx(n), procidx(n)
! Output (same form):
!   real8  soa(n, 34); int64 mult(n); int32 pidx(n), procidx(n)
!   real8  partsrc_t(0:nnz+1,1:nny+1,1:nnx+1,1:3), partTsrc_t, partHsrc_t, partTEsrc_t
!   int64  num100, numimpos, actnum, denum, gn_iters, lm_iters, solves, steps
!   int64  droplets kept by read_part_res of the file save_particles wrote
!   real8  the largest |radius difference| between the droplets read back and
!          those saved (one column, fetched with gpu_get_column)
!
! The restart round trip (save_particles + read_part_res, particles.f90:
! 1988-2146) writes <output file>.restart.
program ntlp_dropin_driver
  use iso_c_binding
  use ntlp_gpu
  implicit none
  integer(c_int32_t) :: hdr(8), nfield
  integer :: nnx, nny, nnz, n, nsteps, it0, prec, step, uin, uout
  integer(c_int32_t) :: layout
  real(c_double) :: xl, yl
  real(c_double), allocatable, target :: z(:), zz(:), dzw(:), dzu(:)
  real(c_double), allocatable :: u(:), v(:), w(:), T(:), q(:)
  real(c_double), allocatable :: soa(:, :)
  integer(c_int64_t), allocatable :: mult(:)
  integer(c_int32_t), allocatable :: pidx(:), procidx(:)
  real(c_double), allocatable :: partsrc(:, :, :, :), partTsrc(:, :, :), partHsrc(:, :, :), partTEsrc(:, :, :)
  type(ntlp_consts_t) :: consts
  type(ntlp_flags_t) :: flags
  type(ntlp_counters_t) :: cnt
  character(len=4096) :: fin, fout
  integer(c_int64_t) :: kept
  real(c_double), allocatable :: col(:)
  real(c_double) :: maxdiff

  if (command_argument_count() /= 2) then
     write(*, '(a)') 'usage: ntlp_dropin_driver <input> <output>'
     error stop 2
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)

  open(newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(uin) hdr
  nnx = hdr(1); nny = hdr(2); nnz = hdr(3); n = hdr(4); nsteps = hdr(5); it0 = hdr(6); prec = hdr(7)
  layout = hdr(8)
  read(uin) xl, yl
  allocate(z(0:nnz+1), zz(0:nnz+1), dzw(0:nnz+2), dzu(0:nnz+2))
  read(uin) z, zz, dzw, dzu
  read(uin) consts
  read(uin) flags
  read(uin) nfield
  allocate(u(nfield), v(nfield), w(nfield), T(nfield), q(nfield))
  read(uin) u, v, w, T, q
  allocate(soa(n, NTLP_NREAL), mult(n), pidx(n), procidx(n))
  read(uin) soa, mult, pidx, procidx
  close(uin)

  ! particle_setup, then fill_ext and the droplets
  call gpu_setup(nnx, nny, nnz, xl, yl, z, zz, dzw, dzu, consts, flags)
  call gpu_set_fields(u, v, w, T, q, layout)
  call gpu_put(soa, mult, pidx, procidx)
  call gpu_counters(cnt, .true.)

  allocate(partsrc(0:nnz+1, 1:nny+1, 1:nnx+1, 1:3), partTsrc(0:nnz+1, 1:nny+1, 1:nnx+1), &
       partHsrc(0:nnz+1, 1:nny+1, 1:nnx+1), partTEsrc(0:nnz+1, 1:nny+1, 1:nnx+1))
  do step = 0, nsteps - 1
     call gpu_particle_update_BE(it0 + step, prec)
     call gpu_particle_coupling_update(partsrc, partTsrc, partHsrc, partTEsrc, deterministic=.true.)
  end do

  call gpu_get(soa, mult, pidx, procidx)
  call gpu_counters(cnt, .false.)

  ! restart round trip: the records saved, read back, and one column compared
  call gpu_save_particles(trim(fout) // '.restart')
  call gpu_read_part_res(trim(fout) // '.restart', kept)
  allocate(col(kept))
  call gpu_get_column(C_RADIUS, col)
  maxdiff = -1.0_c_double
  if (kept == n) maxdiff = maxval(abs(col - soa(:, C_RADIUS)))
  call gpu_check(ntlp_gpu_finalize(), 'ntlp_gpu_finalize')

  open(newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(uout) soa, mult, pidx, procidx
  write(uout) partsrc, partTsrc, partHsrc, partTEsrc
  write(uout) cnt%num100, cnt%numimpos, cnt%actnum, cnt%denum, cnt%gn_iters, cnt%lm_iters, cnt%solves, cnt%steps
  write(uout) kept, maxdiff
  close(uout)
  write(*, '(a,i0,a,i0,a,i0,a)') 'ntlp_dropin_driver: ', n, ' droplets, ', nsteps, ' steps, ', cnt%solves, ' solves'
end program ntlp_dropin_driver
