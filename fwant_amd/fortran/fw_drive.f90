! Fortran driver through the ISO_C_BINDING interface: the sequence the
! reference driver follows (code/srcF/seis3d_wave_8op_p3.F90:40-157) --
! configuration, grid and metric, media, sources, receivers, PML set-up,
! the time loop, the receiver output -- with libfwant doing the solve.
!
!   fw_drive CASE OUT   run the case file CASE, write the receiver records
!                       (float32 [nsteps][nrec][9]) to OUT
!   fw_drive            probe: print the step count for nt = 1000 and the
!                       status of fw_create (a message, not a stop, when
!                       the machine has no GPU)
!
! Case file (stream, native endianness): int32 magic 'FWNT' (0x544e5746),
! version 1, ni nj nk freenode nsteps nforce ntwf stff nmom ntwm stfm nrec
! tinv; float32 stept; int32 abs_number(6); float32 abs_velocity(6)
! cfs_bmax(6) cfs_amax(6); float64 x(nx) y(ny) z(nz); float32 lambda mu rho
! (nx*ny*nz each, column-major); force: int32 indx(3,n), float32 shift(3,n)
! t0(n) stf_time(ntwf) stf_freq(ntwf) Fx Fy Fz (ntwf*n each); moment: int32
! indx(3,n), float32 shift(3,n) t0(n) stf_time(ntwm) stf_freq(ntwm)
! M6(6*ntwm*n); int32 receiver indx(3,nrec).
program fw_drive
  use iso_c_binding
  use fwant_c
  implicit none
  type(fw_config) :: cfg
  type(c_ptr) :: h
  integer(c_int) :: hdr(15), rc, nrecords
  integer :: nx, ny, nz, nsteps, nf, ntwf, stff, nm, ntwm, stfm, nrec, tinv, u
  real(c_double), allocatable :: x(:), y(:), z(:)
  real(c_float), allocatable :: lam(:), mu(:), rho(:), xsin(:), xcot(:), xi_x(:), eta_y(:), zeta_z(:)
  integer(c_int), allocatable :: findx(:), mindx(:), rindx(:)
  real(c_float), allocatable :: fshift(:), ft0(:), ftime(:), ffreq(:), fx(:), fy(:), fz(:)
  real(c_float), allocatable :: mshift(:), mt0(:), mtime(:), mfreq(:), m6(:), rec(:)
  character(len=4096) :: case_path, out_path

  if (command_argument_count() < 2) then
    write(*, '(a, i0)') 'fw_reference_steps(1000) = ', fw_reference_steps(1000_c_int)
    cfg%ni = 8; cfg%nj = 8; cfg%nk = 8; cfg%stept = 0.1; cfg%freenode = 1
    cfg%abs_number = 0; cfg%abs_velocity = 0.0; cfg%cfs_bmax = 1.0; cfg%cfs_amax = 0.0
    rc = fw_create(cfg, h)
    write(*, '(a, i0, a, a)') 'fw_create status ', rc, ': ', fw_error_message(c_null_ptr)
    if (rc == FW_OK) rc = fw_destroy(h)
    stop
  end if
  call get_command_argument(1, case_path)
  call get_command_argument(2, out_path)

  open(newunit=u, file=trim(case_path), access='stream', form='unformatted', status='old')
  read(u) hdr
  if (hdr(1) /= int(z'544e5746', c_int) .or. hdr(2) /= 1) then
    write(*, '(a)') 'fw_drive: not a version-1 case file'
    stop 1
  end if
  cfg%ni = hdr(3); cfg%nj = hdr(4); cfg%nk = hdr(5); cfg%freenode = hdr(6)
  nsteps = hdr(7); nf = hdr(8); ntwf = hdr(9); stff = hdr(10); nm = hdr(11); ntwm = hdr(12); stfm = hdr(13)
  nrec = hdr(14); tinv = hdr(15)
  read(u) cfg%stept, cfg%abs_number, cfg%abs_velocity, cfg%cfs_bmax, cfg%cfs_amax
  nx = cfg%ni + 6; ny = cfg%nj + 6; nz = cfg%nk + 6
  allocate(x(nx), y(ny), z(nz), lam(nx * ny * nz), mu(nx * ny * nz), rho(nx * ny * nz))
  read(u) x, y, z, lam, mu, rho
  allocate(findx(3 * nf), fshift(3 * nf), ft0(nf), ftime(ntwf), ffreq(ntwf), fx(ntwf * nf), fy(ntwf * nf), &
           fz(ntwf * nf))
  read(u) findx, fshift, ft0, ftime, ffreq, fx, fy, fz
  allocate(mindx(3 * nm), mshift(3 * nm), mt0(nm), mtime(ntwm), mfreq(ntwm), m6(6 * ntwm * nm))
  read(u) mindx, mshift, mt0, mtime, mfreq, m6
  allocate(rindx(3 * nrec))
  read(u) rindx
  close(u)

  call fw_check(fw_create(cfg, h), c_null_ptr, 'fw_create')
  allocate(xsin(nx), xcot(nx), xi_x(nx), eta_y(ny), zeta_z(nz))
  call fw_check(fw_build_metric(h, x, y, z, xsin, xcot, xi_x, eta_y, zeta_z), h, 'fw_build_metric')
  call fw_check(fw_set_media(h, lam, mu, rho), h, 'fw_set_media')
  if (nf > 0) call fw_check(fw_set_force(h, nf, ntwf, stff, findx, fshift, ft0, ftime, ffreq, fx, fy, fz), &
                            h, 'fw_set_force')
  if (nm > 0) call fw_check(fw_set_moment(h, nm, ntwm, stfm, mindx, mshift, mt0, mtime, mfreq, m6), &
                            h, 'fw_set_moment')
  call fw_check(fw_set_receivers(h, nrec, tinv, rindx, nsteps), h, 'fw_set_receivers')
  call fw_check(fw_init(h), h, 'fw_init')
  call fw_check(fw_run(h, 0, nsteps), h, 'fw_run')
  allocate(rec(9 * nrec * nsteps))
  call fw_check(fw_get_receivers(h, rec, nrecords), h, 'fw_get_receivers')
  call fw_check(fw_destroy(h), c_null_ptr, 'fw_destroy')

  open(newunit=u, file=trim(out_path), access='stream', form='unformatted', status='replace')
  write(u) rec(1:9 * nrec * nrecords)
  close(u)
  write(*, '(a, i0, a, i0, a)') 'fw_drive: ', nrecords, ' records of ', nrec, ' receivers'
end program fw_drive
