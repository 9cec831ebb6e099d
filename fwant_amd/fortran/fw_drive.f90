! Fortran driver through the ISO_C_BINDING interface: the sequence the
! reference driver follows (code/srcF/seis3d_wave_8op_p3.F90:40-157) --
! configuration, grid and metric, media, sources, receivers, PML set-up,
! the time loop, the receiver output -- with libfwant doing the solve.
!
!   fw_drive CASE OUT   run the case file CASE, write the receiver records
!                       (float32 [nsteps][nrec][9]) to OUT
!   fw_drive CASE OUT stages
!                       the same, with the time loop written as the
!                       reference's loop body (seis3d_wave_8op_p3.F90:
!                       97-157): macdrp_syn, abs_syn and per RK stage
!                       set_cur_time, src_surface, macdrp_Lx?_Ly?_Lz?,
!                       abs_Lx?_Ly?_Lz?, src_stress, src_force,
!                       macdrp_RK_*, abs_RK_*, calling the drop-in modules of
!                       fwant_shim.f90 (one fw_stage per RK stage)
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
  use macdrp_mod
  use abs_mod
  use src_mod
  implicit none
  ! operator pattern (x << 2 | y << 1 | z, 1 = forward) of the first RK stage
  ! of each step of the 8-step block: BBB FFB FFF BBF BFB FBB FBF BFF
  ! (seis3d_wave_8op_p3.F90:97-591); stages 2 and 4 flip all three
  integer, parameter :: block_pat(0:7) = (/ 0, 6, 7, 1, 2, 4, 5, 3 /)
  integer :: ntime
  logical :: use_shim
  character(len=16) :: mode
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
  mode = ''
  if (command_argument_count() >= 3) call get_command_argument(3, mode)
  use_shim = trim(mode) == 'stages'

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
  if (use_shim) then
    call fwant_shim_attach(h)
    do ntime = 0, nsteps - 1
      call rk_step(ntime, block_pat(mod(ntime, 8)), cfg%stept)
    end do
  else
    call fw_check(fw_run(h, 0, nsteps), h, 'fw_run')
  end if
  allocate(rec(9 * nrec * nsteps))
  call fw_check(fw_get_receivers(h, rec, nrecords), h, 'fw_get_receivers')
  call fw_check(fw_destroy(h), c_null_ptr, 'fw_destroy')

  open(newunit=u, file=trim(out_path), access='stream', form='unformatted', status='replace')
  write(u) rec(1:9 * nrec * nrecords)
  close(u)
  write(*, '(a, i0, a, i0, a)') 'fw_drive: ', nrecords, ' records of ', nrec, ' receivers'

contains

  ! the loop body of seis3d_wave_8op_p3.F90:97-157 for a step whose first
  ! RK stage uses pattern p
  subroutine rk_step(ntime, p, stept)
    integer, intent(in) :: ntime, p
    real(c_float), intent(in) :: stept
    call macdrp_syn
    call abs_syn
    ! the 1th stage
    call set_cur_time(ntime, 0.0_c_float)
    call src_surface(ntime, 0.0_c_float, stept)
    call sweep(p)
    call src_stress(hTxx, hTyy, hTzz, hTxy, hTxz, hTyz, ntime, 0.0_c_float, stept)
    call src_force(hVx, hVy, hVz, ntime, 0.0_c_float, stept)
    call macdrp_RK_beg(firRKa(1), firRKb(1))
    call abs_RK_beg(firRKa(1), firRKb(1))
    ! the 2th stage
    call set_cur_time(ntime, firRKa(1))
    call src_surface(ntime, firRKa(1), stept)
    call sweep(7 - p)
    call src_stress(hTxx, hTyy, hTzz, hTxy, hTxz, hTyz, ntime, firRKa(1), stept)
    call src_force(hVx, hVy, hVz, ntime, firRKa(1), stept)
    call macdrp_RK_inn(firRKa(2), firRKb(2))
    call abs_RK_inn(firRKa(2), firRKb(2))
    ! the 3th stage
    call set_cur_time(ntime, firRKa(2))
    call src_surface(ntime, firRKa(2), stept)
    call sweep(p)
    call src_stress(hTxx, hTyy, hTzz, hTxy, hTxz, hTyz, ntime, firRKa(2), stept)
    call src_force(hVx, hVy, hVz, ntime, firRKa(2), stept)
    call macdrp_RK_inn(firRKa(3), firRKb(3))
    call abs_RK_inn(firRKa(3), firRKb(3))
    ! the 4th stage
    call set_cur_time(ntime, firRKa(3))
    call src_surface(ntime, firRKa(3), stept)
    call sweep(7 - p)
    call src_stress(hTxx, hTyy, hTzz, hTxy, hTxz, hTyz, ntime, firRKa(3), stept)
    call src_force(hVx, hVy, hVz, ntime, firRKa(3), stept)
    call macdrp_RK_fin(firRKb(4))
    call abs_RK_fin(firRKb(4))
  end subroutine

  ! macdrp_Lx?_Ly?_Lz? followed by abs_Lx?_Ly?_Lz? of pattern p
  subroutine sweep(p)
    integer, intent(in) :: p
    select case (p)
    case (0)
      call macdrp_LxB_LyB_LzB; call abs_LxB_LyB_LzB
    case (1)
      call macdrp_LxB_LyB_LzF; call abs_LxB_LyB_LzF
    case (2)
      call macdrp_LxB_LyF_LzB; call abs_LxB_LyF_LzB
    case (3)
      call macdrp_LxB_LyF_LzF; call abs_LxB_LyF_LzF
    case (4)
      call macdrp_LxF_LyB_LzB; call abs_LxF_LyB_LzB
    case (5)
      call macdrp_LxF_LyB_LzF; call abs_LxF_LyB_LzF
    case (6)
      call macdrp_LxF_LyF_LzB; call abs_LxF_LyF_LzB
    case default
      call macdrp_LxF_LyF_LzF; call abs_LxF_LyF_LzF
    end select
  end subroutine

  ! para_mod's current time (mod_para.F90:96-101): the library derives each
  ! stage's source time from ntime and the stage itself
  subroutine set_cur_time(ntime, incr)
    integer, intent(in) :: ntime
    real(c_float), intent(in) :: incr
  end subroutine
end program fw_drive
