! ISO_C_BINDING interface to libfwant.so (include/fwant.h) for Fortran
! drivers of the FWANT kind (code/srcF/seis3d_wave_8op_p3.F90).  Arrays are
! passed as they are stored by the reference: column-major over the ghosted
! index space, 1-based local grid indices in source and receiver tables.
! Every function returns FW_OK (0) or a negative status; fw_check turns a
! nonzero status into the reference's error behaviour (message, stop 1).
module fwant_c
  use iso_c_binding
  implicit none
  private
  public :: fw_config, fw_check, fw_error_message
  public :: fw_create, fw_destroy, fw_last_error, fw_dims, fw_set_grid, fw_build_metric, fw_build_metric3d, fw_set_media
  public :: fw_set_force, fw_set_moment, fw_set_receivers, fw_add_snapshot, fw_init, fw_get_pml_profiles
  public :: fw_step, fw_stage, fw_run, fw_reference_steps, fw_sync, fw_get_receivers, fw_get_snapshot
  public :: fw_get_wavefield, fw_set_wavefield, fw_enable_timing, fw_get_timing

  integer(c_int), parameter, public :: FW_OK = 0
  integer(c_int), parameter, public :: FW_VX = 0, FW_VY = 1, FW_VZ = 2, FW_TXX = 3, FW_TYY = 4, &
                                       FW_TZZ = 5, FW_TXY = 6, FW_TXZ = 7, FW_TYZ = 8

  ! SeisFD3D.conf keys that shape the solve (struct fw_config)
  type, bind(C) :: fw_config
    integer(c_int) :: ni, nj, nk
    real(c_float) :: stept
    integer(c_int) :: freenode
    integer(c_int) :: abs_number(6)
    real(c_float) :: abs_velocity(6)
    real(c_float) :: cfs_bmax(6)
    real(c_float) :: cfs_amax(6)
  end type fw_config

  interface
    integer(c_int) function fw_create(cfg, h) bind(C, name="fw_create")
      import :: c_int, c_ptr, fw_config
      type(fw_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function fw_destroy(h) bind(C, name="fw_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function fw_last_error(h) bind(C, name="fw_last_error")
      import :: c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fw_dims(h, nx, ny, nz, freenode) bind(C, name="fw_dims")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(out) :: nx, ny, nz, freenode
    end function
    integer(c_int) function fw_set_grid(h, x, y, z, xsin, xcot, xi_x, eta_y, zeta_z) bind(C, name="fw_set_grid")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(in) :: x(*), y(*), z(*), xsin(*), xcot(*), xi_x(*), eta_y(*), zeta_z(*)
    end function
    integer(c_int) function fw_build_metric(h, x, y, z, xsin, xcot, xi_x, eta_y, zeta_z) &
        bind(C, name="fw_build_metric")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*), y(*), z(*)
      real(c_float), intent(out) :: xsin(*), xcot(*), xi_x(*), eta_y(*), zeta_z(*)
    end function
    ! mode C (terrain-following grid): the radius r(nx*ny*nz) of every node;
    ! the outputs p, q, zeta_r, 1/r are optional (c_null_ptr or c_loc of an
    ! nx*ny*nz real(c_float) array)
    integer(c_int) function fw_build_metric3d(h, x, y, r, p, q, zeta_r, rr) bind(C, name="fw_build_metric3d")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: x(*), y(*), r(*)
      type(c_ptr), value :: p, q, zeta_r, rr
    end function
    integer(c_int) function fw_set_media(h, lam, mu, rho) bind(C, name="fw_set_media")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(in) :: lam(*), mu(*), rho(*)
    end function
    integer(c_int) function fw_set_force(h, n, ntw, stf_id, indx, shift, t0, stf_time, stf_freq, fx, fy, fz) &
        bind(C, name="fw_set_force")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: n, ntw, stf_id
      integer(c_int), intent(in) :: indx(*)
      real(c_float), intent(in) :: shift(*), t0(*), stf_time(*), stf_freq(*), fx(*), fy(*), fz(*)
    end function
    integer(c_int) function fw_set_moment(h, n, ntw, stf_id, indx, shift, t0, stf_time, stf_freq, m6) &
        bind(C, name="fw_set_moment")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: n, ntw, stf_id
      integer(c_int), intent(in) :: indx(*)
      real(c_float), intent(in) :: shift(*), t0(*), stf_time(*), stf_freq(*), m6(*)
    end function
    integer(c_int) function fw_set_receivers(h, n, tinv, indx, max_records) bind(C, name="fw_set_receivers")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: n, tinv, max_records
      integer(c_int), intent(in) :: indx(*)
    end function
    integer(c_int) function fw_add_snapshot(h, subs, subc, subt, tinv, field_mask, max_records, snap_id) &
        bind(C, name="fw_add_snapshot")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: subs(3), subc(3), subt(3)
      integer(c_int), value :: tinv, field_mask, max_records
      integer(c_int), intent(out) :: snap_id
    end function
    integer(c_int) function fw_init(h) bind(C, name="fw_init")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fw_get_pml_profiles(h, a, b, d) bind(C, name="fw_get_pml_profiles")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(out) :: a(*), b(*), d(*)
    end function
    integer(c_int) function fw_step(h, ntime) bind(C, name="fw_step")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: ntime
    end function
    integer(c_int) function fw_stage(h, ntime, stage, xf, yf, zf) bind(C, name="fw_stage")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: ntime, stage, xf, yf, zf
    end function
    integer(c_int) function fw_run(h, ntime0, nsteps) bind(C, name="fw_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: ntime0, nsteps
    end function
    integer(c_int) function fw_reference_steps(nt) bind(C, name="fw_reference_steps")
      import :: c_int
      integer(c_int), value :: nt
    end function
    integer(c_int) function fw_sync(h) bind(C, name="fw_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function fw_get_receivers(h, out, nrecords) bind(C, name="fw_get_receivers")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      real(c_float), intent(out) :: out(*)
      integer(c_int), intent(out) :: nrecords
    end function
    integer(c_int) function fw_get_snapshot(h, snap_id, out, nrecords) bind(C, name="fw_get_snapshot")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: snap_id
      real(c_float), intent(out) :: out(*)
      integer(c_int), intent(out) :: nrecords
    end function
    integer(c_int) function fw_get_wavefield(h, field, out) bind(C, name="fw_get_wavefield")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: field
      real(c_float), intent(out) :: out(*)
    end function
    integer(c_int) function fw_set_wavefield(h, field, in) bind(C, name="fw_set_wavefield")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: h
      integer(c_int), value :: field
      real(c_float), intent(in) :: in(*)
    end function
    integer(c_int) function fw_enable_timing(h, on) bind(C, name="fw_enable_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: on
    end function
    integer(c_int) function fw_get_timing(h, stage_ms, total_ms, nstage) bind(C, name="fw_get_timing")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: stage_ms, total_ms
      integer(c_int), intent(out) :: nstage
    end function
    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! message of fw_last_error(h) as a Fortran string
  function fw_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = fw_last_error(h)
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

  ! the reference stops the run on any failure (mod_macdrp.F90:162-165)
  subroutine fw_check(rc, h, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: h
    character(len=*), intent(in) :: what
    if (rc /= FW_OK) then
      write(*, '(a, a, i0, a, a)') what, ': status ', rc, ': ', fw_error_message(h)
      stop 1
    end if
  end subroutine

end module fwant_c
