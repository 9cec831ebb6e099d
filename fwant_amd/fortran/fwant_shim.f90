! Drop-in modules for the time loop of code/srcF/seis3d_wave_8op_p3.F90:
! macdrp_mod, abs_mod and src_mod with the routines its loop body calls and
! the reference's argument lists (mod_macdrp.F90:30-49, mod_abs_npml.F90:
! 37-51, mod_src.F90:207, 275, 373), so that the loop body runs as written
! with libfwant doing the work.  Linked in place of mod_macdrp.F90,
! mod_abs_npml.F90 and mod_src.F90 for the loop; set-up goes through the
! fw_* calls (fwant_c) and fwant_shim_attach.
!
! How the calls map onto the library: the sweep of one RK stage, its
! CFS-PML corrections, its point sources and its RK update run as one
! fw_stage launch.  The shim records what a stage's calls ask for -- the
! step (src_surface / src_stress / src_force pass ntime), the operator
! pattern (which macdrp_Lx?_Ly?_Lz? is called) -- and launches the stage
! at its macdrp_RK_* call; abs_syn, abs_L*, abs_RK_* and macdrp_syn do no
! work of their own (the CFS-PML corrections and auxiliary RK run inside
! the stage; the RK "m" register is the step's input set).  The routines
! check what they are given: the source times (ntime + tinc) of the stage,
! the classical RK4 coefficients, the abs_L* pattern matching the
! macdrp_L* one; a mismatch stops the run (the reference's error
! behaviour, mod_macdrp.F90:162-165).  The h arrays the loop passes to
! src_stress / src_force are the public hTxx .. hVz of this macdrp_mod:
! placeholders, the RHS lives in device memory.
module fwant_shim_state
  use iso_c_binding
  use fwant_c
  implicit none
  integer, parameter :: SP = c_float
  type(c_ptr), save :: fw_h = c_null_ptr
  integer, save :: st_next = 0      ! next RK stage of the step (0..3)
  integer, save :: st_ntime = -1    ! step of the stage being assembled
  integer, save :: st_pat = -1      ! its operator pattern (x << 2 | y << 1 | z, 1 = forward)
  integer, save :: st_abs = -1      ! the abs_L* pattern of the stage
  real(SP), parameter :: rk_a(3) = (/ 0.5_SP, 0.5_SP, 1.0_SP /)
  real(SP), parameter :: rk_b(4) = (/ 1.0_SP / 6.0_SP, 1.0_SP / 3.0_SP, 1.0_SP / 3.0_SP, 1.0_SP / 6.0_SP /)
contains
  subroutine shim_stop(what)
    character(len=*), intent(in) :: what
    write(*, '(a, a)') 'fwant shim: ', what
    stop 1
  end subroutine
  ! tinc of RK stage st: 0, a1, a2, a3
  real(SP) function stage_tinc(st)
    integer, intent(in) :: st
    if (st == 0) then
      stage_tinc = 0.0_SP
    else
      stage_tinc = rk_a(st)
    end if
  end function
  subroutine note_time(ntime, tinc)
    integer, intent(in) :: ntime
    real(SP), intent(in) :: tinc
    if (st_next == 0) st_ntime = ntime
    if (ntime /= st_ntime .or. tinc /= stage_tinc(st_next)) &
      call shim_stop('source time (ntime, tinc) does not belong to the current RK stage')
  end subroutine
  subroutine run_stage(st, rka, rkb)
    integer, intent(in) :: st
    real(SP), intent(in) :: rka, rkb
    if (st /= st_next) call shim_stop('RK stages out of order')
    if (st < 3) then
      if (rka /= rk_a(st + 1)) call shim_stop('RK a coefficient differs from firRKa (classical RK4)')
    end if
    if (rkb /= rk_b(st + 1)) call shim_stop('RK b coefficient differs from firRKb (classical RK4)')
    if (st_pat < 0) call shim_stop('no macdrp_Lx?_Ly?_Lz? call in this RK stage')
    if (st_abs >= 0 .and. st_abs /= st_pat) call shim_stop('abs_L* pattern differs from the macdrp_L* pattern')
    if (st_ntime < 0) call shim_stop('no src_* call gave the step of this RK stage')
    call fw_check(fw_stage(fw_h, int(st_ntime, c_int), int(st, c_int), int(iand(ishft(st_pat, -2), 1), c_int), &
                           int(iand(ishft(st_pat, -1), 1), c_int), int(iand(st_pat, 1), c_int)), fw_h, 'fw_stage')
    st_next = mod(st + 1, 4)
    st_pat = -1
    st_abs = -1
  end subroutine
end module fwant_shim_state

module macdrp_mod
  use iso_c_binding
  use fwant_shim_state
  implicit none
  private
  real(SP), allocatable, dimension(:,:,:), public :: hTxx, hTyy, hTxy, hVx, hVy, hTzz, hTxz, hTyz, hVz
  real(SP), public :: firRKa(3), firRKb(4)
  public :: fwant_shim_attach, macdrp_syn, macdrp_RK_beg, macdrp_RK_inn, macdrp_RK_fin
  public :: macdrp_LxF_LyF_LzF, macdrp_LxB_LyB_LzB, macdrp_LxB_LyB_LzF, macdrp_LxF_LyF_LzB, &
            macdrp_LxB_LyF_LzF, macdrp_LxF_LyB_LzB, macdrp_LxF_LyB_LzF, macdrp_LxB_LyF_LzB
contains
  ! the handle the loop's calls act on (set up with fw_create .. fw_init)
  subroutine fwant_shim_attach(h)
    type(c_ptr), intent(in) :: h
    fw_h = h
    st_next = 0; st_ntime = -1; st_pat = -1; st_abs = -1
    if (.not. allocated(hTxx)) allocate(hTxx(1,1,1), hTyy(1,1,1), hTxy(1,1,1), hVx(1,1,1), hVy(1,1,1), &
                                        hTzz(1,1,1), hTxz(1,1,1), hTyz(1,1,1), hVz(1,1,1))
    firRKa = rk_a
    firRKb = rk_b
  end subroutine
  ! m = W: the step's input set is the RK "m" register (mod_macdrp.F90:268-287)
  subroutine macdrp_syn
    if (st_next /= 0) call shim_stop('macdrp_syn inside a step')
    st_ntime = -1
  end subroutine
  subroutine set_pat(p)
    integer, intent(in) :: p
    if (st_pat >= 0) call shim_stop('two operator sweeps in one RK stage')
    st_pat = p
  end subroutine
  subroutine macdrp_LxF_LyF_LzF
    call set_pat(7)
  end subroutine
  subroutine macdrp_LxB_LyB_LzB
    call set_pat(0)
  end subroutine
  subroutine macdrp_LxB_LyB_LzF
    call set_pat(1)
  end subroutine
  subroutine macdrp_LxF_LyF_LzB
    call set_pat(6)
  end subroutine
  subroutine macdrp_LxB_LyF_LzF
    call set_pat(3)
  end subroutine
  subroutine macdrp_LxF_LyB_LzB
    call set_pat(4)
  end subroutine
  subroutine macdrp_LxF_LyB_LzF
    call set_pat(5)
  end subroutine
  subroutine macdrp_LxB_LyF_LzB
    call set_pat(2)
  end subroutine
  subroutine macdrp_RK_beg(rka, rkb)
    real(SP), intent(in) :: rka, rkb
    if (st_next /= 0) call shim_stop('macdrp_RK_beg must end the first RK stage')
    call run_stage(0, rka, rkb)
  end subroutine
  subroutine macdrp_RK_inn(rka, rkb)
    real(SP), intent(in) :: rka, rkb
    if (st_next /= 1 .and. st_next /= 2) call shim_stop('macdrp_RK_inn must end the second or third RK stage')
    call run_stage(st_next, rka, rkb)
  end subroutine
  subroutine macdrp_RK_fin(rkb)
    real(SP), intent(in) :: rkb
    if (st_next /= 3) call shim_stop('macdrp_RK_fin must end the fourth RK stage')
    call run_stage(3, 0.0_SP, rkb)
  end subroutine
end module macdrp_mod

module abs_mod
  use fwant_shim_state
  implicit none
  private
  public :: abs_syn, abs_RK_beg, abs_RK_inn, abs_RK_fin
  public :: abs_LxF_LyF_LzF, abs_LxB_LyB_LzB, abs_LxB_LyB_LzF, abs_LxF_LyF_LzB, &
            abs_LxB_LyF_LzF, abs_LxF_LyB_LzB, abs_LxF_LyB_LzF, abs_LxB_LyF_LzB
contains
  ! the CFS-PML corrections and auxiliary RK run inside the stage launch
  subroutine abs_syn
  end subroutine
  subroutine set_abs(p)
    integer, intent(in) :: p
    st_abs = p
  end subroutine
  subroutine abs_LxF_LyF_LzF
    call set_abs(7)
  end subroutine
  subroutine abs_LxB_LyB_LzB
    call set_abs(0)
  end subroutine
  subroutine abs_LxB_LyB_LzF
    call set_abs(1)
  end subroutine
  subroutine abs_LxF_LyF_LzB
    call set_abs(6)
  end subroutine
  subroutine abs_LxB_LyF_LzF
    call set_abs(3)
  end subroutine
  subroutine abs_LxF_LyB_LzB
    call set_abs(4)
  end subroutine
  subroutine abs_LxF_LyB_LzF
    call set_abs(5)
  end subroutine
  subroutine abs_LxB_LyF_LzB
    call set_abs(2)
  end subroutine
  subroutine abs_RK_beg(rka, rkb)
    real(SP), intent(in) :: rka, rkb
  end subroutine
  subroutine abs_RK_inn(rka, rkb)
    real(SP), intent(in) :: rka, rkb
  end subroutine
  subroutine abs_RK_fin(rkb)
    real(SP), intent(in) :: rkb
  end subroutine
end module abs_mod

module src_mod
  use fwant_shim_state
  implicit none
  private
  public :: src_surface, src_stress, src_force
contains
  ! the surface traction of surface forces (mod_src.F90:373-414): applied by
  ! the stage launch at time (ntime + tinc) stept
  subroutine src_surface(ntime, tinc, stept)
    integer, intent(in) :: ntime
    real(SP), intent(in) :: tinc, stept
    call note_time(ntime, tinc)
  end subroutine
  subroutine src_stress(Txx, Tyy, Tzz, Txy, Txz, Tyz, ntime, tinc, stept)
    real(SP), dimension(:,:,:), intent(inout) :: Txx, Tyy, Tzz, Txy, Txz, Tyz
    integer, intent(in) :: ntime
    real(SP), intent(in) :: tinc, stept
    call note_time(ntime, tinc)
  end subroutine
  subroutine src_force(Vx, Vy, Vz, ntime, tinc, stept)
    real(SP), dimension(:,:,:), intent(inout) :: Vx, Vy, Vz
    integer, intent(in) :: ntime
    real(SP), intent(in) :: tinc, stept
    call note_time(ntime, tinc)
  end subroutine
end module src_mod
