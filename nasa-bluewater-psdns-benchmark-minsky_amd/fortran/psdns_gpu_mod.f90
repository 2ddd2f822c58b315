! ISO_C_BINDING interface to the gfx950 PSDNS RK2 step (include/psdns_gpu.h).
!
! A Fortran caller of the reference (PSDNS/timestep_module.f rksubstep) uses
! this module in place of the FFTW path: phase1 + itransform_vel +
! realspace_vel + transform become psdns_stage_nonlinear, the kept `step`
! (step.f) computes dt and the integrating factors from the returned velmax,
! and wavespace_vel + proc4a/proc4b become psdns_stage_update.
module psdns_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PSDNS_FIELD_U = 0, PSDNS_FIELD_UN = 1, PSDNS_FIELD_UX = 2

  type, bind(C) :: psdns_params
    integer(c_int) :: nx, ny, nz, nc, prec, kshift, mask_kind
    real(c_double) :: viscos
    real(c_double) :: pr(8)
    real(c_double) :: grad(3, 8)
  end type psdns_params

  interface
    integer(c_int) function psdns_create(p, ctx) bind(C, name="psdns_create")
      import :: c_int, c_ptr, psdns_params
      type(psdns_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function psdns_create

    integer(c_int) function psdns_destroy(ctx) bind(C, name="psdns_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function psdns_destroy

    integer(c_int) function psdns_geometry(ctx, num_al, max_al_x) bind(C, name="psdns_geometry")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: num_al, max_al_x
    end function psdns_geometry

    integer(c_int) function psdns_put_state(ctx, uny, layout) bind(C, name="psdns_put_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: uny
      integer(c_int), value :: layout
    end function psdns_put_state

    integer(c_int) function psdns_get_state(ctx, uny, layout) bind(C, name="psdns_get_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: uny
      integer(c_int), value :: layout
    end function psdns_get_state

    integer(c_int) function psdns_set_shifts(ctx, gsh) bind(C, name="psdns_set_shifts")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: gsh(3, 2)
    end function psdns_set_shifts

    integer(c_int) function psdns_stage_nonlinear(ctx, kstep, velmax) bind(C, name="psdns_stage_nonlinear")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kstep
      real(c_double), intent(out) :: velmax
    end function psdns_stage_nonlinear

    integer(c_int) function psdns_stage_update(ctx, kstep, dt, xfac, yfac, zfac, forcing) &
        bind(C, name="psdns_stage_update")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kstep
      real(c_double), value :: dt
      real(c_double), intent(in) :: xfac(*), yfac(*), zfac(*)
      type(c_ptr), value :: forcing
    end function psdns_stage_update

    integer(c_int) function psdns_ek_shell(ctx, kfor, ek) bind(C, name="psdns_ek_shell")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kfor
      real(c_double), intent(out) :: ek(*)
    end function psdns_ek_shell

    integer(c_int) function psdns_ek_force(ctx, kf_shell, kfor, ekinf, ek_out) bind(C, name="psdns_ek_force")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kf_shell, kfor
      real(c_double), intent(in) :: ekinf(*)
      real(c_double), intent(out) :: ek_out(*)
    end function psdns_ek_force

    integer(c_int) function psdns_shell_stats(ctx, ek, dk, nshell, tke, eps) bind(C, name="psdns_shell_stats")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ek(*), dk(*)
      integer(c_int), value :: nshell
      real(c_double), intent(out) :: tke, eps
    end function psdns_shell_stats

    integer(c_int) function psdns_vg_moments(ctx, dudx, dwdz) bind(C, name="psdns_vg_moments")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dudx(8), dwdz(8)
    end function psdns_vg_moments

    ! RK4 substep kstep = 1..4 after psdns_stage_nonlinear(ctx, kstep, ...)
    ! (homogeneous/advanc_1.f .. advanc_4.f); factors read on kstep 1 only
    integer(c_int) function psdns_rk4_update(ctx, kstep, dt, xfac, yfac, zfac) bind(C, name="psdns_rk4_update")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kstep
      real(c_double), value :: dt
      real(c_double), intent(in) :: xfac(*), yfac(*), zfac(*)
    end function psdns_rk4_update

    ! EPFOR forcing (force.f increments for(kfor, k2fo, k2fo, 3), complex(8)):
    ! box and radius, the efkz(nxh, 3) tally of the last proc4a / advanc_1,
    ! and the increments of the RK4 substeps of a step
    integer(c_int) function psdns_set_epfor(ctx, kfor, kforce) bind(C, name="psdns_set_epfor")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: kfor
      real(c_double), value :: kforce
    end function psdns_set_epfor

    integer(c_int) function psdns_epfor_tally(ctx, efkz) bind(C, name="psdns_epfor_tally")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: efkz(*)
    end function psdns_epfor_tally

    integer(c_int) function psdns_rk4_forcing(ctx, forcing) bind(C, name="psdns_rk4_forcing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: forcing
    end function psdns_rk4_forcing

    ! pencil checkpoint of un (homogeneous/outpen_write.f:59-76, inpen_read.f);
    ! path is a NUL-terminated string: trim(name)//c_null_char
    integer(c_int) function psdns_outpen(ctx, path, file_prec) bind(C, name="psdns_outpen")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: file_prec
    end function psdns_outpen

    integer(c_int) function psdns_inpen(ctx, path) bind(C, name="psdns_inpen")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function psdns_inpen

    integer(c_int) function psdns_init_field(ctx, seed, kpeak, tke) bind(C, name="psdns_init_field")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      integer(c_long_long), value :: seed
      real(c_double), value :: kpeak, tke
    end function psdns_init_field

    integer(c_int) function psdns_stage_times(ctx, enable, times) bind(C, name="psdns_stage_times")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      real(c_double), intent(out) :: times(6)
    end function psdns_stage_times

    ! per-substep pass times (timestep_module.f tt(1..10) at pass granularity):
    ! times(p + 1, st + 1), p = 0..10 passes, st = 0..3 substeps
    integer(c_int) function psdns_pass_times(ctx, enable, times, ntimes) bind(C, name="psdns_pass_times")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable, ntimes
      real(c_double), intent(out) :: times(*)
    end function psdns_pass_times

    type(c_ptr) function psdns_last_error_c() bind(C, name="psdns_last_error")
      import :: c_ptr
    end function psdns_last_error_c

    integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! abrt.f-style failure: print the library's message and stop
  subroutine psdns_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: i, l
    character(len=512) :: text
    if (rc == 0) return
    p = psdns_last_error_c()
    l = int(c_strlen(p))
    call c_f_pointer(p, msg, [l])
    text = ' '
    do i = 1, min(l, 512)
      text(i:i) = msg(i)
    end do
    write (6, '(a,a,a,i0,a,a)') 'abort: ', where, ' rc=', rc, ': ', trim(text)
    error stop 1
  end subroutine psdns_check

end module psdns_gpu
