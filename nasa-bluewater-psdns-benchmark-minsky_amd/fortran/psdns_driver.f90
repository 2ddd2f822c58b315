! RK2 driver over the gfx950 C ABI, mirroring the reference time loop:
!   main.f:392-685      loop, shifts, per-step timing line in `log`
!   timestep_module.f   two substeps: nonlinear terms, step, update
!   step.f:64-85        dt = cfl / velmax (velmax reduced in FP64)
!   step.f:366-385      integrating factors xfac, yfac, zfac
!   shifts.f:40-178     phase-shift offsets (kshift rules)
!   ek_force.f          FEK forcing after the corrector
!   force.f, vfuo.f     EPFOR increments after step (timestep_module.f:263),
!                       white-noise case (tforce = 0) of the UO process
!   forcdt.f            energy input rate from the efkz tally
!
! usage: psdns_driver key=value ...
!   n=64 prec=8 nsteps=20 cfl=3.14159265 viscos=0.01 kshift=1 kf_shell=3
!   seed=122929298 ic=<file> out=<file> shifts=<file> log=<file>
!   inpen=<file> outpen=<file> rk=2|4
!   kforce=<r> epsfor=<r> kfor=<i> epfor_out=<file>   EPFOR forcing (kforce > 0)
! ic/out: raw uny(ny,num_al,3) complex at precision prec (stream access)
! inpen/outpen: pencil checkpoint files of the reference (incond.f:120-133
!   reads one with inpen, main.f:604-608 writes un with outpen at the end)
! shifts: raw gsh(3,2) per step, real(8), overriding the kshift rule
! epfor_out: the increments for(kfor,k2fo,k2fo,3) of every step, complex(8)
program psdns_driver
  use iso_c_binding
  use psdns_gpu
  implicit none

  integer :: n = 64, prec = 8, nsteps = 10, kshift = 1, kf_shell = 0, rk = 2, istep, i, k, nargs
  integer(c_long_long) :: seed = 122929298_c_long_long
  real(c_double) :: cfl = 3.14159265358979d0, viscos = 0.01d0, velmax, velmax2, dt, time
  real(c_double) :: gsh(3, 2), tke, eps, r3(3), stimes(6)
  character(len=512) :: arg, key, val, ic_file = '', out_file = '', shift_file = '', log_file = 'log'
  character(len=512) :: inpen_file = '', outpen_file = '', epfor_file = ''
  real(c_double) :: kforce = 0d0, epsfor = 0.1d0, erate
  integer(c_int) :: kfor_ep = 0, k2fo
  complex(c_double_complex), allocatable, target :: forc(:, :, :, :)
  real(c_double), allocatable :: efkz(:, :)
  type(c_ptr) :: fptr
  type(psdns_params) :: p
  type(c_ptr) :: ctx
  integer(c_int) :: num_al, max_al_x, kfor
  real(c_double), allocatable :: xfac(:), yfac(:), zfac(:), ekinf(:), ekout(:), ek(:), dk(:)
  complex(c_double_complex), allocatable, target :: u8(:, :, :)
  complex(c_float_complex), allocatable, target :: u4(:, :, :)
  integer(8) :: c0, c1, crate
  real(8) :: tcpu, tmin, tmax, tsum
  integer :: nt, shift_unit, seedsz
  integer, allocatable :: rseed(:)
  logical :: first

  nargs = command_argument_count()
  do i = 1, nargs
    call get_command_argument(i, arg)
    k = index(arg, '=')
    if (k == 0) cycle
    key = arg(1:k - 1)
    val = arg(k + 1:)
    select case (trim(key))
    case ('n'); read (val, *) n
    case ('prec'); read (val, *) prec
    case ('nsteps'); read (val, *) nsteps
    case ('cfl'); read (val, *) cfl
    case ('viscos'); read (val, *) viscos
    case ('kshift'); read (val, *) kshift
    case ('kf_shell'); read (val, *) kf_shell
    case ('seed'); read (val, *) seed
    case ('ic'); ic_file = val
    case ('out'); out_file = val
    case ('shifts'); shift_file = val
    case ('log'); log_file = val
    case ('inpen'); inpen_file = val
    case ('outpen'); outpen_file = val
    case ('rk'); read (val, *) rk
    case ('kforce'); read (val, *) kforce
    case ('epsfor'); read (val, *) epsfor
    case ('kfor'); read (val, *) kfor_ep
    case ('epfor_out'); epfor_file = val
    end select
  end do

  p%nx = n; p%ny = n; p%nz = n; p%nc = 0; p%prec = prec; p%kshift = kshift; p%mask_kind = 0
  p%viscos = viscos; p%pr = 1d0; p%grad = 0d0
  call psdns_check(psdns_create(p, ctx), 'psdns_create')
  call psdns_check(psdns_geometry(ctx, num_al, max_al_x), 'psdns_geometry')
  write (6, '(a,i0,a,i0,a,i0,a,i0)') 'psdns_driver: n=', n, ' prec=', prec, ' num_al=', num_al, &
      ' max_al_x=', max_al_x

  ! initial condition: pencil checkpoint, file (reference cylinder layout)
  ! or the device generator
  if (len_trim(inpen_file) > 0) then
    call psdns_check(psdns_inpen(ctx, trim(inpen_file)//c_null_char), 'psdns_inpen')
  else if (len_trim(ic_file) > 0) then
    if (prec == 8) then
      allocate (u8(n, num_al, 3))
      open (11, file=trim(ic_file), access='stream', form='unformatted', status='old')
      read (11) u8
      close (11)
      call psdns_check(psdns_put_state(ctx, c_loc(u8), 0), 'psdns_put_state')
    else
      allocate (u4(n, num_al, 3))
      open (11, file=trim(ic_file), access='stream', form='unformatted', status='old')
      read (11) u4
      close (11)
      call psdns_check(psdns_put_state(ctx, c_loc(u4), 0), 'psdns_put_state')
    end if
  else
    call psdns_check(psdns_init_field(ctx, seed, 3d0, 0.5d0), 'psdns_init_field')
  end if

  kfor = 0
  if (kf_shell > 0) then
    kfor = kf_shell + 1
    allocate (ekinf(kfor), ekout(kfor))
    ! target shell energies: those of the initial field (ek_init.f role)
    call psdns_check(psdns_ek_shell(ctx, kfor, ekinf), 'psdns_ek_shell')
  end if

  ! EPFOR: kfor = kforce / min(beta) + 1.5 (param_set.f:33) unless given
  fptr = c_null_ptr
  if (kforce > 0d0) then
    if (kfor_ep == 0) kfor_ep = int(kforce + 1.5d0)
    k2fo = 2 * kfor_ep - 1
    allocate (forc(kfor_ep, k2fo, k2fo, 3), efkz(n / 2, 3))
    call psdns_check(psdns_set_epfor(ctx, kfor_ep, kforce), 'psdns_set_epfor')
    fptr = c_loc(forc)
    if (len_trim(epfor_file) > 0) &
      open (14, file=trim(epfor_file), access='stream', form='unformatted', status='replace')
  end if

  allocate (xfac(n / 2), yfac(n), zfac(n))
  shift_unit = 0
  if (len_trim(shift_file) > 0) then
    shift_unit = 12
    open (shift_unit, file=trim(shift_file), access='stream', form='unformatted', status='old')
  end if
  call random_seed(size=seedsz)
  allocate (rseed(seedsz))
  rseed = int(seed) + 12345
  call random_seed(put=rseed)
  gsh = 0d0
  if (kshift == 1 .or. kshift == 3) then
    call random_number(r3)
    gsh(:, 2) = r3
  end if

  open (7, file=trim(log_file), status='replace')
  call psdns_check(psdns_stage_times(ctx, 1, stimes), 'psdns_stage_times')
  time = 0d0
  tmin = huge(1d0); tmax = 0d0; tsum = 0d0; nt = 0
  call system_clock(count_rate=crate)
  do istep = 1, nsteps
    call system_clock(c0)
    ! shifts.f:65-84
    if (shift_unit /= 0) then
      read (shift_unit) gsh
    else
      call random_number(r3)
      do i = 1, 3
        select case (kshift)
        case (1)
          gsh(i, 1) = mod(gsh(i, 2) + 0.5d0, 1d0); gsh(i, 2) = r3(i)
        case (2)
          gsh(i, 1) = r3(i); gsh(i, 2) = mod(gsh(i, 1) + 0.5d0, 1d0)
        case (3)
          gsh(i, 1) = mod(gsh(i, 2) + 0.5d0, 1d0); gsh(i, 2) = mod(gsh(i, 1) + 0.5d0, 1d0)
        case default
          gsh(i, 1) = 0d0; gsh(i, 2) = 0d0
        end select
      end do
    end if
    call psdns_check(psdns_set_shifts(ctx, gsh), 'psdns_set_shifts')

    ! substep 1: nonlinear terms, step (dt, integrating factors), predictor
    call psdns_check(psdns_stage_nonlinear(ctx, 1, velmax), 'psdns_stage_nonlinear(1)')
    if (velmax < 1d-3) then
      write (6, *) 'velmax gotten zero at istep=', istep, velmax
      error stop 1
    end if
    dt = cfl / velmax
    if (dt > 10d0 .or. dt < 1d-7) then
      write (6, *) 'dt gotten out of control in step, dt=', dt
      error stop 1
    end if
    do i = 1, n / 2
      xfac(i) = exp(-viscos * dble(i - 1)**2 * dt)
    end do
    do i = 1, n
      k = i - 1
      if (i > n / 2 + 1) k = i - 1 - n
      yfac(i) = exp(-viscos * dble(k)**2 * dt)
      zfac(i) = yfac(i)
    end do
    ! force (timestep_module.f:263): this step's EPFOR increments
    if (kforce > 0d0) then
      call epfor_increments(forc, kfor_ep, kforce, epsfor, dt)
      if (len_trim(epfor_file) > 0) write (14) forc
    end if
    if (rk == 4) then
      ! rkmethod = 4 (timestep_module.f:293-310): advanc_1 .. advanc_4
      if (kforce > 0d0) call psdns_check(psdns_rk4_forcing(ctx, fptr), 'psdns_rk4_forcing')
      call psdns_check(psdns_rk4_update(ctx, 1, dt, xfac, yfac, zfac), 'psdns_rk4_update(1)')
      do k = 2, 4
        call psdns_check(psdns_stage_nonlinear(ctx, k, velmax2), 'psdns_stage_nonlinear(rk4)')
        call psdns_check(psdns_rk4_update(ctx, k, dt, xfac, yfac, zfac), 'psdns_rk4_update')
      end do
    else
      call psdns_check(psdns_stage_update(ctx, 1, dt, xfac, yfac, zfac, fptr), 'psdns_stage_update(1)')

      ! substep 2: corrector
      call psdns_check(psdns_stage_nonlinear(ctx, 2, velmax2), 'psdns_stage_nonlinear(2)')
      call psdns_check(psdns_stage_update(ctx, 2, dt, xfac, yfac, zfac, fptr), 'psdns_stage_update(2)')
    end if
    ! forcdt.f:18-45: energy input rate of the forcing (efkz / dt)
    if (kforce > 0d0) then
      call psdns_check(psdns_epfor_tally(ctx, efkz), 'psdns_epfor_tally')
      erate = 0.5d0 * sum(efkz) / dt
      write (6, '(a,i6,1p,e22.14)') ' forcing input: istep, erate =', istep, erate
    end if
    if (kfor > 0) call psdns_check(psdns_ek_force(ctx, kf_shell, kfor, ekinf, ekout), 'psdns_ek_force')
    time = time + dt

    call system_clock(c1)
    tcpu = dble(c1 - c0) / dble(crate)
    ! main.f:651-657 (one task: min = ave = max)
    write (7, '(a,i7,a,3f11.3,a)') 'cpu at step', istep, ' :min/ave/max=', tcpu, tcpu, tcpu, '  secs.'
    write (6, '(a,i6,1p,4e13.5)') ' istep,dt,time,cfl,velmax=', istep, dt, time, cfl, velmax
    if (istep > 1) then
      tmin = min(tmin, tcpu); tmax = max(tmax, tcpu); tsum = tsum + tcpu; nt = nt + 1
    end if
  end do
  close (7)
  if (shift_unit /= 0) close (shift_unit)
  if (kforce > 0d0 .and. len_trim(epfor_file) > 0) close (14)

  allocate (ek(n), dk(n))
  call psdns_check(psdns_shell_stats(ctx, ek, dk, n, tke, eps), 'psdns_shell_stats')
  write (6, '(a,1p,2e22.14)') 'final tke, epsilon = ', tke, eps
  if (nt > 0) write (6, '(a,3f12.6)') 's/step min/mean/max (steps 2..n) = ', tmin, tsum / nt, tmax
  call psdns_check(psdns_stage_times(ctx, 0, stimes), 'psdns_stage_times')
  write (6, '(a,6f10.3)') 'pass ms (y-inv z-inv x z-fwd y-fwd+upd total) = ', stimes

  if (len_trim(out_file) > 0) then
    open (13, file=trim(out_file), access='stream', form='unformatted', status='replace')
    if (prec == 8) then
      if (.not. allocated(u8)) allocate (u8(n, num_al, 3))
      call psdns_check(psdns_get_state(ctx, c_loc(u8), 0), 'psdns_get_state')
      write (13) u8
    else
      if (.not. allocated(u4)) allocate (u4(n, num_al, 3))
      call psdns_check(psdns_get_state(ctx, c_loc(u4), 0), 'psdns_get_state')
      write (13) u4
    end if
    close (13)
  end if
  if (len_trim(outpen_file) > 0) &
    call psdns_check(psdns_outpen(ctx, trim(outpen_file)//c_null_char, 8), 'psdns_outpen')
  call psdns_check(psdns_destroy(ctx), 'psdns_destroy')
  first = .true.

contains

  ! force.f:70-160 with vfuo.f at tforce = 0 (white noise: the velocity
  ! increment over dt has variance 2 epsfor dt per real component): alpha
  ! normal to k and z, beta normal to alpha and k; zero outside kforce and
  ! at k = 0; conjugate symmetry on kx = 0.  Gaussians by Box-Muller.
  subroutine epfor_increments(f, kfor, kforce, epsfor, dt)
    integer(c_int), intent(in) :: kfor
    complex(c_double_complex), intent(out) :: f(kfor, 2 * kfor - 1, 2 * kfor - 1, 3)
    real(c_double), intent(in) :: kforce, epsfor, dt
    integer :: x, y, z, i, k2f
    real(8) :: r(4), amp, k1u, k2u, k3u, k, k12
    complex(8) :: alpha, beta
    k2f = 2 * kfor - 1
    amp = sqrt(2d0 * epsfor * dt)
    f = (0d0, 0d0)
    do z = 1, k2f
      k3u = dble(z - 1)
      if (z > kfor) k3u = dble(z - 2 * kfor)
      do y = 1, k2f
        k2u = dble(y - 1)
        if (y > kfor) k2u = dble(y - 2 * kfor)
        do x = 1, kfor
          call random_number(r)
          r(1) = 1d0 - r(1); r(3) = 1d0 - r(3)
          alpha = amp * sqrt(-2d0 * log(r(1))) * cmplx(cos(6.283185307179586d0 * r(2)), &
                                                        sin(6.283185307179586d0 * r(2)), kind=8)
          beta = amp * sqrt(-2d0 * log(r(3))) * cmplx(cos(6.283185307179586d0 * r(4)), &
                                                       sin(6.283185307179586d0 * r(4)), kind=8)
          k1u = dble(x - 1)
          k = sqrt(k1u**2 + k2u**2 + k3u**2)
          k12 = sqrt(k1u**2 + k2u**2)
          if (k > kforce .or. k == 0d0) cycle
          if (k12 == 0d0) then
            f(x, y, z, 1) = alpha
            f(x, y, z, 2) = beta
          else
            f(x, y, z, 1) = (alpha * k * k2u + beta * k1u * k3u) / (k * k12)
            f(x, y, z, 2) = (beta * k2u * k3u - alpha * k * k1u) / (k * k12)
            f(x, y, z, 3) = -beta * k12 / k
          end if
        end do
      end do
    end do
    do i = 1, 3
      do y = kfor + 1, k2f
        f(1, y, 1, i) = conjg(f(1, k2f + 2 - y, 1, i))
      end do
      do z = kfor + 1, k2f
        f(1, 1, z, i) = conjg(f(1, 1, k2f + 2 - z, i))
        do y = 2, k2f
          f(1, y, z, i) = conjg(f(1, k2f + 2 - y, k2f + 2 - z, i))
        end do
      end do
    end do
  end subroutine epfor_increments
end program psdns_driver
