! mhd2d_run.f90 -- the reference's main program (main.f90:1-22) on the GPU.
!
!   mhd2d_run <namelist file>
!
! reads one &run namelist, projects the initial condition on the device
! (init_data), advances it with RK3 (the fused step, mode = 'fused', or the
! reference's call sequence subroutine by subroutine, mode = 'subroutines'),
! then writes what main.f90 writes (save_XY, save_solution) when qtxt is
! true and, when out is set, the final uh, Bx, By as raw little-endian
! float64 in the reference layouts, preceded by 4 float64: t, count, Nx, Ny.
! Without xa < xb the domain, boundaries and M come from the initial
! condition's initN.txt (mhd2d_ic_defaults).  restart = <file> continues from
! a checkpoint instead of projecting the initial condition (no prologue);
! checkpoint = <file> writes one at the end; l2err calls calculate_L2error
! with t20 = tend if tend <= 5, else 0 (main.f90:13-19).  frames sets the
! host's flash (1: Q1..Q8flash.txt and T.txt every tend/frame_max as
! RK3.f90:29-53,154-172 write them; 0: the t = 0 frame only; -1: none).

program mhd2d_run
  use mhd2d_common
  implicit none
  integer :: Nx_ = 32, Ny_ = 32, flux_type = 4, ic = 103, bc = 1, pp = 0, indicator = 1, precision = 0
  integer :: max_steps = 0, nextra = 0, u, frames = -1, frame_max = 200
  real(c_double) :: xa = 0, xb = 0, ya = 0, yb = 0, M = 10, CFL = 0.15d0, Ck = 1, mlp_lo = 0.2d0, mlp_hi = 0.8d0
  real(c_double) :: tend_ = 0.1d0, extra(64) = 0
  character(len=16) :: mode = 'fused'
  character(len=512) :: out = '', nml = '', restart = '', checkpoint = ''
  logical :: qtxt = .false., quiet = .false., l2err = .false.
  type(mhd2d_config) :: c
  type(mhd2d_ic_info) :: info
  namelist /run/ Nx_, Ny_, flux_type, ic, bc, pp, indicator, precision, max_steps, xa, xb, ya, yb, M, CFL, Ck, &
                 mlp_lo, mlp_hi, tend_, extra, nextra, mode, out, qtxt, quiet, restart, checkpoint, l2err, &
                 frames, frame_max
  interface
    subroutine init_data_ic(ic, extra, nextra)
      import :: c_double
      integer, intent(in) :: ic, nextra
      real(c_double), intent(in) :: extra(*)
    end subroutine
    subroutine RK3_subroutines(max_steps)
      integer, intent(in) :: max_steps
    end subroutine
    subroutine load_checkpoint(path)
      character(len=*), intent(in) :: path
    end subroutine
    subroutine save_checkpoint(path)
      character(len=*), intent(in) :: path
    end subroutine
  end interface

  call get_command_argument(1, nml)
  if (len_trim(nml) == 0) then
    write(0, '(a)') 'usage: mhd2d_run <namelist file with &run ... />'
    error stop 2
  end if
  open(newunit=u, file=trim(nml), status='old', action='read')
  read(u, nml=run)
  close(u)
  if (xb <= xa) then       ! the initial condition's own domain, boundaries and M
    if (mhd2d_ic_defaults(int(ic, c_int32_t), info) /= 0) then
      write(0, '(a,i0)') 'unknown initial condition ', ic
      error stop 2
    end if
    xa = info%xa
    xb = info%xb
    ya = info%ya
    yb = info%yb
    bc = info%bcR
    M = info%M
  end if

  c = mhd2d_config(Nx_, Ny_, flux_type, bc, bc, bc, bc, pp, indicator, precision, 0, &
                   xa, xb, ya, yb, M, CFL, Ck, mlp_lo, mlp_hi)
  call mhd2d_open(c)
  tend = tend_
  verbose = .not. quiet
  flash = frames
  frameMAX = frame_max

  if (len_trim(restart) > 0) then
    call load_checkpoint(restart)
    do while (t < tend .and. (max_steps <= 0 .or. count < max_steps))
      call check(mhd2d_step(ctx, tend, t), 'RK3 step')
      call sync_time()
    end do
  else
    call init_data_ic(ic, extra, nextra)
  end if
  if (len_trim(restart) > 0) then
    continue
  else if (trim(mode) == 'subroutines') then
    call RK3_subroutines(max_steps)
  else if (max_steps > 0) then
    ! fused step, bounded step count (same loop test as RK3)
    call check(mhd2d_rk3_prologue(ctx), 'RK3 prologue')
    t = 0
    do while (t < tend .and. count < max_steps)
      call check(mhd2d_step(ctx, tend, t), 'RK3 step')
      count = count + 1
    end do
    call sync_time()
  else
    call RK3()
  end if
  call calculate_umax()
  call calculate_totaldiv()
  call calculate_pmin()
  print '(a,es24.16,a,i0,a,es12.4,a,es12.4,a,es12.4)', 't = ', t, '  steps = ', count, '  umax = ', umax, &
        '  totaldiv = ', totaldiv, '  pmin = ', pmin

  if (qtxt) then
    call save_XY()
    call save_solution()
  end if
  if (l2err) then
    t20 = 0
    if (tend <= 5) t20 = tend
    call calculate_L2error()
  end if
  if (len_trim(checkpoint) > 0) call save_checkpoint(checkpoint)
  if (len_trim(out) > 0) then
    call gpu_download()
    open(newunit=u, file=trim(out), access='stream', form='unformatted', status='replace')
    write(u) t, real(count, c_double), real(Nx, c_double), real(Ny, c_double)
    write(u) uh, Bx, By
    close(u)
  end if
  call mhd2d_close()
end program mhd2d_run
