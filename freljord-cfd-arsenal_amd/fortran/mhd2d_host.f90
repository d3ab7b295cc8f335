! mhd2d_host.f90 -- Fortran host of the gfx950 MHD2D library.
!
! The reference (GDF_Balsara_P2_0316) is a set of argument-less external
! subroutines that share their state through COMMON blocks (com.txt:19-54).
! This file keeps that shape for a Fortran caller:
!
!   module mhd2d_capi    iso_c_binding interfaces of include/mhd2d.h
!   module mhd2d_common  the host-side state a caller of the reference
!                        reads: uh, Bx, By in the reference bounds, t, dt,
!                        tend, umax, totaldiv, pmin, count, plus the
!                        opaque device context
!   external subroutines set_bc, Lh, LhEz, TVB_Limiter_P2, pp_Limiter,
!                        KXRCF_Detector, div_free_Balsara, calculate_dt,
!                        calculate_umax, calculate_totaldiv, calculate_pmin,
!                        init_data, RK3, save_solution, save_XY
!
! Every subroutine body is one call into the C ABI on the context; the
! numerical state lives on the GPU between calls.  uh/Bx/By on the host are
! only filled by gpu_download (and pushed by gpu_upload), the explicit
! host<->device transfers that replace direct COMMON access.
! Errors: every status is checked; a failing call prints the library's
! message and stops the program (the reference has no error path).

module mhd2d_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: MHD2D_LF = 1, MHD2D_HLL = 2, MHD2D_HLLC = 3, MHD2D_HLLD = 4
  integer(c_int32_t), parameter :: MHD2D_IND_NONE = 0, MHD2D_IND_ALL = 1, MHD2D_IND_KXRCF = 2, &
                                   MHD2D_IND_MLP = 3
  integer(c_int32_t), parameter :: MHD2D_FP64 = 0, MHD2D_MIXED = 1
  integer(c_int32_t), parameter :: MHD2D_IC_VORTEX = 1, MHD2D_IC_ORSZAG_TANG = 3, MHD2D_IC_ROTOR = 4, &
                                   MHD2D_IC_BLAST = 5, MHD2D_IC_CLOUD_SHOCK = 6, MHD2D_IC_CLOUD_SHOCK_2 = 61, &
                                   MHD2D_IC_MAGNETIC_LOOP = 7, MHD2D_IC_BRIO_WU = 8, MHD2D_IC_BRIO_WU_Y = 81, &
                                   MHD2D_IC_RYU_JONES = 9, MHD2D_IC_RYU_JONES_Y = 91, &
                                   MHD2D_IC_OT_PERTURBED = 103, MHD2D_IC_RIEMANN4 = 104

  ! struct mhd2d_ic_info (include/mhd2d.h)
  type, bind(C) :: mhd2d_ic_info
    real(c_double) :: xa, xb, ya, yb, tend, M
    integer(c_int32_t) :: bcR, bcL, bcU, bcD, shock, pad
  end type mhd2d_ic_info

  ! struct mhd2d_config (include/mhd2d.h)
  type, bind(C) :: mhd2d_config
    integer(c_int32_t) :: Nx, Ny, flux_type, bcR, bcL, bcU, bcD, pp, indicator, precision, reserved
    real(c_double) :: xa, xb, ya, yb, M, CFL, Ck, mlp_lo, mlp_hi
  end type mhd2d_config

  interface
    integer(c_int) function mhd2d_create(cfg, ctx) bind(C, name='mhd2d_create')
      import :: c_int, c_ptr, mhd2d_config
      type(mhd2d_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function mhd2d_destroy(ctx) bind(C, name='mhd2d_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function mhd2d_last_error(ctx) bind(C, name='mhd2d_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_sync(ctx) bind(C, name='mhd2d_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_upload(ctx, uh, Bx, By) bind(C, name='mhd2d_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: uh(*), Bx(*), By(*)
    end function
    integer(c_int) function mhd2d_download(ctx, uh, Bx, By) bind(C, name='mhd2d_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: uh(*), Bx(*), By(*)
    end function
    integer(c_int) function mhd2d_init_data(ctx, ic, extra, nextra) bind(C, name='mhd2d_init_data')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: ic, nextra
      real(c_double), intent(in) :: extra(*)
    end function
    integer(c_int) function mhd2d_set_bc(ctx) bind(C, name='mhd2d_set_bc')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_lh(ctx) bind(C, name='mhd2d_lh')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_lhez(ctx) bind(C, name='mhd2d_lhez')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_rk_stage(ctx, stage, dt) bind(C, name='mhd2d_rk_stage')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: stage
      real(c_double), value :: dt
    end function
    integer(c_int) function mhd2d_tvb_limiter_p2(ctx) bind(C, name='mhd2d_tvb_limiter_p2')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_pp_limiter(ctx) bind(C, name='mhd2d_pp_limiter')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_kxrcf_detector(ctx, nflag) bind(C, name='mhd2d_kxrcf_detector')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: nflag
    end function
    integer(c_int) function mhd2d_div_free_balsara(ctx) bind(C, name='mhd2d_div_free_balsara')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_calculate_dt(ctx, dt) bind(C, name='mhd2d_calculate_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function mhd2d_diagnostics(ctx, umax, totaldiv, pmin) bind(C, name='mhd2d_diagnostics')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: umax, totaldiv, pmin
    end function
    integer(c_int) function mhd2d_rk3_prologue(ctx) bind(C, name='mhd2d_rk3_prologue')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_step(ctx, tend, t_out) bind(C, name='mhd2d_step')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: tend
      real(c_double), intent(out) :: t_out
    end function
    integer(c_int) function mhd2d_step_graph(ctx, tend, nsteps, t_out) bind(C, name='mhd2d_step_graph')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: tend
      integer(c_int64_t), value :: nsteps
      real(c_double), intent(out) :: t_out
    end function
    integer(c_int) function mhd2d_get_time(ctx, t, dt, count) bind(C, name='mhd2d_get_time')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: t, dt
      integer(c_int64_t), intent(out) :: count
    end function
    integer(c_int) function mhd2d_set_time(ctx, t, count) bind(C, name='mhd2d_set_time')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      integer(c_int64_t), value :: count
    end function
    integer(c_int) function mhd2d_ic_defaults(ic, info) bind(C, name='mhd2d_ic_defaults')
      import :: c_int, c_int32_t, mhd2d_ic_info
      integer(c_int32_t), value :: ic
      type(mhd2d_ic_info), intent(out) :: info
    end function
    integer(c_int) function mhd2d_calculate_az(ctx) bind(C, name='mhd2d_calculate_az')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function mhd2d_calculate_l2error(ctx, t20, L2) bind(C, name='mhd2d_calculate_l2error')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: t20
      real(c_double), intent(out) :: L2(8)
    end function
    integer(c_int) function mhd2d_checkpoint_save(ctx, path) bind(C, name='mhd2d_checkpoint_save')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function
    integer(c_int) function mhd2d_checkpoint_load(ctx, path) bind(C, name='mhd2d_checkpoint_load')
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: path(*)
    end function
  end interface

contains

  ! copy the NUL-terminated message of mhd2d_last_error into a Fortran string
  function last_error(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    msg = ''
    p = mhd2d_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, s, [256])
    do n = 1, 256
      if (s(n) == c_null_char) exit
      msg(n:n) = s(n)
    end do
  end function last_error

end module mhd2d_capi


module mhd2d_common
  use mhd2d_capi
  implicit none

  type(c_ptr) :: ctx = c_null_ptr
  type(mhd2d_config) :: cfg
  integer :: Nx = 0, Ny = 0
  integer, parameter :: dimPk = 10, NumEq = 8, NumGLP = 5, kdeg = 2
  ! com.txt:22,48 bounds
  real(c_double), allocatable :: uh(:, :, :, :), Bx(:, :, :), By(:, :, :)
  ! com.txt:20 /time/ and the diagnostics of com.txt:22,43,54
  real(c_double) :: t = 0, dt = 0, tend = 0, umax = 0, totaldiv = 0, pmin = 0
  ! com.txt:20 t20 and com.txt:24 L2 (calculate_L2error)
  real(c_double) :: t20 = 0, L2(8) = 0
  integer(c_int64_t) :: count = 0, nflag = 0
  logical :: verbose = .true.

contains

  ! negative status: a HIP or usage error, stop.  Positive: the numerical
  ! flags (1 non-finite value, 2 collapsed dt, 4 negative pressure), which
  ! the reference would not notice; reported once per new combination.
  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    integer(c_int), save :: reported = 0
    if (rc < 0) then
      write(0, '(a,a,i0,a,a)') what, ': status ', rc, ': ', trim(last_error(ctx))
      error stop 1
    else if (rc > 0 .and. rc /= reported) then
      write(0, '(a,a,i0,a)') what, ': numerical flags ', rc, &
        ' (1 non-finite value, 2 dt below 1e-7, 4 negative pressure)'
      reported = rc
    end if
  end subroutine check

  ! create the device context and the host arrays in the reference bounds
  subroutine mhd2d_open(c)
    type(mhd2d_config), intent(in) :: c
    integer(c_int) :: rc
    cfg = c
    Nx = c%Nx
    Ny = c%Ny
    rc = mhd2d_create(cfg, ctx)
    if (rc /= 0) then
      write(0, '(a,i0)') 'mhd2d_create failed with status ', rc
      error stop 1
    end if
    allocate(uh(0:Nx + 1, 0:Ny + 1, dimPk, NumEq), Bx(0:Nx, 1:Ny, kdeg + 1), By(1:Nx, 0:Ny, kdeg + 1))
    uh = 0
    Bx = 0
    By = 0
  end subroutine mhd2d_open

  subroutine mhd2d_close()
    integer(c_int) :: rc
    if (c_associated(ctx)) rc = mhd2d_destroy(ctx)
    ctx = c_null_ptr
    if (allocated(uh)) deallocate(uh, Bx, By)
  end subroutine mhd2d_close

  subroutine gpu_upload()
    call check(mhd2d_upload(ctx, uh, Bx, By), 'mhd2d_upload')
  end subroutine gpu_upload

  subroutine gpu_download()
    call check(mhd2d_download(ctx, uh, Bx, By), 'mhd2d_download')
  end subroutine gpu_download

  ! /time/ from the device (t, dt and count live there during RK3)
  subroutine sync_time()
    call check(mhd2d_get_time(ctx, t, dt, count), 'mhd2d_get_time')
  end subroutine sync_time

end module mhd2d_common


! ---------------------------------------------------------------------------
! the reference's subroutines, one C-ABI call each

subroutine set_bc()                ! set_bc.f90:5-27
  use mhd2d_common
  call check(mhd2d_set_bc(ctx), 'set_bc')
end subroutine set_bc

subroutine Lh()                    ! Lh.f90:1-305 (du, face traces and fluxes)
  use mhd2d_common
  call check(mhd2d_lh(ctx), 'Lh')
end subroutine Lh

subroutine LhEz()                  ! LhEz.f90:1-96 (dEz1, dEz2; needs Lh first)
  use mhd2d_common
  call check(mhd2d_lhez(ctx), 'LhEz')
end subroutine LhEz

subroutine TVB_Limiter_P2()        ! TVB_Limiter_P2.f90:1-191 (+ pp_Limiter if cfg%pp)
  use mhd2d_common
  call check(mhd2d_tvb_limiter_p2(ctx), 'TVB_Limiter_P2')
end subroutine TVB_Limiter_P2

subroutine pp_Limiter()            ! pp_Limiter.f90:1-83
  use mhd2d_common
  call check(mhd2d_pp_limiter(ctx), 'pp_Limiter')
end subroutine pp_Limiter

subroutine KXRCF_Detector()        ! KXRCF_Detector.f90:1-108 (flags kept on the device)
  use mhd2d_common
  call check(mhd2d_kxrcf_detector(ctx, nflag), 'KXRCF_Detector')
end subroutine KXRCF_Detector

subroutine div_free_Balsara()      ! div_free_Balsara.f90:1-79
  use mhd2d_common
  call check(mhd2d_div_free_balsara(ctx), 'div_free_Balsara')
end subroutine div_free_Balsara

subroutine calculate_dt()          ! calculate_dt.f90:5-27
  use mhd2d_common
  call check(mhd2d_calculate_dt(ctx, dt), 'calculate_dt')
end subroutine calculate_dt

subroutine calculate_umax()        ! calculate_umax.f90:5-13
  use mhd2d_common
  real(c_double) :: a, b
  call check(mhd2d_diagnostics(ctx, umax, a, b), 'calculate_umax')
end subroutine calculate_umax

subroutine calculate_totaldiv()    ! calculate_totaldiv.f90:5-23
  use mhd2d_common
  real(c_double) :: a, b
  call check(mhd2d_diagnostics(ctx, a, totaldiv, b), 'calculate_totaldiv')
end subroutine calculate_totaldiv

subroutine calculate_pmin()        ! calculate_pmin.f90:5-26
  use mhd2d_common
  real(c_double) :: a, b
  call check(mhd2d_diagnostics(ctx, a, b, pmin), 'calculate_pmin')
end subroutine calculate_pmin

! init_data.f90:21-121 on the device for the ICs of include/mhd2d.h
subroutine init_data(ic, extra, nextra)
  use mhd2d_common
  integer, intent(in) :: ic, nextra
  real(c_double), intent(in) :: extra(*)
  call check(mhd2d_init_data(ctx, int(ic, c_int32_t), extra, int(nextra, c_int32_t)), 'init_data')
end subroutine init_data

! calculate_Az.f90:1-26 (magnetic loop: mode 1 of Bx, By from Az)
subroutine calculate_Az()
  use mhd2d_common
  call check(mhd2d_calculate_az(ctx), 'calculate_Az')
end subroutine calculate_Az

! calculate_L2error.f90:1-37: L2 against the vortex translated by t20
subroutine calculate_L2error()
  use mhd2d_common
  call check(mhd2d_calculate_l2error(ctx, t20, L2), 'calculate_L2error')
  if (verbose) print *, L2
end subroutine calculate_L2error

! binary checkpoint incl. the face fields and (t, count) (include/mhd2d.h)
subroutine save_checkpoint(path)
  use mhd2d_common
  character(len=*), intent(in) :: path
  call check(mhd2d_checkpoint_save(ctx, trim(path) // c_null_char), 'save_checkpoint')
end subroutine save_checkpoint

subroutine load_checkpoint(path)
  use mhd2d_common
  character(len=*), intent(in) :: path
  call check(mhd2d_checkpoint_load(ctx, trim(path) // c_null_char), 'load_checkpoint')
  call sync_time()
end subroutine load_checkpoint

! RK3.f90:5-174 without frame output.  t, dt and count stay on the device
! inside the loop (the step clips dt to tend, and a step that starts at
! t >= tend changes nothing, exactly as the reference's loop test stops
! it).  The steps run as HIP graph launches, 16 per read of t; with verbose
! one per read, and the per-step line of RK3.f90:150 is printed (it costs
! one extra reduction per step).
subroutine RK3()
  use mhd2d_common
  integer(c_int64_t) :: chunk
  chunk = 16
  if (verbose) chunk = 1
  call check(mhd2d_rk3_prologue(ctx), 'RK3 prologue')
  call calculate_umax()
  call calculate_totaldiv()
  call calculate_pmin()
  t = 0
  count = 0
  if (verbose) print *, t, "  ", umax, "  ", totaldiv
  do while (t < tend)
    call check(mhd2d_step_graph(ctx, tend, chunk, t), 'RK3 step')
    if (verbose) then
      call check(mhd2d_diagnostics(ctx, umax, totaldiv, pmin), 'RK3 diagnostics')
      print *, t, "  ", umax, "  ", totaldiv
    end if
  end do
  call sync_time()
end subroutine RK3

! RK3 written with the reference's own call sequence (RK3.f90:60-149):
! every whole-array statement of the reference is one mhd2d_rk_stage call.
subroutine RK3_subroutines(max_steps)
  use mhd2d_common
  integer, intent(in) :: max_steps
  integer :: n
  t = 0
  count = 0
  call check(mhd2d_set_time(ctx, 0d0, 0_c_int64_t), 'set_time')
  call set_bc()
  call TVB_Limiter_P2()
  call div_free_Balsara()
  call calculate_umax()
  call calculate_totaldiv()
  call calculate_pmin()
  if (verbose) print *, t, "  ", umax, "  ", totaldiv
  n = 0
  do while (t < tend .and. (max_steps <= 0 .or. n < max_steps))
    call calculate_dt()
    if (t + dt >= tend) then
      dt = tend - t
      t = tend
    else
      t = t + dt
    end if
    if (dt < 1e-7) t = tend
    ! stage 1
    call set_bc()
    call Lh()
    call LhEz()
    call check(mhd2d_rk_stage(ctx, 1_c_int32_t, dt), 'stage 1')
    call TVB_Limiter_P2()
    call div_free_Balsara()
    call set_bc()
    ! stage 2
    call Lh()
    call LhEz()
    call check(mhd2d_rk_stage(ctx, 2_c_int32_t, dt), 'stage 2')
    call TVB_Limiter_P2()
    call div_free_Balsara()
    ! stage 3
    call set_bc()
    call Lh()
    call LhEz()
    call check(mhd2d_rk_stage(ctx, 3_c_int32_t, dt), 'stage 3')
    call TVB_Limiter_P2()
    call div_free_Balsara()
    count = count + 1
    n = n + 1
    if (verbose) then
      call calculate_umax()
      call calculate_totaldiv()
      call calculate_pmin()
      print *, t, "  ", umax, "  ", totaldiv
    end if
  end do
  call check(mhd2d_set_time(ctx, t, count), 'set_time')
end subroutine RK3_subroutines

! save_solution.f90:5-37: Q1..Q8.txt, one value per line, d -> j -> i order
subroutine save_solution()
  use mhd2d_common
  integer :: i, j, d, n
  character(len=8) :: fname
  call gpu_download()
  do n = 1, NumEq
    write(fname, '(a,i0,a)') 'Q', n, '.txt'
    open(unit=10 + n, file=trim(fname), status='replace')
  end do
  do d = 1, dimPk
    do j = 1, Ny
      do i = 1, Nx
        do n = 1, NumEq
          write(10 + n, *) uh(i, j, d, n)
        end do
      end do
    end do
  end do
  do n = 1, NumEq
    close(10 + n)
  end do
end subroutine save_solution

! save_XY.f90:5-30: Gauss-point coordinates, nodes and weights
subroutine save_XY()
  use mhd2d_common
  integer :: i, j, q
  real(c_double) :: hx, hy, lam(5), w(5)
  ! 5-point Gauss-Legendre nodes and weights on [-1, 1]
  lam = [-0.906179845938663992797627d0, -0.538469310105683091036314d0, 0d0, &
         0.538469310105683091036314d0, 0.906179845938663992797627d0]
  w = [0.236926885056189087514264d0, 0.478628670499366468041292d0, 0.568888888888888888888889d0, &
       0.478628670499366468041292d0, 0.236926885056189087514264d0]
  hx = (cfg%xb - cfg%xa) / Nx
  hy = (cfg%yb - cfg%ya) / Ny
  open(unit=11, file='Xc.txt', status='replace')
  open(unit=12, file='Yc.txt', status='replace')
  open(unit=13, file='lambda.txt', status='replace')
  open(unit=14, file='weight.txt', status='replace')
  do i = 1, Nx
    do q = 1, NumGLP
      write(11, *) cfg%xa + (i - 0.5d0) * hx + 0.5d0 * hx * lam(q)
    end do
  end do
  do j = 1, Ny
    do q = 1, NumGLP
      write(12, *) cfg%ya + (j - 0.5d0) * hy + 0.5d0 * hy * lam(q)
    end do
  end do
  do q = 1, NumGLP
    write(13, *) lam(q)
    write(14, *) w(q)
  end do
  close(11)
  close(12)
  close(13)
  close(14)
end subroutine save_XY
