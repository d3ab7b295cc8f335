! mhd2d_host.f90 -- Fortran host of the gfx950 MHD2D library.
!
! The reference (GDF_Balsara_P2_0316) is a set of argument-less external
! subroutines that share their state through COMMON blocks (com.txt:19-54).
! This file keeps that shape for a Fortran caller:
!
!   module mhd2d_capi    iso_c_binding interfaces of include/mhd2d.h
!                        (mhd2d_capi.f90)
!   module mhd2d_common  the host-side state a caller of the reference
!                        reads: uh, Bx, By in the reference bounds, t, dt,
!                        tend, umax, totaldiv, pmin, count, plus the
!                        opaque device context
!   external subroutines set_bc, Lh, LhEz, TVB_Limiter_P2, pp_Limiter,
!                        KXRCF_Detector, div_free_Balsara, calculate_dt,
!                        calculate_umax, calculate_totaldiv, calculate_pmin,
!                        init_data, calculate_Ez, calculate_Az,
!                        calculate_L2error, RK3 (with the frame files of
!                        RK3.f90:29-53,154-172), save_solution, save_XY
!
! Every subroutine body is one call into the C ABI on the context; the
! numerical state lives on the GPU between calls.  uh/Bx/By on the host are
! only filled by gpu_download (and pushed by gpu_upload), the explicit
! host<->device transfers that replace direct COMMON access.
! Errors: every status is checked; a failing call prints the library's
! message and stops the program (the reference has no error path).



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
  ! the initial condition init_data projects: the reference picks it at
  ! compile time (include 'initN.txt', init_data.f90:19); here it is set
  ! before the call (ids and parameters of mhd2d_init_data in mhd2d.h)
  integer :: ic_select = MHD2D_IC_ORSZAG_TANG, ic_nextra = 0
  real(c_double) :: ic_extra(64) = 0
  ! frame files of RK3 (com.txt:1 flash, frameMAX): flash = 1 writes
  ! Q1..Q8flash.txt and T.txt as RK3.f90:29-53,154-172 do, flash = 0 only
  ! the t = 0 frame (the reference opens and writes it whatever flash is),
  ! flash = -1 (this host's default) no frame files at all
  integer :: flash = -1, frameMAX = 200
  integer, private :: frame_unit(9) = 0

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

  ! RK3.f90:29-37: the frame files (replaced if they exist)
  subroutine frame_open()
    integer :: n
    character(len=16) :: fname
    do n = 1, NumEq
      write(fname, '(a,i0,a)') 'Q', n, 'flash.txt'
      open(newunit=frame_unit(n), file=trim(fname), status='replace')
    end do
    open(newunit=frame_unit(9), file='T.txt', status='replace')
  end subroutine frame_open

  ! RK3.f90:39-53 / :156-170: one frame, every mode d of every interior
  ! cell in d -> j -> i order, one value per line and file, then t.  Only a
  ! frame downloads the state.
  subroutine frame_write()
    integer :: i, j, d, n
    call gpu_download()
    do d = 1, dimPk
      do j = 1, Ny
        do i = 1, Nx
          do n = 1, NumEq
            write(frame_unit(n), *) uh(i, j, d, n)
          end do
        end do
      end do
    end do
    write(frame_unit(9), *) t
  end subroutine frame_write

  subroutine frame_close()
    integer :: n
    do n = 1, 9
      if (frame_unit(n) /= 0) close(frame_unit(n))
      frame_unit(n) = 0
    end do
  end subroutine frame_close

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

! init_data.f90:1-121 on the device, argument-less like the reference: the
! initial condition is ic_select with ic_extra(1:ic_nextra) (mhd2d_common)
subroutine init_data()
  use mhd2d_common
  call check(mhd2d_init_data(ctx, int(ic_select, c_int32_t), ic_extra, int(ic_nextra, c_int32_t)), 'init_data')
end subroutine init_data

! the same for an explicit initial condition (ids and parameters of
! mhd2d_init_data in include/mhd2d.h)
subroutine init_data_ic(ic, extra, nextra)
  use mhd2d_common
  integer, intent(in) :: ic, nextra
  real(c_double), intent(in) :: extra(*)
  call check(mhd2d_init_data(ctx, int(ic, c_int32_t), extra, int(nextra, c_int32_t)), 'init_data')
end subroutine init_data_ic

! calculate_Ez.f90:1-7: the corner electric field Ez = u2*B1 - u1*B2, the
! same function the vertex kernels evaluate on the device
subroutine calculate_Ez(Ez, u1, u2, B1, B2)
  use mhd2d_common
  real(c_double), intent(out) :: Ez
  real(c_double), intent(in) :: u1, u2, B1, B2
  real(c_double) :: a(1), b(1), c(1), d(1), e(1)
  a(1) = u1
  b(1) = u2
  c(1) = B1
  d(1) = B2
  call check(mhd2d_calculate_ez(1_c_int64_t, a, b, c, d, e), 'calculate_Ez')
  Ez = e(1)
end subroutine calculate_Ez

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

! RK3.f90:5-174.  t, dt and count stay on the device inside the loop (the
! step clips dt to tend, and a step that starts at t >= tend changes
! nothing, exactly as the reference's loop test stops it).  The steps run
! as HIP graph launches, 16 per read of t; with verbose or frame files one
! per read, and with verbose the per-step line of RK3.f90:151 is printed (it
! costs one extra reduction per step).  Frames (flash, see mhd2d_common) are
! the only downloads of the loop.
subroutine RK3()
  use mhd2d_common
  integer(c_int64_t) :: chunk
  real(c_double) :: tt
  integer :: countf
  chunk = 16
  if (verbose .or. flash == 1) chunk = 1
  call check(mhd2d_rk3_prologue(ctx), 'RK3 prologue')
  call calculate_umax()
  call calculate_totaldiv()
  call calculate_pmin()
  t = 0
  count = 0
  if (verbose) print *, t, "  ", umax, "  ", totaldiv
  if (flash >= 0) then
    call frame_open()
    call frame_write()
  end if
  tt = tend / frameMAX
  countf = 1
  do while (t < tend)
    call check(mhd2d_step_graph(ctx, tend, chunk, t), 'RK3 step')
    if (verbose) then
      call check(mhd2d_diagnostics(ctx, umax, totaldiv, pmin), 'RK3 diagnostics')
      print *, t, "  ", umax, "  ", totaldiv
    end if
    if ((t > tt * countf .or. t == tend) .and. flash == 1) then   ! RK3.f90:154
      call frame_write()
      countf = countf + 1
    end if
  end do
  if (flash >= 0) call frame_close()
  call sync_time()
end subroutine RK3

! RK3 written with the reference's own call sequence (RK3.f90:60-149):
! every whole-array statement of the reference is one mhd2d_rk_stage call.
subroutine RK3_subroutines(max_steps)
  use mhd2d_common
  integer, intent(in) :: max_steps
  integer :: n, countf
  real(c_double) :: tt
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
  if (flash >= 0) then
    call frame_open()
    call frame_write()
  end if
  tt = tend / frameMAX
  countf = 1
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
    if ((t > tt * countf .or. t == tend) .and. flash == 1) then   ! RK3.f90:154
      call frame_write()
      countf = countf + 1
    end if
  end do
  if (flash >= 0) call frame_close()
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
