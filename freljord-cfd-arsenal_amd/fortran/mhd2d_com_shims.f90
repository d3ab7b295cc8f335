! mhd2d_com_shims.f90 -- the hot-path subroutines of GDF_Balsara_P2_0316 for
! a build of the reference itself: compile this file with the reference's
! com.txt on the include path, in place of set_bc.f90, Lh.f90, LhEz.f90,
! TVB_Limiter_P2.f90, pp_Limiter.f90, KXRCF_Detector.f90,
! div_free_Balsara.f90, calculate_dt.f90, calculate_umax.f90,
! calculate_totaldiv.f90 and calculate_pmin.f90, and link mhd2d_capi.o and
! libmhd2d.so.  The reference's own main.f90, RK3.f90, init_data.f90,
! get_GLP.f90, get_basis.f90, save_*.f90 and calculate_L2error.f90 stay as
! they are: they talk to these subroutines through COMMON (com.txt:19-54).
!
! Each subroutine pushes the COMMON arrays it reads to the device context,
! makes one C-ABI call and pulls what the reference routine writes back
! into COMMON, in the reference layouts.  This is the per-call mode: every
! call crosses PCIe, so it suits checking and small grids; the state-resident
! mode is the host of mhd2d_host.f90.  The context is created on the first
! call from the compile-time configuration (com.txt:1 Nx, Ny, flux_type) and
! the run-time one in COMMON (/mesh/ domain, /bc/ boundaries, M), with the
! reference's shipped choices: TVB on every cell, no positivity limiter
! (TVB_Limiter_P2.f90:12,22,123 are comments), CFL 0.15 (calculate_dt.f90:8).
! Compile with the reference's own flags (-fdefault-real-8
! -fdefault-double-8): com.txt relies on implicit REAL(8).

module mhd2d_com_ctx
  use mhd2d_capi
  implicit none
  type(c_ptr), save :: ctx = c_null_ptr
contains
  ! negative status: stop with the library's message; positive: numerical
  ! flags, which the reference would not notice, reported once per value
  subroutine com_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    integer(c_int), save :: reported = 0
    if (rc < 0) then
      write(0, '(a,a,i0,a,a)') what, ': status ', rc, ': ', trim(last_error(ctx))
      error stop 1
    else if (rc > 0 .and. rc /= reported) then
      write(0, '(a,a,i0)') what, ': numerical flags ', rc
      reported = rc
    end if
  end subroutine com_check
end module mhd2d_com_ctx

! the context, from the COMMON configuration, on first use
subroutine com_open()
  use mhd2d_com_ctx
  include 'com.txt'
  type(mhd2d_config) :: c
  integer(c_int) :: rc
  if (c_associated(ctx)) return
  c = mhd2d_config(Nx, Ny, flux_type, nint(bcR), nint(bcL), nint(bcU), nint(bcD), 0, MHD2D_IND_ALL, MHD2D_FP64, 0, &
                   xa, xb, ya, yb, M, 0.15d0, 1d0, 0.2d0, 0.8d0)
  rc = mhd2d_create(c, ctx)
  if (rc /= 0) then
    write(0, '(a,i0)') 'mhd2d_create failed with status ', rc
    error stop 1
  end if
end subroutine com_open

! COMMON uh, Bx, By -> device
subroutine com_push()
  use mhd2d_com_ctx
  include 'com.txt'
  call com_open()
  call com_check(mhd2d_upload(ctx, uh, Bx, By), 'upload')
end subroutine com_push

! device -> COMMON uh, Bx, By
subroutine com_pull()
  use mhd2d_com_ctx
  include 'com.txt'
  call com_check(mhd2d_download(ctx, uh, Bx, By), 'download')
end subroutine com_pull

subroutine set_bc                  ! set_bc.f90:1-28
  use mhd2d_com_ctx
  call com_push()
  call com_check(mhd2d_set_bc(ctx), 'set_bc')
  call com_pull()
end subroutine set_bc

subroutine Lh                      ! Lh.f90:1-305: du
  use mhd2d_com_ctx
  include 'com.txt'
  call com_push()
  call com_check(mhd2d_lh(ctx), 'Lh')
  call com_check(mhd2d_debug_get(ctx, MHD2D_DU, du), 'Lh du')
end subroutine Lh

subroutine LhEz                    ! LhEz.f90:1-97: dEz1, dEz2, EzVertex (after Lh)
  use mhd2d_com_ctx
  include 'com.txt'
  call com_check(mhd2d_lhez(ctx), 'LhEz')
  call com_check(mhd2d_debug_get(ctx, MHD2D_DEZ1, dEz1), 'LhEz dEz1')
  call com_check(mhd2d_debug_get(ctx, MHD2D_DEZ2, dEz2), 'LhEz dEz2')
  call com_check(mhd2d_debug_get(ctx, MHD2D_EZVERTEX, EzVertex), 'LhEz EzVertex')
end subroutine LhEz

subroutine TVB_Limiter_P2          ! TVB_Limiter_P2.f90:1-191 (its set_bc included)
  use mhd2d_com_ctx
  call com_push()
  call com_check(mhd2d_tvb_limiter_p2(ctx), 'TVB_Limiter_P2')
  call com_pull()
end subroutine TVB_Limiter_P2

subroutine pp_Limiter              ! pp_Limiter.f90:1-125
  use mhd2d_com_ctx
  call com_push()
  call com_check(mhd2d_pp_limiter(ctx), 'pp_Limiter')
  call com_pull()
end subroutine pp_Limiter

subroutine KXRCF_Detector          ! KXRCF_Detector.f90:1-110: Is_Trouble_Cell
  use mhd2d_com_ctx
  include 'com.txt'
  integer(c_int64_t) :: nflag
  real(c_double), allocatable :: flags(:, :)
  call com_push()
  call com_check(mhd2d_kxrcf_detector(ctx, nflag), 'KXRCF_Detector')
  allocate(flags(Nx, Ny))
  call com_check(mhd2d_debug_get(ctx, MHD2D_TROUBLE, flags), 'KXRCF_Detector flags')
  Is_Trouble_Cell = flags
end subroutine KXRCF_Detector

subroutine div_free_Balsara        ! div_free_Balsara.f90:1-81
  use mhd2d_com_ctx
  call com_push()
  call com_check(mhd2d_div_free_balsara(ctx), 'div_free_Balsara')
  call com_pull()
end subroutine div_free_Balsara

subroutine calculate_dt            ! calculate_dt.f90:1-28: dt
  use mhd2d_com_ctx
  include 'com.txt'
  real(c_double) :: dtc
  call com_push()
  call com_check(mhd2d_calculate_dt(ctx, dtc), 'calculate_dt')
  dt = dtc
end subroutine calculate_dt

subroutine calculate_umax          ! calculate_umax.f90:1-14
  use mhd2d_com_ctx
  include 'com.txt'
  real(c_double) :: a, b, c
  call com_push()
  call com_check(mhd2d_diagnostics(ctx, a, b, c), 'calculate_umax')
  umax = a
end subroutine calculate_umax

subroutine calculate_totaldiv      ! calculate_totaldiv.f90:1-24
  use mhd2d_com_ctx
  include 'com.txt'
  real(c_double) :: a, b, c
  call com_push()
  call com_check(mhd2d_diagnostics(ctx, a, b, c), 'calculate_totaldiv')
  totaldiv = b
end subroutine calculate_totaldiv

subroutine calculate_pmin          ! calculate_pmin.f90:1-27
  use mhd2d_com_ctx
  include 'com.txt'
  real(c_double) :: a, b, c
  call com_push()
  call com_check(mhd2d_diagnostics(ctx, a, b, c), 'calculate_pmin')
  pmin = c
end subroutine calculate_pmin
