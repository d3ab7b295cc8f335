! mhd2d_capi.f90 -- iso_c_binding interfaces of include/mhd2d.h, one per
! export, plus the constants of the header.  This module is the whole
! binding: the Fortran host (mhd2d_host.f90) and the COMMON-block shims a
! reference build links in place of its hot-path sources
! (mhd2d_com_shims.f90) both call the library through it.

module mhd2d_capi
  use iso_c_binding
  implicit none

  integer(c_int32_t), parameter :: MHD2D_LF = 1, MHD2D_HLL = 2, MHD2D_HLLC = 3, MHD2D_HLLD = 4
  integer(c_int32_t), parameter :: MHD2D_IND_NONE = 0, MHD2D_IND_ALL = 1, MHD2D_IND_KXRCF = 2, &
                                   MHD2D_IND_MLP = 3
  integer(c_int32_t), parameter :: MHD2D_FP64 = 0, MHD2D_MIXED = 1
  ! arrays of mhd2d_debug_get
  integer(c_int32_t), parameter :: MHD2D_DU = 0, MHD2D_DEZ1 = 1, MHD2D_DEZ2 = 2, MHD2D_EZVERTEX = 3, &
                                   MHD2D_TROUBLE = 4
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
    integer(c_int) function mhd2d_debug_get(ctx, which, host) bind(C, name='mhd2d_debug_get')
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: which
      real(c_double), intent(out) :: host(*)
    end function
    integer(c_int) function mhd2d_face_jump(ctx, jump) bind(C, name='mhd2d_face_jump')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: jump
    end function
    ! uh and out are device addresses (c_ptr); stream a hipStream_t or c_null_ptr
    integer(c_int) function mhd2d_diagnostics_noalloc(Nx, Ny, xa, xb, ya, yb, uh, out, stream) &
        bind(C, name='mhd2d_diagnostics_noalloc')
      import :: c_int, c_int32_t, c_double, c_ptr
      integer(c_int32_t), value :: Nx, Ny
      real(c_double), value :: xa, xb, ya, yb
      type(c_ptr), value :: uh, out, stream
    end function
    integer(c_int) function mhd2d_calculate_ez(n, u1, u2, B1, B2, Ez) bind(C, name='mhd2d_calculate_ez')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n
      real(c_double), intent(in) :: u1(*), u2(*), B1(*), B2(*)
      real(c_double), intent(out) :: Ez(*)
    end function
    ! live kernel timing
    integer(c_int) function mhd2d_profile_enable(ctx, on) bind(C, name='mhd2d_profile_enable')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function
    integer(c_int) function mhd2d_profile_get(ctx, ms, launches, nclass) bind(C, name='mhd2d_profile_get')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(*)
      integer(c_int64_t), intent(out) :: launches(*)
      integer(c_int32_t), value :: nclass
    end function
    integer(c_int) function mhd2d_stream_triad(n, unroll, reps, gbs, ms_per_launch) &
        bind(C, name='mhd2d_stream_triad')
      import :: c_int, c_int32_t, c_int64_t, c_double
      integer(c_int64_t), value :: n
      integer(c_int32_t), value :: unroll, reps
      real(c_double), intent(out) :: gbs, ms_per_launch
    end function
    ! learned troubled-cell indicator
    integer(c_int) function mhd2d_mlp_set_weights(ctx, params, n) bind(C, name='mhd2d_mlp_set_weights')
      import :: c_int, c_ptr, c_int32_t, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(in) :: params(*)
      integer(c_int32_t), value :: n
    end function
    integer(c_int) function mhd2d_mlp_probability(ctx, host_p) bind(C, name='mhd2d_mlp_probability')
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(out) :: host_p(*)
    end function
    integer(c_int) function mhd2d_indicator_stats(ctx, stats) bind(C, name='mhd2d_indicator_stats')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: stats(8)
    end function
    ! FP32 training of the indicator (a trainer is an opaque handle like ctx)
    integer(c_int) function mhd2d_trainer_create(capacity, max_batch, tr) bind(C, name='mhd2d_trainer_create')
      import :: c_int, c_ptr, c_int32_t, c_int64_t
      integer(c_int64_t), value :: capacity
      integer(c_int32_t), value :: max_batch
      type(c_ptr), intent(out) :: tr
    end function
    integer(c_int) function mhd2d_trainer_destroy(tr) bind(C, name='mhd2d_trainer_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: tr
    end function
    type(c_ptr) function mhd2d_trainer_last_error(tr) bind(C, name='mhd2d_trainer_last_error')
      import :: c_ptr
      type(c_ptr), value :: tr
    end function
    integer(c_int) function mhd2d_trainer_collect(tr, ctx, neg_keep, seed, n_total, n_pos) &
        bind(C, name='mhd2d_trainer_collect')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: tr, ctx
      real(c_double), value :: neg_keep
      integer(c_int64_t), value :: seed
      integer(c_int64_t), intent(out) :: n_total, n_pos
    end function
    integer(c_int) function mhd2d_trainer_size(tr, n, npos) bind(C, name='mhd2d_trainer_size')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: tr
      integer(c_int64_t), intent(out) :: n, npos
    end function
    integer(c_int) function mhd2d_trainer_clear(tr) bind(C, name='mhd2d_trainer_clear')
      import :: c_int, c_ptr
      type(c_ptr), value :: tr
    end function
    integer(c_int) function mhd2d_trainer_get_data(tr, offset, count, X, Y, idx) bind(C, name='mhd2d_trainer_get_data')
      import :: c_int, c_ptr, c_int64_t, c_float
      type(c_ptr), value :: tr
      integer(c_int64_t), value :: offset, count
      real(c_float), intent(out) :: X(*), Y(*)
      integer(c_int64_t), intent(out) :: idx(*)
    end function
    integer(c_int) function mhd2d_trainer_set_data(tr, X, Y, n) bind(C, name='mhd2d_trainer_set_data')
      import :: c_int, c_ptr, c_int64_t, c_float
      type(c_ptr), value :: tr
      real(c_float), intent(in) :: X(*), Y(*)
      integer(c_int64_t), value :: n
    end function
    integer(c_int) function mhd2d_trainer_grad(tr, params, offset, count, pos_weight, grad, loss) &
        bind(C, name='mhd2d_trainer_grad')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
      type(c_ptr), value :: tr
      real(c_float), intent(in) :: params(*)
      integer(c_int64_t), value :: offset
      integer(c_int32_t), value :: count
      real(c_float), value :: pos_weight
      real(c_float), intent(out) :: grad(*), loss
    end function
    integer(c_int) function mhd2d_trainer_fit(tr, params, steps, batch, lr, pos_weight, seed, losses) &
        bind(C, name='mhd2d_trainer_fit')
      import :: c_int, c_ptr, c_int32_t, c_int64_t, c_float
      type(c_ptr), value :: tr
      real(c_float), intent(inout) :: params(*)
      integer(c_int32_t), value :: steps, batch
      real(c_float), value :: lr, pos_weight
      integer(c_int64_t), value :: seed
      real(c_float), intent(out) :: losses(*)
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
