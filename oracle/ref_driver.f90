! ref_driver.f90 -- a driver of this project's own, linked against the
! reference Fortran sources of GDF_Balsara_P2_0316 (built by
! oracle/build_ref.sh into oracle/_ref/).  It loads a state in the
! reference's COMMON layout, calls ONE reference subroutine (or the
! reference's RK3 time loop) and dumps the result, so that the CPU
! restatement (oracle/mhd2d_oracle.c) and the HIP path can be checked
! against the reference itself.  TEST INFRASTRUCTURE ONLY.
!
! usage: ref_driver MODE INPUT OUTPUT
!   INPUT : stream file = 16 doubles header + uh + Bx + By (reference layout)
!           header: xa xb ya yb bcR bcL bcU bcD M tend t dt nsteps 0 0 0
!   MODE  : setbc | lh | tvb | pp | divfree | kxrcf | dt | diag | run | bench
!           | tables | eig | init1 | init3 | init4 | init5 | init6 | init6b
!           | init7 | init8 | init8b | init9 | init9b | az | l2err
!           (l2err: t20 = header slot 14; the 8 errors go to OUTPUT.l2)
!   OUTPUT: 8 scalars, 8 mesh values, uh, Bx, By, du, dEz1, dEz2, EzVertex,
!           Is_Trouble_Cell (as doubles)
program ref_driver
    include 'com.txt'
    character(len=512) :: mode, fin, fout
    real(kind=8) :: hdr(16), scal(8)
    real(kind=8) :: evr(8,8), evl(8,8), cellU(8)
    real(kind=8), allocatable :: eigs(:,:,:,:,:,:)
    integer :: istep, nsteps
    integer(kind=8) :: c0, c1, crate

    call get_command_argument(1, mode)
    call get_command_argument(2, fin)
    call get_command_argument(3, fout)

    open(10, file=trim(fin), access='stream', form='unformatted', status='old')
    read(10) hdr
    read(10) uh, Bx, By
    close(10)

    xa = hdr(1)
    xb = hdr(2)
    ya = hdr(3)
    yb = hdr(4)
    bcR = hdr(5)
    bcL = hdr(6)
    bcU = hdr(7)
    bcD = hdr(8)
    M = hdr(9)
    tend = hdr(10)
    t = hdr(11)
    dt = hdr(12)
    nsteps = int(hdr(13))

    hx = (xb - xa)/Nx
    hy = (yb - ya)/Ny
    hx1 = 0.5d0*hx
    hy1 = 0.5d0*hy

    call get_GLP
    call get_basis

    scal = 0
    select case (trim(mode))
    case ('setbc')
        call set_bc
    case ('lh')
        call set_bc
        call Lh
        call LhEz
    case ('tvb')
        call TVB_Limiter_P2
    case ('pp')
        call pp_Limiter
    case ('divfree')
        call div_free_Balsara
    case ('kxrcf')
        call KXRCF_Detector
    case ('dt')
        call calculate_dt
        scal(2) = dt
        scal(6) = alphax
        scal(7) = alphay
    case ('diag')
        call calculate_umax
        call calculate_totaldiv
        call calculate_pmin
    case ('run')
        call RK3
    case ('bench')
        ! the RK3.f90 prologue and nsteps passes of its loop body, with the
        ! array combinations of RK3.f90:80-135; no frame output
        t = 0
        count = 0
        call set_bc
        call TVB_Limiter_P2
        call div_free_Balsara
        call system_clock(c0, crate)
        do istep = 1, nsteps
            call calculate_dt
            t = t + dt
            call set_bc
            call Lh
            call LhEz
            uI = uh + dt*du
            BxI = Bx + dt*dEz1
            ByI = By + dt*dEz2
            uh0 = uh
            Bx0 = Bx
            By0 = By
            uh = uI
            Bx = BxI
            By = ByI
            call TVB_Limiter_P2
            call div_free_Balsara
            call set_bc
            call Lh
            call LhEz
            uh = uI + dt*du
            Bx = BxI + dt*dEz1
            By = ByI + dt*dEz2
            uII = (3d0/4d0)*uh0 + (1d0/4d0)*uh
            BxII = (3d0/4d0)*Bx0 + (1d0/4d0)*Bx
            ByII = (3d0/4d0)*By0 + (1d0/4d0)*By
            uh = uII
            Bx = BxII
            By = ByII
            call TVB_Limiter_P2
            call div_free_Balsara
            call set_bc
            call Lh
            call LhEz
            Bx = BxII + dt*dEz1
            By = ByII + dt*dEz2
            uh = uII + dt*du
            uh = (1d0/3d0)*uh0 + (2d0/3d0)*uh
            Bx = (1d0/3d0)*Bx0 + (2d0/3d0)*Bx
            By = (1d0/3d0)*By0 + (2d0/3d0)*By
            call TVB_Limiter_P2
            call div_free_Balsara
            count = count + 1
        end do
        call system_clock(c1)
        scal(8) = dble(c1 - c0)/dble(crate)
        call calculate_umax
        call calculate_totaldiv
        call calculate_pmin
    case ('tables')
        ! quadrature and basis tables (get_GLP.f90, get_basis.f90)
        open(12, file=trim(fout)//'.tab', access='stream', form='unformatted', status='replace')
        write(12) lambda, weight, lambdaL, mm, phiG, phiGLL, phixG, phiyG, phiGR, phiGL, phiGU, phiGD
        write(12) phiRU, phiLU, phiRD, phiLD, EzG, EzxG, EzyG, mmE, EzR, EzL, EzU, EzD
        close(12)
    case ('eig')
        ! eigenmatrix.f90 at every interior cell average, n = x and n = y
        allocate(eigs(8,8,2,2,Nx,Ny))
        do j = 1,Ny
            do i = 1,Nx
                cellU = uh(i,j,1,:)
                call eigenmatrix(evr,evl,cellU(1),cellU(2)/cellU(1),cellU(3)/cellU(1),cellU(4)/cellU(1), &
                                 cellU(5),cellU(6),cellU(7),cellU(8),1.0d0,0.0d0,0.0d0)
                eigs(:,:,1,1,i,j) = evr
                eigs(:,:,2,1,i,j) = evl
                call eigenmatrix(evr,evl,cellU(1),cellU(2)/cellU(1),cellU(3)/cellU(1),cellU(4)/cellU(1), &
                                 cellU(5),cellU(6),cellU(7),cellU(8),0.0d0,1.0d0,0.0d0)
                eigs(:,:,1,2,i,j) = evr
                eigs(:,:,2,2,i,j) = evl
            end do
        end do
        open(12, file=trim(fout)//'.eig', access='stream', form='unformatted', status='replace')
        write(12) eigs
        close(12)
    case ('init3')
        call init_data_3
    case ('init1')
        call init_data_1
    case ('init4')
        call init_data_4
    case ('init5')
        call init_data_5
    case ('init6')
        call init_data_6
    case ('init6b')
        call init_data_6b
    case ('init7')
        call init_data_7
    case ('init8')
        call init_data_8
    case ('init8b')
        call init_data_8b
    case ('init9')
        call init_data_9
    case ('init9b')
        call init_data_9b
    case ('az')
        call calculate_Az
    case ('l2err')
        do i = 1, Nx
            Xc(i) = xa + (i - 0.5)*hx
        end do
        do j = 1, Ny
            Yc(j) = ya + (j - 0.5)*hy
        end do
        t20 = hdr(14)
        call calculate_L2error
        open(12, file=trim(fout)//'.l2', access='stream', form='unformatted', status='replace')
        write(12) L2
        close(12)
    case default
        print *, 'unknown mode ', trim(mode)
        stop 2
    end select

    if (trim(mode) /= 'dt') then
        scal(1) = t
        scal(2) = dt
        scal(3) = umax
        scal(4) = totaldiv
        scal(5) = pmin
        if (trim(mode) /= 'bench') scal(8) = count
    end if

    open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(11) scal
    write(11) xa, xb, ya, yb, hx, hy, M, tend
    write(11) uh, Bx, By, du, dEz1, dEz2, EzVertex
    write(11) dble(Is_Trouble_Cell)
    close(11)
end program ref_driver
