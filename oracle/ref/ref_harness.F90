!> Golden-vector harness around the reference serial modules (test infrastructure).
!!
!! Links the UNMODIFIED reference sources src/common/mom6_types.F90 and
!! src/serial/mom6_{continuity,coriolis,barotropic,vert_visc,hor_visc}_serial.F90
!! (compiled in place from /root/reference by oracle/ref/Makefile) and replays the
!! step order of app/rk2_serial_driver.F90:218-399 on a state read from a file.
!! Differences from running that driver binary, all deliberate:
!!   * every array the reference allocates without initialising is zeroed
!!     (the value glibc's fresh pages give it), so small grids are deterministic;
!!   * the state, wind stress, Coriolis parameter and masks come from the input;
!!   * the htmp copy runs over isd..ied (rk2_serial_driver.F90:375 says jed);
!!   * for periodic grids the halo is refilled at the exchange points of
!!     app/rk2_mpi_serial_driver.F90 and btstep uses the split API there.
!! Usage: mom6_ref_harness <config.txt>
!! dump_mode < 0: no output file; each iteration's wall time is printed
!! ("RK2 iteration N: t s", as the reference driver prints it) instead.
program mom6_ref_harness
    use iso_fortran_env, only: dp => real64, int64, int32
    use mom6_types, only: ocean_grid_type, verticalGrid_type, init_ocean_grid, init_verticalGrid, &
                          BT_cont_type, alloc_BT_cont_type, mech_forcing_type, vertvisc_type, &
                          init_mech_forcing, init_vertvisc_visc
    use mom6_continuity_serial, only: continuity_CS, continuity_init, continuity_PPM
    use mom6_coriolis_serial, only: coriolis_CS, coriolis_init, CorAdCalc
    use mom6_barotropic_serial, only: barotropic_CS, barotropic_init, btstep, btstep_init_state, &
                                      btstep_do_step, btstep_get_output
    use mom6_vert_visc_serial, only: vert_visc_CS, vert_visc_init, vert_visc_coef, vert_visc_apply
    use mom6_hor_visc_serial, only: hor_visc_CS, hor_visc_init, hor_visc
    implicit none

    type(ocean_grid_type) :: G
    type(verticalGrid_type) :: GV
    type(continuity_CS) :: cont_CS
    type(coriolis_CS) :: cor_CS
    type(barotropic_CS) :: bt_CS
    type(vert_visc_CS) :: visc_CS
    type(hor_visc_CS) :: hvisc_CS
    type(BT_cont_type), pointer :: BT_cont
    type(mech_forcing_type) :: forces
    type(vertvisc_type) :: visc

    real(dp), allocatable, dimension(:, :, :) :: u, v, h, h0, htmp, uh, vh, CAu, CAv, up, vp, diffu, diffv
    real(dp), allocatable, dimension(:, :, :) :: u_cor, por_face_areaU, visc_rem_u
    real(dp), allocatable, dimension(:, :) :: eta, ubt, vbt, ubt_av, vbt_av, eta_av, uhbt, du_cor

    integer :: ni, nj, nk, niter, bt_nsteps, bc_x, bc_y, dump_mode
    integer :: lap, bih, smag, noslip, bbound, leith, scheme, mono, coupled, ray3d
    real(dp) :: dt, dx_km, Kh, Ah, Kv, Kv_ml, Kv_bbl, Hmix, Hbbl, ray_bot
    character(len=512) :: cfgfile, infile, outfile
    integer :: iu, ou, iter, k, i, j
    integer(int64) :: clk0, clk1, clk_rate
    logical :: per

    call get_command_argument(1, cfgfile)
    open (newunit=iu, file=trim(cfgfile), status='old', action='read')
    read (iu, *) ni, nj, nk, niter, bt_nsteps
    read (iu, *) dt, dx_km
    read (iu, *) bc_x, bc_y
    read (iu, *) Kh, Ah, lap, bih, smag, noslip, bbound, leith
    read (iu, *) Kv, Kv_ml, Kv_bbl, Hmix, Hbbl, ray_bot, ray3d
    read (iu, *) scheme, mono, coupled
    read (iu, '(A)') infile
    read (iu, '(A)') outfile
    read (iu, *) dump_mode
    close (iu)
    per = (bc_x == 1) .or. (bc_y == 1)

    call init_ocean_grid(G, ni, nj, nk, dx_km, 45.0_dp)
    call init_verticalGrid(GV, nk)

    allocate (u(G%isd:G%ied, G%jsd:G%jed, nk), source=0.0_dp)
    allocate (v, h, h0, htmp, uh, vh, CAu, CAv, up, vp, diffu, diffv, mold=u)
    v = 0; h = 0; h0 = 0; htmp = 0; uh = 0; vh = 0; CAu = 0; CAv = 0; up = 0; vp = 0
    diffu = 0; diffv = 0
    allocate (eta(G%isd:G%ied, G%jsd:G%jed), source=0.0_dp)
    allocate (ubt, vbt, ubt_av, vbt_av, eta_av, mold=eta)
    ubt = 0; vbt = 0; ubt_av = 0; vbt_av = 0; eta_av = 0

    call init_mech_forcing(forces, G)
    call init_vertvisc_visc(visc, G, GV, use_rayleigh=.true.)
    visc%Ray_u = 0.0_dp; visc%Ray_v = 0.0_dp
    visc%Ray_u(:, :, nk) = ray_bot; visc%Ray_v(:, :, nk) = ray_bot
    if (ray3d /= 0) then   ! Rayleigh drag at every level: ray_bot*k/nk
        do k = 1, nk
            visc%Ray_u(:, :, k) = ray_bot * real(k, dp) / real(nk, dp)
            visc%Ray_v(:, :, k) = ray_bot * real(k, dp) / real(nk, dp)
        end do
    end if

    ! ---- inputs: u v h0 (3D); eta ubt vbt taux tauy CoriolisBu masks T Cu Cv Bu (2D)
    open (newunit=iu, file=trim(infile), access='stream', form='unformatted', status='old', action='read')
    read (iu) u, v, h0, eta, ubt, vbt, forces%taux, forces%tauy, G%CoriolisBu, &
        G%mask2dT, G%mask2dCu, G%mask2dCv, G%mask2dBu
    close (iu)
    h = h0

    call continuity_init(cont_CS, G, GV, uhbt, u_cor, du_cor, por_face_areaU, visc_rem_u)
    uhbt = 0.0_dp; u_cor = 0.0_dp; du_cor = 0.0_dp
    cont_CS%monotonic = (mono /= 0)
    call alloc_BT_cont_type(BT_cont, G, GV)
    call coriolis_init(cor_CS, G, scheme)
    cor_CS%dvdx = 0; cor_CS%dudy = 0; cor_CS%rel_vort = 0; cor_CS%abs_vort = 0; cor_CS%q = 0
    cor_CS%Ih_q = 0; cor_CS%hArea_u = 0; cor_CS%hArea_v = 0; cor_CS%KE = 0
    cor_CS%a = 0; cor_CS%b = 0; cor_CS%c = 0; cor_CS%d = 0
    call barotropic_init(bt_CS, G, dt, bt_nsteps)
    bt_CS%eta = 0; bt_CS%eta_pred = 0; bt_CS%ubt = 0; bt_CS%vbt = 0; bt_CS%ubt_prev = 0
    bt_CS%vbt_prev = 0; bt_CS%uhbt = 0; bt_CS%vhbt = 0; bt_CS%PFu = 0; bt_CS%PFv = 0
    bt_CS%Cor_u = 0; bt_CS%Cor_v = 0; bt_CS%ubt_av = 0; bt_CS%vbt_av = 0
    bt_CS%uhbt_av = 0; bt_CS%vhbt_av = 0
    if (bc_y == 2) bt_CS%Datv = bt_CS%Datv*G%mask2dCv
    call vert_visc_init(visc_CS, G, GV, Kv=Kv, Kv_ml=Kv_ml, Kv_extra_bbl=Kv_bbl, Hmix=Hmix, Hbbl=Hbbl)
    visc_CS%a_u = 0; visc_CS%a_v = 0; visc_CS%h_u = 0; visc_CS%h_v = 0
    visc_CS%visc_rem_u = 0; visc_CS%visc_rem_v = 0
    call hor_visc_init(hvisc_CS, G, GV, Kh=Kh, Ah=Ah, Laplacian=(lap /= 0), biharmonic=(bih /= 0), &
                       Smagorinsky=(smag /= 0), Leith=(leith /= 0), no_slip=(noslip /= 0), &
                       better_bound=(bbound /= 0))

    if (per) then
        call fill3(u); call fill3(v); call fill3(h); call fill3(h0)
        call fill2(eta); call fill2(ubt); call fill2(vbt)
    end if

    if (dump_mode >= 0) &
        open (newunit=ou, file=trim(outfile), access='stream', form='unformatted', status='replace')

    do iter = 1, niter
        call system_clock(clk0, clk_rate)
        h = h0
        ! ---------------- predictor
        call transports(u, v, h)
        if (per) then; call fill3(uh); call fill3(vh); end if
        if (dump_mode > 0 .and. iter == 1) then
            call put3('p_uh', uh); call put3('p_vh', vh)
        end if
        call hor_visc(u, v, h, diffu, diffv, G, GV, hvisc_CS, uh, vh)
        call CorAdCalc(u, v, h, uh, vh, CAu, CAv, G, GV, cor_CS)
        if (dump_mode > 0 .and. iter == 1) then
            call put3('p_diffu', diffu); call put3('p_diffv', diffv)
            call put3('p_CAu', CAu); call put3('p_CAv', CAv)
        end if
        do k = 1, nk; do j = G%jsc, G%jec; do i = G%isc, G%iec - 1
            up(i, j, k) = u(i, j, k) + dt*(CAu(i, j, k) + diffu(i, j, k))
        end do; end do; end do
        do k = 1, nk; do j = G%jsc, G%jec - 1; do i = G%isc, G%iec
            vp(i, j, k) = v(i, j, k) + dt*(CAv(i, j, k) + diffv(i, j, k))
        end do; end do; end do
        if (dump_mode > 0 .and. iter == 1) then
            call put3('p_up0', up); call put3('p_vp0', vp)
        end if
        call vert_visc_coef(up, vp, h, visc_CS, G, GV)
        call vert_visc_apply(up, vp, h, dt, visc_CS, G, GV, forces, visc)
        if (dump_mode > 0 .and. iter == 1) then
            call put3('p_up', up); call put3('p_vp', vp)
            call put2('p_taux_bot', visc_CS%taux_bot); call put2('p_tauy_bot', visc_CS%tauy_bot)
        end if
        call barotropic(eta, ubt, vbt, ubt_av, vbt_av, eta_av)
        ! coupled form: continuity's target is btstep's time-averaged transport
        if (coupled /= 0) uhbt = bt_CS%Datu*ubt_av
        if (dump_mode > 0 .and. iter == 1) then
            call put2('p_ubt_av', ubt_av); call put2('p_vbt_av', vbt_av); call put2('p_eta_av', eta_av)
        end if
        if (per) then; call fill3(up); call fill3(vp); end if
        call continuity_PPM(up, h0, h, uh, dt, G, GV, cont_CS, por_face_areaU, uhbt, &
                            visc_rem_u, u_cor, BT_cont, du_cor)
        if (per) call fill3(h)
        if (dump_mode > 0 .and. iter == 1) then
            call put3('p_h', h); call put3('p_uhc', uh); call put3('p_u_cor', u_cor)
            call put2('p_du_cor', du_cor); call put_btc('p_')
        end if
        ! ---------------- corrector
        call transports(up, vp, h)
        if (per) then; call fill3(uh); call fill3(vh); end if
        call hor_visc(up, vp, h, diffu, diffv, G, GV, hvisc_CS, uh, vh)
        call CorAdCalc(up, vp, h, uh, vh, CAu, CAv, G, GV, cor_CS)
        do k = 1, nk; do j = G%jsc, G%jec; do i = G%isc, G%iec - 1
            u(i, j, k) = u(i, j, k) + 0.5_dp*dt*(CAu(i, j, k) + diffu(i, j, k))
        end do; end do; end do
        do k = 1, nk; do j = G%jsc, G%jec - 1; do i = G%isc, G%iec
            v(i, j, k) = v(i, j, k) + 0.5_dp*dt*(CAv(i, j, k) + diffv(i, j, k))
        end do; end do; end do
        call vert_visc_coef(u, v, h, visc_CS, G, GV)
        call vert_visc_apply(u, v, h, 0.5_dp*dt, visc_CS, G, GV, forces, visc)
        call barotropic(eta_av, ubt_av, vbt_av, ubt_av, vbt_av, eta)
        if (coupled /= 0) uhbt = bt_CS%Datu*ubt_av
        htmp = h
        call continuity_PPM(u, htmp, h, uh, 0.5_dp*dt, G, GV, cont_CS, por_face_areaU, uhbt, &
                            visc_rem_u, u_cor, BT_cont, du_cor)
        if (per) then
            call fill3(u); call fill3(v); call fill3(h)
            call fill2(eta); call fill2(ubt); call fill2(vbt)
        end if
        call system_clock(clk1)
        if (dump_mode < 0) print '(A,I4,A,F10.6,A)', '  RK2 iteration ', iter, ':  ', &
            real(clk1 - clk0, dp) / real(clk_rate, dp), ' s'
    end do
    if (dump_mode < 0) stop

    call put3('u', u); call put3('v', v); call put3('h', h); call put3('h0', h0)
    call put3('up', up); call put3('vp', vp); call put3('uh', uh); call put3('vh', vh)
    call put3('CAu', CAu); call put3('CAv', CAv); call put3('diffu', diffu); call put3('diffv', diffv)
    call put2('eta', eta); call put2('ubt_av', ubt_av); call put2('vbt_av', vbt_av)
    call put2('eta_av', eta_av); call put3('u_cor', u_cor); call put2('du_cor', du_cor)
    if (coupled /= 0) call put2('uhbt', uhbt)
    call put_btc('')
    call put2('taux_bot', visc_CS%taux_bot); call put2('tauy_bot', visc_CS%tauy_bot)
    close (ou)

contains

    subroutine transports(uu, vv, hh)  ! rk2_serial_driver.F90:544-561
        real(dp), intent(in) :: uu(G%isd:, G%jsd:, :), vv(G%isd:, G%jsd:, :), hh(G%isd:, G%jsd:, :)
        integer :: ii, jj, kk
        do kk = 1, nk; do jj = G%jsd, G%jed; do ii = G%isd, G%ied
            uh(ii, jj, kk) = uu(ii, jj, kk)*0.5_dp*(hh(ii, jj, kk) + hh(min(ii + 1, G%ied), jj, kk))*G%dyCu(ii, jj)
            vh(ii, jj, kk) = vv(ii, jj, kk)*0.5_dp*(hh(ii, jj, kk) + hh(ii, min(jj + 1, G%jed), kk))*G%dxCv(ii, jj)
        end do; end do; end do
    end subroutine transports

    subroutine barotropic(e_in, u_in, v_in, ua, va, e_out)
        real(dp), intent(in) :: e_in(G%isd:, G%jsd:), u_in(G%isd:, G%jsd:), v_in(G%isd:, G%jsd:)
        real(dp), intent(inout) :: ua(G%isd:, G%jsd:), va(G%isd:, G%jsd:), e_out(G%isd:, G%jsd:)
        integer :: n
        real(dp), allocatable :: ei(:, :), ui(:, :), vi(:, :)
        if (.not. per) then
            ! btstep's in/out arguments alias in the corrector call; pass copies
            ei = e_in; ui = u_in; vi = v_in
            call btstep(ei, ui, vi, ua, va, e_out, G, bt_CS)
        else
            call btstep_init_state(bt_CS, G, e_in, u_in, v_in)
            do n = 1, bt_CS%nstep
                call btstep_do_step(bt_CS, G, n)
                if (n < bt_CS%nstep) then
                    call fill2(bt_CS%ubt); call fill2(bt_CS%vbt); call fill2(bt_CS%eta)
                end if
            end do
            call btstep_get_output(bt_CS, G, ua, va, e_out)
        end if
    end subroutine barotropic

    subroutine fill3(a)
        real(dp), intent(inout) :: a(G%isd:, G%jsd:, :)
        integer :: kk
        do kk = 1, size(a, 3)
            call fill2(a(:, :, kk))
        end do
    end subroutine fill3

    subroutine fill2(a)  ! periodic refill of the 3-wide halo, x then y
        real(dp), intent(inout) :: a(G%isd:, G%jsd:)
        integer :: ii, jj, P, Q
        P = G%iec - G%isc + 1; Q = G%jec - G%jsc + 1
        if (bc_x == 1) then
            do jj = G%jsd, G%jed
                do ii = G%isd, G%isc - 1
                    a(ii, jj) = a(ii + P, jj)
                end do
                do ii = G%iec + 1, G%ied
                    a(ii, jj) = a(ii - P, jj)
                end do
            end do
        end if
        if (bc_y == 1) then
            do ii = G%isd, G%ied
                do jj = G%jsd, G%jsc - 1
                    a(ii, jj) = a(ii, jj + Q)
                end do
                do jj = G%jec + 1, G%jed
                    a(ii, jj) = a(ii, jj - Q)
                end do
            end do
        end if
    end subroutine fill2

    subroutine put3(name, a)
        character(len=*), intent(in) :: name
        real(dp), intent(in) :: a(:, :, :)
        character(len=16) :: nm
        nm = name
        write (ou) nm, int(size(a), int64), a
    end subroutine put3

    subroutine put2(name, a)
        character(len=*), intent(in) :: name
        real(dp), intent(in) :: a(:, :)
        character(len=16) :: nm
        nm = name
        write (ou) nm, int(size(a), int64), a
    end subroutine put2

    subroutine put_btc(pfx)
        character(len=*), intent(in) :: pfx
        call put2(pfx//'FA_u_EE', BT_cont%FA_u_EE); call put2(pfx//'FA_u_E0', BT_cont%FA_u_E0)
        call put2(pfx//'FA_u_W0', BT_cont%FA_u_W0); call put2(pfx//'FA_u_WW', BT_cont%FA_u_WW)
        call put2(pfx//'uBT_WW', BT_cont%uBT_WW); call put2(pfx//'uBT_EE', BT_cont%uBT_EE)
        call put3(pfx//'BT_h_u', BT_cont%h_u)
    end subroutine put_btc

end program mom6_ref_harness
