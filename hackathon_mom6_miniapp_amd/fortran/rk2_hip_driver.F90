!> RK2 driver on libmom6hip: the split-RK2 step of app/rk2_serial_driver.F90
!! (:218-399) with every module on the GPU through the Fortran module API of
!! mom6_hip.F90 — the counterpart of app/rk2_cuda_c_driver.F90 for gfx950.
!!
!! Usage: rk2_hip_driver [ni] [nj] [nk] [niter] [bt_nsteps] [diag_mode] [infile] [outfile] [options]
!!        Defaults: 180 180 75 10 30 0 (as the reference drivers).
!!   infile  (optional) raw FP64 u, v, h0 (3D), eta, ubt, vbt, CoriolisBu (2D)
!!           over the data domain, Fortran order, replacing the analytic state
!!           ('-' for the analytic state)
!!   outfile (optional) receives raw FP64 u, v, h (3D), eta (2D) after the run
!!           ('-' for none)
!!   options (optional) comma-separated switches of the serial modules that
!!           the reference driver leaves off: smagorinsky, leith, bound
!!           (hor_visc_init's Smagorinsky=, Leith=, better_bound=) and ray3d
!!           (Rayleigh drag at every level, Ray(k) = (1e-4 k)/nk, through the
!!           3D vertvisc fields instead of the bottom layer's 2D coefficient)
!!
!! Configuration of the reference driver: static halos, the beta plane of
!! init_ocean_grid at 45 deg, dx = 10 km, dt = 300 s, Kh = 100 m2/s, Kv = 1e-4,
!! Kv_ml = 1e-2, Kv_extra_bbl = 1e-2, Hmix = 50 m, positive-definite PPM
!! limiter, Sadourny energy Coriolis, taux = 0.1 sin(pi (j-1)/nj) Pa and
!! Rayleigh drag 1e-4 m/s in the bottom layer.  Prints per-iteration wall
!! times ("RK2 iteration N: t s") and the verify_state summary (:563-623).
program rk2_hip_driver
    use iso_c_binding
    use iso_fortran_env, only: dp => real64, int64
    use mom6_hip_capi
    use mom6_hip_common
    use mom6_continuity_hip
    use mom6_coriolis_hip
    use mom6_barotropic_hip
    use mom6_vert_visc_hip
    use mom6_hor_visc_hip
    implicit none

    real(dp), parameter :: G_EARTH = 9.80_dp, OMEGA = 7.2921e-5_dp, EARTH_RADIUS = 6.371e6_dp
    real(dp), parameter :: PI = 3.14159265358979323846_dp, PI_STATE = 3.14159_dp
    real(dp), parameter :: DEPTH = 4000.0_dp

    integer :: ni, nj, nk, niter, bt_nsteps, diag_mode
    integer :: isd, ied, jsd, jed, isc, iec, jsc, jec, n1, n2, i, j, k, iter
    integer(c_size_t) :: n3v
    real(dp) :: dt, dx, f0, beta
    character(len=512) :: arg, infile, outfile, options
    logical :: opt_smag, opt_leith, opt_bound, opt_ray3d

    ! host arrays: grid, forcing, state
    real(dp), allocatable, target :: one(:, :), areaT(:, :), IareaT(:, :), dxv(:, :), Idxv(:, :)
    real(dp), allocatable, target :: CoriolisBu(:, :), Datu(:, :), gtot(:, :), rem(:, :), f_4(:, :, :)
    real(dp), allocatable, target :: dx2(:, :), ratio(:, :), taux(:, :), tauy(:, :), ray(:, :)
    real(dp), allocatable, target :: u(:, :, :), v(:, :, :), h(:, :, :), h0(:, :, :)
    real(dp), allocatable, target :: eta(:, :), ubt(:, :), vbt(:, :), ray3(:, :, :)

    ! device arrays
    type(c_ptr) :: d_u, d_v, d_h, d_h0, d_htmp, d_uh, d_vh, d_CAu, d_CAv, d_up, d_vp, d_diffu, d_diffv
    type(c_ptr) :: d_eta, d_ubt, d_vbt, d_ubt_av, d_vbt_av, d_eta_av, d_taux, d_tauy, d_ray, d_dyCu, d_dxCv
    type(c_ptr) :: d_ray3
    ! diagnostics: the shared 2D work array, KE, mass, reduction workspace, 4 statistics
    type(c_ptr) :: d_work2d, d_ke, d_mass, d_red, d_stats

    type(continuity_CS_hip) :: cont_CS
    type(BT_cont_type_hip), target :: BT_cont
    type(coriolis_CS_hip) :: cor_CS
    type(barotropic_CS_hip) :: bt_CS
    type(vert_visc_CS_hip) :: visc_CS
    type(hor_visc_CS_hip) :: hvisc_CS

    integer(int64) :: c0, c1, rate
    real(dp) :: t_iter, t_total
    real(dp) :: mass_init, mass_final, max_u, max_v, max_eta, ke

    ni = 180; nj = 180; nk = 75; niter = 10; bt_nsteps = 30; diag_mode = 0
    infile = ''; outfile = ''; options = ''
    if (command_argument_count() >= 1) then; call get_command_argument(1, arg); read (arg, *) ni; end if
    if (command_argument_count() >= 2) then; call get_command_argument(2, arg); read (arg, *) nj; end if
    if (command_argument_count() >= 3) then; call get_command_argument(3, arg); read (arg, *) nk; end if
    if (command_argument_count() >= 4) then; call get_command_argument(4, arg); read (arg, *) niter; end if
    if (command_argument_count() >= 5) then; call get_command_argument(5, arg); read (arg, *) bt_nsteps; end if
    if (command_argument_count() >= 6) then; call get_command_argument(6, arg); read (arg, *) diag_mode; end if
    if (command_argument_count() >= 7) call get_command_argument(7, infile)
    if (command_argument_count() >= 8) call get_command_argument(8, outfile)
    if (command_argument_count() >= 9) call get_command_argument(9, options)
    if (trim(infile) == '-') infile = ''
    if (trim(outfile) == '-') outfile = ''
    opt_smag = index(options, 'smagorinsky') > 0
    opt_leith = index(options, 'leith') > 0
    opt_bound = index(options, 'bound') > 0
    opt_ray3d = index(options, 'ray3d') > 0
    if (ni /= nj) then
        ! the reference driver's htmp copy loop runs i over isd..jed (:375)
        print '(A)', 'rk2_hip_driver: ni must equal nj, as for the reference serial driver'
        stop 1
    end if
    dt = 300.0_dp
    dx = 10.0_dp * 1000.0_dp

    print '(A)', '================================================================'
    print '(A)', 'MOM6 Split RK2 — HIP (gfx950) Driver'
    print '(A)', '================================================================'
    print '(A,I5,A,I5,A,I4)', 'Grid: ', ni, ' x ', nj, ' x ', nk
    print '(A,I4)', 'RK2 iterations: ', niter
    print '(A,I4)', 'Barotropic substeps: ', bt_nsteps
    if (diag_mode > 0) then
        print '(A)', 'Diagnostics: ENABLED'
    else
        print '(A)', 'Diagnostics: disabled'
    end if
    if (len_trim(options) > 0) print '(A,A)', 'Options: ', trim(options)

    ! grid bounds (init_ocean_grid, src/common/mom6_types.F90:145-148)
    isd = 1; ied = ni + 2; jsd = 1; jed = nj + 2
    isc = 4; iec = ni - 1; jsc = 4; jec = nj - 1
    n1 = ied - isd + 1; n2 = jed - jsd + 1
    n3v = int(n1, c_size_t) * int(n2, c_size_t) * int(nk, c_size_t)

    ! uniform metrics and the beta plane (:189-226)
    allocate (one(isd:ied, jsd:jed), areaT(isd:ied, jsd:jed), IareaT(isd:ied, jsd:jed))
    allocate (dxv(isd:ied, jsd:jed), Idxv(isd:ied, jsd:jed), CoriolisBu(isd:ied, jsd:jed), Datu(isd:ied, jsd:jed))
    allocate (gtot(isd:ied, jsd:jed), rem(isd:ied, jsd:jed), f_4(4, isd:ied, jsd:jed), dx2(isd:ied, jsd:jed))
    allocate (ratio(isd:ied, jsd:jed), taux(isd:ied, jsd:jed), tauy(isd:ied, jsd:jed), ray(isd:ied, jsd:jed))
    one = 1.0_dp
    areaT = dx * dx; IareaT = 1.0_dp / (dx * dx)
    dxv = dx; Idxv = 1.0_dp / dx
    dx2 = dx * dx
    ratio = dx * (1.0_dp / dx)
    f0 = 2.0_dp * OMEGA * sin(45.0_dp * PI / 180.0_dp)
    beta = 2.0_dp * OMEGA * cos(45.0_dp * PI / 180.0_dp) / EARTH_RADIUS
    do j = jsd, jed
        CoriolisBu(:, j) = f0 + beta * (real(j - 1 - nj / 2, dp) - 0.5_dp) * dx
    end do

    ! analytic state (initialize_state, :466-506) or the given one
    allocate (u(isd:ied, jsd:jed, nk), v(isd:ied, jsd:jed, nk), h(isd:ied, jsd:jed, nk), h0(isd:ied, jsd:jed, nk))
    allocate (eta(isd:ied, jsd:jed), ubt(isd:ied, jsd:jed), vbt(isd:ied, jsd:jed))
    if (len_trim(infile) > 0) then
        open (newunit=i, file=trim(infile), access='stream', form='unformatted', status='old', action='read')
        read (i) u, v, h0, eta, ubt, vbt, CoriolisBu
        close (i)
    else
        do k = 1, nk
            do j = jsd, jed
                do i = isd, ied
                    h0(i, j, k) = DEPTH / nk + 10.0_dp * sin(real(i - 1, dp) / ni * PI_STATE) * &
                                  cos(real(j - 1, dp) / nj * PI_STATE) * exp(-real(k, dp) / 20.0_dp)
                    u(i, j, k) = 0.1_dp * sin(real(j - 1, dp) / nj * PI_STATE * 2.0_dp) * exp(-real(k, dp) / 30.0_dp)
                    v(i, j, k) = 0.1_dp * cos(real(i - 1, dp) / ni * PI_STATE * 2.0_dp) * exp(-real(k, dp) / 30.0_dp)
                end do
            end do
        end do
        do j = jsd, jed
            do i = isd, ied
                eta(i, j) = 0.5_dp * sin(real(i - 1, dp) / ni * PI_STATE * 2.0_dp) * &
                            cos(real(j - 1, dp) / nj * PI_STATE * 2.0_dp)
                ubt(i, j) = 0.05_dp * sin(real(j - 1, dp) / nj * PI_STATE)
                vbt(i, j) = 0.05_dp * cos(real(i - 1, dp) / ni * PI_STATE)
            end do
        end do
    end if
    h = h0
    ! forcing (initialize_forces_visc, :508-542)
    do j = jsd, jed
        taux(:, j) = 0.1_dp * sin(real(j - 1, dp) / nj * PI_STATE)
    end do
    tauy = 0.0_dp
    ray = 1.0e-4_dp

    call check_bt_cfl(dt, bt_nsteps, 10.0_dp, 4000.0_dp)

    ! control structures (the module inits of :130-175)
    call hip_check(mom6h_set_device(0_c_int), 'mom6h_set_device')
    call continuity_init_hip(cont_CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, IareaT, Idxv, dxv, one, &
                             .false., dxv, areaT, dxv, one)
    call alloc_BT_cont_type_hip(BT_cont, isd, ied, jsd, jed, nk)
    call coriolis_init_hip(cor_CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, areaT, IareaT, CoriolisBu, one, &
                           dxv, dxv, dxv, dxv, Idxv, Idxv, SADOURNY75_ENERGY_HIP)
    Datu = DEPTH * dx
    gtot = G_EARTH
    rem = 0.999_dp
    do j = jsd, jed
        do i = isd, ied
            f_4(:, i, j) = 0.25_dp * CoriolisBu(i, j)
        end do
    end do
    call barotropic_init_hip(bt_CS, isd, ied, jsd, jed, isc, iec, jsc, jec, bt_nsteps, dt, 0.2_dp, 0, Datu, Datu, &
                             gtot, gtot, gtot, gtot, f_4, f_4, rem, rem, IareaT, Idxv, Idxv)
    call vert_visc_init_hip(visc_CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, one, one, 1.0e-4_dp, 1.0e-2_dp, &
                            1.0e-2_dp, 50.0_dp, 10.0_dp)
    call vert_visc_keep_coefficients_hip(visc_CS, .false.)   ! the driver never reads a_u, h_u, ...
    call hor_visc_init_hip(hvisc_CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, 100.0_dp, 1.0e-3_dp, ratio, ratio, &
                           ratio, ratio, Idxv, Idxv, Idxv, Idxv, IareaT, IareaT, one, one, one, one, dx2, dx2, &
                           dx2, dx2)
    if (opt_smag .or. opt_leith .or. opt_bound) &
        call hor_visc_set_closures_hip(hvisc_CS, isd, ied, jsd, jed, opt_smag, opt_leith, opt_bound, areaT, dxv, &
                                       dxv, Idxv, Idxv)

    ! device state (up, vp and the work arrays start at zero, as the
    ! reference's never-written host arrays do)
    d_u = alloc_copy_3d(u, n1, n2, nk); d_v = alloc_copy_3d(v, n1, n2, nk)
    d_h0 = alloc_copy_3d(h0, n1, n2, nk); d_h = alloc_copy_3d(h, n1, n2, nk)
    d_htmp = alloc_zero_3d(n1, n2, nk); d_uh = alloc_zero_3d(n1, n2, nk); d_vh = alloc_zero_3d(n1, n2, nk)
    d_CAu = alloc_zero_3d(n1, n2, nk); d_CAv = alloc_zero_3d(n1, n2, nk)
    d_up = alloc_zero_3d(n1, n2, nk); d_vp = alloc_zero_3d(n1, n2, nk)
    d_diffu = alloc_zero_3d(n1, n2, nk); d_diffv = alloc_zero_3d(n1, n2, nk)
    d_eta = alloc_copy_2d(eta, n1, n2); d_ubt = alloc_copy_2d(ubt, n1, n2); d_vbt = alloc_copy_2d(vbt, n1, n2)
    d_ubt_av = alloc_zero_2d(n1, n2); d_vbt_av = alloc_zero_2d(n1, n2); d_eta_av = alloc_zero_2d(n1, n2)
    d_taux = alloc_copy_2d(taux, n1, n2); d_tauy = alloc_copy_2d(tauy, n1, n2); d_ray = alloc_copy_2d(ray, n1, n2)
    d_dyCu = alloc_copy_2d(dxv, n1, n2); d_dxCv = alloc_copy_2d(dxv, n1, n2)
    d_ray3 = c_null_ptr
    if (opt_ray3d) then   ! Ray(k) = (1e-4 k)/nk at every level, Ray_u = Ray_v
        allocate (ray3(isd:ied, jsd:jed, nk))
        do k = 1, nk
            ray3(:, :, k) = (1.0e-4_dp * real(k, dp)) / real(nk, dp)
        end do
        d_ray3 = alloc_copy_3d(ray3, n1, n2, nk)
    end if
    if (diag_mode > 0) then   ! diag_init (mom6_diag.F90:76-114): work_2d starts at zero
        d_work2d = alloc_zero_2d(n1, n2); d_ke = alloc_zero_3d(n1, n2, nk); d_mass = alloc_zero_2d(n1, n2)
        d_red = alloc_zero_2d(int(mom6h_diag_stats_work_doubles()), 1); d_stats = alloc_zero_2d(4, 1)
    end if
    call hip_check(mom6h_device_synchronize(), 'init')

    t_total = 0.0_dp
    do iter = 1, niter
        call system_clock(c0, rate)
        ! h <- h0 (:222-228) is folded into the predictor continuity call
        ! ---- predictor (:231-305)
        call hip_check(mom6h_transports(isd, ied, jsd, jed, nk, d_u, d_v, d_h0, d_dyCu, d_dxCv, d_uh, d_vh, &
                                        c_null_ptr), 'transports')
        call hor_visc_hip(d_u, d_v, d_h0, d_diffu, d_diffv, hvisc_CS)
        call CorAdCalc_hip(d_u, d_v, d_h0, d_uh, d_vh, d_CAu, d_CAv, cor_CS)
        call hip_check(mom6h_vel_update(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, dt, d_u, d_v, d_CAu, d_diffu, &
                                        d_CAv, d_diffv, d_up, d_vp, c_null_ptr), 'vel_update')
        if (opt_ray3d) then
            call vert_visc_cra_ray3d_hip(d_up, d_vp, d_h0, dt, visc_CS, d_taux, d_tauy, d_ray3, d_ray3)
        else
            call vert_visc_cra_hip(d_up, d_vp, d_h0, dt, visc_CS, d_taux, d_tauy, d_ray, d_ray)
        end if
        call btstep_hip(d_eta, d_ubt, d_vbt, d_ubt_av, d_vbt_av, d_eta_av, bt_CS)
        call continuity_PPM_hip(d_up, d_h0, d_h, d_uh, dt, cont_CS, BT_cont=BT_cont, hin_halo_to_h=.true.)
        if (diag_mode > 0) then   ! predictor diagnostics (:300-308)
            call hip_check(mom6h_diag_product_sum(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, d_diffu, d_h, 0, &
                                                  d_work2d, c_null_ptr), 'diag u')
            call post_stats('diffu_sum', d_work2d, 1)
            call hip_check(mom6h_diag_product_sum(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, d_diffv, d_h, 1, &
                                                  d_work2d, c_null_ptr), 'diag v')
            call post_stats('diffv_sum', d_work2d, 1)
        end if
        ! ---- corrector (:308-392)
        call hip_check(mom6h_transports(isd, ied, jsd, jed, nk, d_up, d_vp, d_h, d_dyCu, d_dxCv, d_uh, d_vh, &
                                        c_null_ptr), 'transports')
        call hor_visc_hip(d_up, d_vp, d_h, d_diffu, d_diffv, hvisc_CS)
        call CorAdCalc_hip(d_up, d_vp, d_h, d_uh, d_vh, d_CAu, d_CAv, cor_CS)
        call hip_check(mom6h_vel_update(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, 0.5_dp * dt, d_u, d_v, d_CAu, &
                                        d_diffu, d_CAv, d_diffv, d_u, d_v, c_null_ptr), 'vel_update')
        if (opt_ray3d) then
            call vert_visc_cra_ray3d_hip(d_u, d_v, d_h, 0.5_dp * dt, visc_CS, d_taux, d_tauy, d_ray3, d_ray3)
        else
            call vert_visc_cra_hip(d_u, d_v, d_h, 0.5_dp * dt, visc_CS, d_taux, d_tauy, d_ray, d_ray)
        end if
        call btstep_hip(d_eta_av, d_ubt_av, d_vbt_av, d_ubt_av, d_vbt_av, d_eta, bt_CS)
        call copy_d2d(d_htmp, d_h, n3v)
        call continuity_PPM_hip(d_u, d_htmp, d_h, d_uh, 0.5_dp * dt, cont_CS, BT_cont=BT_cont)
        if (diag_mode > 0 .and. iter == niter) then   ! end-of-run diagnostics (:386-394)
            call hip_check(mom6h_diag_ke(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, d_u, d_v, d_ke, c_null_ptr), &
                           'diag KE')
            call post_stats('KE', d_ke, nk)
            call hip_check(mom6h_diag_mass(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, d_h, d_mass, c_null_ptr), &
                           'diag mass')
            call post_stats('mass', d_mass, 1)
        end if
        call hip_check(mom6h_device_synchronize(), 'step')
        call system_clock(c1)
        t_iter = real(c1 - c0, dp) / real(rate, dp)
        t_total = t_total + t_iter
        print '(A,I4,A,F10.6,A)', 'RK2 iteration ', iter, ': ', t_iter, ' s'
    end do

    ! verify_state (:563-623)
    call copy_to_host_3d(u, d_u, n1, n2, nk)
    call copy_to_host_3d(v, d_v, n1, n2, nk)
    call copy_to_host_3d(h, d_h, n1, n2, nk)
    call copy_to_host_2d(eta, d_eta, n1, n2)
    mass_init = 0.0_dp; mass_final = 0.0_dp; max_u = 0.0_dp; max_v = 0.0_dp; max_eta = 0.0_dp; ke = 0.0_dp
    do k = 1, nk
        do j = jsc, jec
            do i = isc, iec
                mass_init = mass_init + h0(i, j, k) * areaT(i, j)
                mass_final = mass_final + h(i, j, k) * areaT(i, j)
                max_u = max(max_u, abs(u(i, j, k)))
                max_v = max(max_v, abs(v(i, j, k)))
                ke = ke + 0.5_dp * (u(i, j, k)**2 + v(i, j, k)**2) * h(i, j, k) * areaT(i, j)
            end do
        end do
    end do
    do j = jsc, jec
        do i = isc, iec
            max_eta = max(max_eta, abs(eta(i, j)))
        end do
    end do
    print '(A)', '================================================================'
    print '(A)', 'Verification:'
    print '(A,ES20.12)', '  Initial mass:     ', mass_init
    print '(A,ES20.12)', '  Final mass:       ', mass_final
    print '(A,ES20.12)', '  Relative error:   ', abs(mass_final - mass_init) / mass_init
    print '(A,ES20.12)', '  Max |u|:          ', max_u
    print '(A,ES20.12)', '  Max |v|:          ', max_v
    print '(A,ES20.12)', '  Max |eta|:        ', max_eta
    print '(A,ES20.12)', '  Kinetic energy:   ', ke
    print '(A,F12.6,A)', 'Total time in RK2 iterations: ', t_total, ' s'
    if (niter > 0) print '(A,ES14.6)', 'Cell-updates per second: ', real(ni, dp) * nj * nk * niter / t_total
    if (len_trim(outfile) > 0) then
        open (newunit=i, file=trim(outfile), access='stream', form='unformatted', status='replace', action='write')
        write (i) u, v, h, eta
        close (i)
    end if

    call continuity_end_hip(cont_CS)
    call dealloc_BT_cont_type_hip(BT_cont)
    call coriolis_end_hip(cor_CS)
    call barotropic_end_hip(bt_CS)
    call vert_visc_end_hip(visc_CS)
    call hor_visc_end_hip(hvisc_CS)
    call hip_free(d_u); call hip_free(d_v); call hip_free(d_h); call hip_free(d_h0); call hip_free(d_htmp)
    call hip_free(d_uh); call hip_free(d_vh); call hip_free(d_CAu); call hip_free(d_CAv)
    call hip_free(d_up); call hip_free(d_vp); call hip_free(d_diffu); call hip_free(d_diffv)
    call hip_free(d_eta); call hip_free(d_ubt); call hip_free(d_vbt)
    call hip_free(d_ubt_av); call hip_free(d_vbt_av); call hip_free(d_eta_av)
    call hip_free(d_taux); call hip_free(d_tauy); call hip_free(d_ray); call hip_free(d_dyCu); call hip_free(d_dxCv)
    call hip_free(d_ray3)

contains

    ! post_data_2d/3d with output mode DIAG_STATS (src/common/mom6_diag.F90:166-274):
    ! the statistics of a device field over the compute domain, printed as the reference prints them
    subroutine post_stats(name, field, nlev)
        character(len=*), intent(in) :: name
        type(c_ptr), intent(in) :: field
        integer, intent(in) :: nlev
        real(dp), target :: st(4)
        call hip_check(mom6h_diag_stats(isd, ied, jsd, jed, isc, iec, jsc, jec, nlev, field, d_red, d_stats, &
                                        c_null_ptr), 'diag stats')
        call hip_check(mom6h_device_synchronize(), 'diag stats')
        call hip_check(mom6h_memcpy_d2h(c_loc(st), d_stats, 32_c_size_t), 'diag stats')
        print '(A,A,A,4ES14.6)', '  Diag: ', trim(name), ' min/max/mean/rms: ', st
    end subroutine

    !> The reference driver's barotropic CFL report (rk2_serial_driver.F90:625-662),
    !> numbers from mom6h_bt_cfl
    subroutine check_bt_cfl(dt, nsteps, dx_km, depth)
        real(dp), intent(in) :: dt, dx_km, depth
        integer, intent(in) :: nsteps
        real(c_double) :: cfl, cfl_2d
        integer(c_int) :: min_nsteps, status
        call hip_check(mom6h_bt_cfl(dt, int(nsteps, c_int), dx_km, depth, cfl, cfl_2d, min_nsteps, status), &
                       'mom6h_bt_cfl')
        print '(A)', ''
        print '(A)', 'Barotropic CFL Check:'
        print '(A,F8.1,A)', '  Gravity wave speed: ', sqrt(9.80_dp * depth), ' m/s'
        print '(A,F8.3,A)', '  BT timestep:        ', dt / real(nsteps, dp), ' s'
        print '(A,F8.4)', '  CFL (1D):           ', cfl
        print '(A,F8.4)', '  CFL (2D diagonal):  ', cfl_2d
        if (status == 2) then
            print '(A)', ''
            print '(A)', '  *** WARNING: CFL > 0.9 - UNSTABLE! ***'
            print '(A,I4,A)', '  Recommend at least ', min_nsteps, ' substeps'
        else if (status == 1) then
            print '(A)', ''
            print '(A)', '  * CAUTION: CFL > 0.7 - marginally stable'
            print '(A,I4,A)', '  Recommend at least ', min_nsteps, ' substeps'
        else
            print '(A)', '  CFL OK'
        end if
    end subroutine check_bt_cfl

end program rk2_hip_driver
