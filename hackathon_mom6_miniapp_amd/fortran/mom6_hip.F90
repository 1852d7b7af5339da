!> Fortran host side of libmom6hip (gfx950): iso_c_binding interfaces and the
!! module-level API a Fortran RK2 driver calls.
!!
!! The modules mirror the reference's src/cuda_c Fortran wrappers (one module
!! per physics component, a control structure per module, init / compute / end)
!! with the suffix _hip in place of _cuda_c:
!!   mom6_continuity_hip  <- src/cuda_c/mom6_continuity_cuda_c.F90
!!   mom6_coriolis_hip    <- src/cuda_c/mom6_coriolis_cuda_c.F90
!!   mom6_barotropic_hip  <- src/cuda_c/mom6_barotropic_cuda_c.F90
!!   mom6_vert_visc_hip   <- src/cuda_c/mom6_vert_visc_cuda_c.F90
!!   mom6_hor_visc_hip    <- src/cuda_c/mom6_hor_visc_cuda_c.F90
!!   mom6_hip_common      <- src/cuda_c/mom6_cuda_c_common.F90
!! Init routines take host arrays over (isd:ied, jsd:jed) and copy them to the
!! device; compute routines take device addresses (type(c_ptr)).  Every library
!! call is checked: a failure prints the library's message and stops, where the
!! reference ignores the return codes.  The arithmetic behind each call follows
!! the reference's serial modules (src/serial/*_serial.F90).
module mom6_hip_capi
    use iso_c_binding
    implicit none
    public

    interface
        function mom6h_last_error() bind(c, name='mom6h_last_error') result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        integer(c_int) function mom6h_set_device(dev) bind(c, name='mom6h_set_device')
            import :: c_int
            integer(c_int), value :: dev
        end function
        integer(c_int) function mom6h_device_synchronize() bind(c, name='mom6h_device_synchronize')
            import :: c_int
        end function
        integer(c_int) function mom6h_malloc(p, nbytes) bind(c, name='mom6h_malloc')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), intent(out) :: p
            integer(c_size_t), value :: nbytes
        end function
        integer(c_int) function mom6h_free(p) bind(c, name='mom6h_free')
            import :: c_int, c_ptr
            type(c_ptr), value :: p
        end function
        integer(c_int) function mom6h_memcpy_h2d(dst, src, nbytes) bind(c, name='mom6h_memcpy_h2d')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: nbytes
        end function
        integer(c_int) function mom6h_memcpy_d2h(dst, src, nbytes) bind(c, name='mom6h_memcpy_d2h')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: nbytes
        end function
        integer(c_int) function mom6h_memcpy_d2d(dst, src, nbytes) bind(c, name='mom6h_memcpy_d2d')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: dst, src
            integer(c_size_t), value :: nbytes
        end function
        integer(c_int) function mom6h_memset(p, val, nbytes) bind(c, name='mom6h_memset')
            import :: c_int, c_ptr, c_size_t
            type(c_ptr), value :: p
            integer(c_int), value :: val
            integer(c_size_t), value :: nbytes
        end function

        integer(c_int) function mom6h_continuity_init(cs, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, &
                IareaT, IdxT, dy_Cu, mask2dT, monotonic, dxT, dxCu, mask2dCu) &
                bind(c, name='mom6h_continuity_init')
            import :: c_int, c_ptr, c_double
            type(c_ptr), intent(out) :: cs
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk, monotonic
            real(c_double), intent(in) :: IareaT(*), IdxT(*), dy_Cu(*), mask2dT(*), dxT(*), dxCu(*), mask2dCu(*)
        end function
        integer(c_int) function mom6h_continuity_ppm(cs, u, hin, h, uh, dt, uhbt, bt_cont, u_cor, du_cor, &
                hin_halo_to_h, stream) bind(c, name='mom6h_continuity_ppm')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cs, u, hin, h, uh, uhbt, bt_cont, u_cor, du_cor, stream
            real(c_double), value :: dt
            integer(c_int), value :: hin_halo_to_h
        end function
        integer(c_int) function mom6h_continuity_end(cs) bind(c, name='mom6h_continuity_end')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs
        end function
        integer(c_int) function mom6h_bt_cont_alloc(bt, isd, ied, jsd, jed, nk) bind(c, name='mom6h_bt_cont_alloc')
            import :: c_int, c_ptr
            type(c_ptr), value :: bt
            integer(c_int), value :: isd, ied, jsd, jed, nk
        end function
        integer(c_int) function mom6h_bt_cont_free(bt) bind(c, name='mom6h_bt_cont_free')
            import :: c_int, c_ptr
            type(c_ptr), value :: bt
        end function

        integer(c_int) function mom6h_coriolis_init(cs, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, &
                areaT, IareaBu, CoriolisBu, mask2dBu, dyCv, dxCu, dyCu, dxCv, IdxCu, IdyCv, scheme) &
                bind(c, name='mom6h_coriolis_init')
            import :: c_int, c_ptr, c_double
            type(c_ptr), intent(out) :: cs
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk, scheme
            real(c_double), intent(in) :: areaT(*), IareaBu(*), CoriolisBu(*), mask2dBu(*), dyCv(*), &
                                          dxCu(*), dyCu(*), dxCv(*), IdxCu(*), IdyCv(*)
        end function
        integer(c_int) function mom6h_coradcalc(cs, u, v, h, uh, vh, CAu, CAv, stream) bind(c, name='mom6h_coradcalc')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs, u, v, h, uh, vh, CAu, CAv, stream
        end function
        integer(c_int) function mom6h_coriolis_end(cs) bind(c, name='mom6h_coriolis_end')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs
        end function

        integer(c_int) function mom6h_barotropic_init(cs, isd, ied, jsd, jed, isc, iec, jsc, jec, nstep, dt, &
                bebt, first_direction, Datu, Datv, gtot_E, gtot_W, gtot_N, gtot_S, f_4_u, f_4_v, &
                bt_rem_u, bt_rem_v, IareaT, IdxCu, IdyCv, periodic_x, periodic_y) &
                bind(c, name='mom6h_barotropic_init')
            import :: c_int, c_ptr, c_double
            type(c_ptr), intent(out) :: cs
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nstep, first_direction
            integer(c_int), value :: periodic_x, periodic_y
            real(c_double), value :: dt, bebt
            real(c_double), intent(in) :: Datu(*), Datv(*), gtot_E(*), gtot_W(*), gtot_N(*), gtot_S(*), &
                                          f_4_u(*), f_4_v(*), bt_rem_u(*), bt_rem_v(*), IareaT(*), &
                                          IdxCu(*), IdyCv(*)
        end function
        integer(c_int) function mom6h_btstep(cs, eta_in, ubt_in, vbt_in, u_av, v_av, eta_av, stream) &
                bind(c, name='mom6h_btstep')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs, eta_in, ubt_in, vbt_in, u_av, v_av, eta_av, stream
        end function
        integer(c_int) function mom6h_barotropic_end(cs) bind(c, name='mom6h_barotropic_end')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs
        end function

        integer(c_int) function mom6h_vert_visc_init(cs, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, &
                mask2dCu, mask2dCv, Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl) bind(c, name='mom6h_vert_visc_init')
            import :: c_int, c_ptr, c_double
            type(c_ptr), intent(out) :: cs
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
            real(c_double), intent(in) :: mask2dCu(*), mask2dCv(*)
            real(c_double), value :: Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl
        end function
        integer(c_int) function mom6h_vert_visc(cs, u, v, h, dt, taux, tauy, ray_u, ray_v, stream) &
                bind(c, name='mom6h_vert_visc')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cs, u, v, h, taux, tauy, ray_u, ray_v, stream
            real(c_double), value :: dt
        end function
        integer(c_int) function mom6h_vert_visc_end(cs) bind(c, name='mom6h_vert_visc_end')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs
        end function
        integer(c_int) function mom6h_vert_visc_ray3d(cs, u, v, h, dt, taux, tauy, ray_u, ray_v, stream) &
                bind(c, name='mom6h_vert_visc_ray3d')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cs, u, v, h, taux, tauy, ray_u, ray_v, stream
            real(c_double), value :: dt
        end function
        integer(c_int) function mom6h_vert_visc_keep_coefficients(cs, keep) &
                bind(c, name='mom6h_vert_visc_keep_coefficients')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs
            integer(c_int), value :: keep
        end function

        integer(c_int) function mom6h_hor_visc_init(cs, isd, ied, jsd, jed, isc, iec, jsc, jec, nz, Kh, &
                h_neglect, DY_dxT, DX_dyT, DY_dxBu, DX_dyBu, IdyCu, IdxCu, IdyCv, IdxCv, IareaCu, IareaCv, &
                mask2dT, mask2dBu, reduction_xx, reduction_xy, dy2h, dx2h, dy2q, dx2q) &
                bind(c, name='mom6h_hor_visc_init')
            import :: c_int, c_ptr, c_double
            type(c_ptr), intent(out) :: cs
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nz
            real(c_double), value :: Kh, h_neglect
            real(c_double), intent(in) :: DY_dxT(*), DX_dyT(*), DY_dxBu(*), DX_dyBu(*), IdyCu(*), IdxCu(*), &
                                          IdyCv(*), IdxCv(*), IareaCu(*), IareaCv(*), mask2dT(*), &
                                          mask2dBu(*), reduction_xx(*), reduction_xy(*), dy2h(*), &
                                          dx2h(*), dy2q(*), dx2q(*)
        end function
        integer(c_int) function mom6h_hor_visc_set_options(cs, laplacian, biharmonic, Ah, no_slip, &
                Idx2dyCu, Idxdy2u, Idx2dyCv, Idxdy2v) bind(c, name='mom6h_hor_visc_set_options')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cs
            integer(c_int), value :: laplacian, biharmonic, no_slip
            real(c_double), value :: Ah
            real(c_double), intent(in) :: Idx2dyCu(*), Idxdy2u(*), Idx2dyCv(*), Idxdy2v(*)
        end function
        integer(c_int) function mom6h_hor_visc(cs, u, v, h, diffu, diffv, stream) bind(c, name='mom6h_hor_visc')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs, u, v, h, diffu, diffv, stream
        end function
        integer(c_int) function mom6h_hor_visc_set_closures(cs, smagorinsky, leith, better_bound, areaT, dxBu, &
                dyBu, IdxBu, IdyBu) bind(c, name='mom6h_hor_visc_set_closures')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: cs
            integer(c_int), value :: smagorinsky, leith, better_bound
            real(c_double), intent(in) :: areaT(*), dxBu(*), dyBu(*), IdxBu(*), IdyBu(*)
        end function
        integer(c_int) function mom6h_hor_visc_end(cs) bind(c, name='mom6h_hor_visc_end')
            import :: c_int, c_ptr
            type(c_ptr), value :: cs
        end function

        integer(c_int) function mom6h_transports(isd, ied, jsd, jed, nk, u, v, h, dyCu, dxCv, uh, vh, stream) &
                bind(c, name='mom6h_transports')
            import :: c_int, c_ptr
            integer(c_int), value :: isd, ied, jsd, jed, nk
            type(c_ptr), value :: u, v, h, dyCu, dxCv, uh, vh, stream
        end function
        integer(c_int) function mom6h_vel_update(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, fac_dt, bu, bv, &
                CAu, diffu, CAv, diffv, ou, ov, stream) bind(c, name='mom6h_vel_update')
            import :: c_int, c_ptr, c_double
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
            real(c_double), value :: fac_dt
            type(c_ptr), value :: bu, bv, CAu, diffu, CAv, diffv, ou, ov, stream
        end function
        integer(c_int) function mom6h_halo_fill(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, periodic_x, &
                periodic_y, a, stream) bind(c, name='mom6h_halo_fill')
            import :: c_int, c_ptr
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk, periodic_x, periodic_y
            type(c_ptr), value :: a, stream
        end function
        integer(c_int) function mom6h_bt_cfl(dt, nsteps, dx_km, depth, cfl, cfl_2d, min_nsteps, status) &
                bind(c, name='mom6h_bt_cfl')
            import :: c_int, c_double
            real(c_double), value :: dt, dx_km, depth
            integer(c_int), value :: nsteps
            real(c_double), intent(out) :: cfl, cfl_2d
            integer(c_int), intent(out) :: min_nsteps, status
        end function
        ! diagnostics (src/common/mom6_diag.F90, app/rk2_serial_driver.F90:664-723)
        integer(c_int) function mom6h_diag_product_sum(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, a, b, &
                position, work_2d, stream) bind(c, name='mom6h_diag_product_sum')
            import :: c_int, c_ptr
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk, position
            type(c_ptr), value :: a, b, work_2d, stream
        end function
        integer(c_int) function mom6h_diag_ke(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, u, v, ke, stream) &
                bind(c, name='mom6h_diag_ke')
            import :: c_int, c_ptr
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
            type(c_ptr), value :: u, v, ke, stream
        end function
        integer(c_int) function mom6h_diag_mass(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, h, mass, stream) &
                bind(c, name='mom6h_diag_mass')
            import :: c_int, c_ptr
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
            type(c_ptr), value :: h, mass, stream
        end function
        integer(c_long) function mom6h_diag_stats_work_doubles() bind(c, name='mom6h_diag_stats_work_doubles')
            import :: c_long
        end function
        integer(c_int) function mom6h_diag_stats(isd, ied, jsd, jed, isc, iec, jsc, jec, nk, field, work, out, &
                stream) bind(c, name='mom6h_diag_stats')
            import :: c_int, c_ptr
            integer(c_int), value :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
            type(c_ptr), value :: field, work, out, stream
        end function
    end interface
end module mom6_hip_capi

!> Allocation, transfer and error helpers (mom6_cuda_c_common.F90:74-128)
module mom6_hip_common
    use iso_c_binding
    use iso_fortran_env, only: dp => real64, error_unit
    use mom6_hip_capi
    implicit none
    private
    public :: hip_check, alloc_copy_2d, alloc_copy_3d, alloc_zero_2d, alloc_zero_3d, hip_free, &
              copy_to_host_3d, copy_to_host_2d, copy_d2d

contains

    !> Stops with the library's message when rc is not 0.
    subroutine hip_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer :: n
        if (rc == 0) return
        call c_f_pointer(mom6h_last_error(), msg, [1024])
        n = 0
        do while (n < 1024)
            if (msg(n + 1) == c_null_char) exit
            n = n + 1
        end do
        write (error_unit, '(A,A,I0,A,1024A1)') trim(where), ': libmom6hip error ', rc, ': ', msg(1:n)
        stop 1
    end subroutine

    function alloc_bytes(nbytes) result(p)
        integer(c_size_t), intent(in) :: nbytes
        type(c_ptr) :: p
        call hip_check(mom6h_malloc(p, nbytes), 'mom6h_malloc')
    end function

    function alloc_copy_2d(a, n1, n2) result(p)
        integer, intent(in) :: n1, n2
        real(dp), intent(in), target :: a(n1, n2)
        type(c_ptr) :: p
        integer(c_size_t) :: nb
        nb = int(n1, c_size_t) * int(n2, c_size_t) * 8_c_size_t
        p = alloc_bytes(nb)
        call hip_check(mom6h_memcpy_h2d(p, c_loc(a), nb), 'alloc_copy_2d')
    end function

    function alloc_copy_3d(a, n1, n2, n3) result(p)
        integer, intent(in) :: n1, n2, n3
        real(dp), intent(in), target :: a(n1, n2, n3)
        type(c_ptr) :: p
        integer(c_size_t) :: nb
        nb = int(n1, c_size_t) * int(n2, c_size_t) * int(n3, c_size_t) * 8_c_size_t
        p = alloc_bytes(nb)
        call hip_check(mom6h_memcpy_h2d(p, c_loc(a), nb), 'alloc_copy_3d')
    end function

    function alloc_zero_2d(n1, n2) result(p)
        integer, intent(in) :: n1, n2
        type(c_ptr) :: p
        integer(c_size_t) :: nb
        nb = int(n1, c_size_t) * int(n2, c_size_t) * 8_c_size_t
        p = alloc_bytes(nb)
        call hip_check(mom6h_memset(p, 0_c_int, nb), 'alloc_zero_2d')
    end function

    function alloc_zero_3d(n1, n2, n3) result(p)
        integer, intent(in) :: n1, n2, n3
        type(c_ptr) :: p
        integer(c_size_t) :: nb
        nb = int(n1, c_size_t) * int(n2, c_size_t) * int(n3, c_size_t) * 8_c_size_t
        p = alloc_bytes(nb)
        call hip_check(mom6h_memset(p, 0_c_int, nb), 'alloc_zero_3d')
    end function

    subroutine hip_free(p)
        type(c_ptr), intent(inout) :: p
        if (.not. c_associated(p)) return
        call hip_check(mom6h_free(p), 'mom6h_free')
        p = c_null_ptr
    end subroutine

    subroutine copy_to_host_3d(a, p, n1, n2, n3)
        integer, intent(in) :: n1, n2, n3
        real(dp), intent(out), target :: a(n1, n2, n3)
        type(c_ptr), intent(in) :: p
        call hip_check(mom6h_device_synchronize(), 'copy_to_host_3d')
        call hip_check(mom6h_memcpy_d2h(c_loc(a), p, int(n1, c_size_t) * int(n2, c_size_t) * &
                                        int(n3, c_size_t) * 8_c_size_t), 'copy_to_host_3d')
    end subroutine

    subroutine copy_to_host_2d(a, p, n1, n2)
        integer, intent(in) :: n1, n2
        real(dp), intent(out), target :: a(n1, n2)
        type(c_ptr), intent(in) :: p
        call hip_check(mom6h_device_synchronize(), 'copy_to_host_2d')
        call hip_check(mom6h_memcpy_d2h(c_loc(a), p, int(n1, c_size_t) * int(n2, c_size_t) * 8_c_size_t), &
                       'copy_to_host_2d')
    end subroutine

    subroutine copy_d2d(dst, src, nvals)
        type(c_ptr), intent(in) :: dst, src
        integer(c_size_t), intent(in) :: nvals
        call hip_check(mom6h_memcpy_d2d(dst, src, nvals * 8_c_size_t), 'copy_d2d')
    end subroutine
end module mom6_hip_common

!> continuity_init_cuda_c / continuity_PPM_cuda_c / continuity_end_cuda_c
!! (src/cuda_c/mom6_continuity_cuda_c.F90:269-617)
module mom6_continuity_hip
    use iso_c_binding
    use iso_fortran_env, only: dp => real64
    use mom6_hip_capi
    use mom6_hip_common
    implicit none
    private
    public :: continuity_CS_hip, BT_cont_type_hip, continuity_init_hip, continuity_PPM_hip, &
              continuity_end_hip, alloc_BT_cont_type_hip, dealloc_BT_cont_type_hip

    type :: continuity_CS_hip
        type(c_ptr) :: handle = c_null_ptr
        logical :: initialized = .false.
    end type

    !> BT_cont_type (src/common/mom6_types.F90:89-110) on the device; the layout
    !! of the first seven members is that of the C struct mom6h_bt_cont.
    type, bind(c) :: BT_cont_type_hip
        type(c_ptr) :: FA_u_EE = c_null_ptr, FA_u_E0 = c_null_ptr, FA_u_W0 = c_null_ptr, FA_u_WW = c_null_ptr
        type(c_ptr) :: uBT_WW = c_null_ptr, uBT_EE = c_null_ptr, h_u = c_null_ptr
    end type

contains

    subroutine continuity_init_hip(CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, IareaT, IdxT, dy_Cu, &
                                   mask2dT, monotonic, dxT, areaT, dxCu, mask2dCu)
        type(continuity_CS_hip), intent(inout) :: CS
        integer, intent(in) :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
        real(dp), intent(in) :: IareaT(isd:ied, jsd:jed), IdxT(isd:ied, jsd:jed)
        real(dp), intent(in) :: dy_Cu(isd:ied, jsd:jed), mask2dT(isd:ied, jsd:jed)
        logical, intent(in) :: monotonic
        real(dp), intent(in) :: dxT(isd:ied, jsd:jed), areaT(isd:ied, jsd:jed)
        real(dp), intent(in) :: dxCu(isd:ied, jsd:jed), mask2dCu(isd:ied, jsd:jed)
        integer(c_int) :: mono
        ! areaT is accepted for signature parity with continuity_init_cuda_c; the
        ! PPM path does not use it
        mono = merge(1_c_int, 0_c_int, monotonic)
        call hip_check(mom6h_continuity_init(CS%handle, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, IareaT, IdxT, &
                                             dy_Cu, mask2dT, mono, dxT, dxCu, mask2dCu), 'continuity_init_hip')
        CS%initialized = .true.
    end subroutine

    !> u_p, hin_p -> h_p (compute domain), uh_p.  uhbt_p = c_null_ptr: zero
    !! barotropic-transport target.  hin_halo_to_h: also copy hin into the halo
    !! of h (the driver's h reset folded into the call).
    subroutine continuity_PPM_hip(u_p, hin_p, h_p, uh_p, dt, CS, uhbt_p, BT_cont, u_cor_p, du_cor_p, &
                                  hin_halo_to_h, stream)
        type(c_ptr), intent(in) :: u_p, hin_p, h_p, uh_p
        real(dp), intent(in) :: dt
        type(continuity_CS_hip), intent(inout) :: CS
        type(c_ptr), intent(in), optional :: uhbt_p, u_cor_p, du_cor_p, stream
        type(BT_cont_type_hip), intent(inout), target, optional :: BT_cont
        logical, intent(in), optional :: hin_halo_to_h
        type(c_ptr) :: ub, bt, uc, duc, st
        integer(c_int) :: halo
        ub = c_null_ptr; bt = c_null_ptr; uc = c_null_ptr; duc = c_null_ptr; st = c_null_ptr; halo = 0
        if (present(uhbt_p)) ub = uhbt_p
        if (present(BT_cont)) bt = c_loc(BT_cont)
        if (present(u_cor_p)) uc = u_cor_p
        if (present(du_cor_p)) duc = du_cor_p
        if (present(stream)) st = stream
        if (present(hin_halo_to_h)) halo = merge(1_c_int, 0_c_int, hin_halo_to_h)
        call hip_check(mom6h_continuity_ppm(CS%handle, u_p, hin_p, h_p, uh_p, dt, ub, bt, uc, duc, halo, st), &
                       'continuity_PPM_hip')
    end subroutine

    subroutine continuity_end_hip(CS)
        type(continuity_CS_hip), intent(inout) :: CS
        if (.not. CS%initialized) return
        call hip_check(mom6h_continuity_end(CS%handle), 'continuity_end_hip')
        CS%handle = c_null_ptr
        CS%initialized = .false.
    end subroutine

    subroutine alloc_BT_cont_type_hip(BT_cont, isd, ied, jsd, jed, nz)
        type(BT_cont_type_hip), intent(inout), target :: BT_cont
        integer, intent(in) :: isd, ied, jsd, jed, nz
        call hip_check(mom6h_bt_cont_alloc(c_loc(BT_cont), isd, ied, jsd, jed, nz), 'alloc_BT_cont_type_hip')
    end subroutine

    subroutine dealloc_BT_cont_type_hip(BT_cont)
        type(BT_cont_type_hip), intent(inout), target :: BT_cont
        call hip_check(mom6h_bt_cont_free(c_loc(BT_cont)), 'dealloc_BT_cont_type_hip')
    end subroutine
end module mom6_continuity_hip

!> coriolis_init_cuda_c / CorAdCalc_cuda_c / coriolis_end_cuda_c
!! (src/cuda_c/mom6_coriolis_cuda_c.F90:164-356)
module mom6_coriolis_hip
    use iso_c_binding
    use iso_fortran_env, only: dp => real64
    use mom6_hip_capi
    use mom6_hip_common
    implicit none
    private
    public :: coriolis_CS_hip, coriolis_init_hip, CorAdCalc_hip, coriolis_end_hip
    integer, parameter, public :: SADOURNY75_ENERGY_HIP = 1, ARAKAWA_HSU90_HIP = 2, ARAKAWA_LAMB81_HIP = 3

    type :: coriolis_CS_hip
        type(c_ptr) :: handle = c_null_ptr
        logical :: initialized = .false.
    end type

contains

    subroutine coriolis_init_hip(CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, areaT, IareaBu, CoriolisBu, &
                                 mask2dBu, dyCv, dxCu, dyCu, dxCv, IdxCu, IdyCv, scheme)
        type(coriolis_CS_hip), intent(inout) :: CS
        integer, intent(in) :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
        real(dp), intent(in), dimension(isd:ied, jsd:jed) :: areaT, IareaBu, CoriolisBu, mask2dBu, dyCv, dxCu, &
                                                             dyCu, dxCv, IdxCu, IdyCv
        integer, intent(in), optional :: scheme
        integer(c_int) :: sch
        sch = SADOURNY75_ENERGY_HIP
        if (present(scheme)) sch = scheme
        call hip_check(mom6h_coriolis_init(CS%handle, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, areaT, IareaBu, &
                                           CoriolisBu, mask2dBu, dyCv, dxCu, dyCu, dxCv, IdxCu, IdyCv, sch), &
                       'coriolis_init_hip')
        CS%initialized = .true.
    end subroutine

    subroutine CorAdCalc_hip(u_p, v_p, h_p, uh_p, vh_p, CAu_p, CAv_p, CS, stream)
        type(c_ptr), intent(in) :: u_p, v_p, h_p, uh_p, vh_p, CAu_p, CAv_p
        type(coriolis_CS_hip), intent(inout) :: CS
        type(c_ptr), intent(in), optional :: stream
        type(c_ptr) :: st
        st = c_null_ptr
        if (present(stream)) st = stream
        call hip_check(mom6h_coradcalc(CS%handle, u_p, v_p, h_p, uh_p, vh_p, CAu_p, CAv_p, st), 'CorAdCalc_hip')
    end subroutine

    subroutine coriolis_end_hip(CS)
        type(coriolis_CS_hip), intent(inout) :: CS
        if (.not. CS%initialized) return
        call hip_check(mom6h_coriolis_end(CS%handle), 'coriolis_end_hip')
        CS%handle = c_null_ptr
        CS%initialized = .false.
    end subroutine
end module mom6_coriolis_hip

!> barotropic_init_cuda_c / btstep_cuda_c / barotropic_end_cuda_c
!! (src/cuda_c/mom6_barotropic_cuda_c.F90:226-669)
module mom6_barotropic_hip
    use iso_c_binding
    use iso_fortran_env, only: dp => real64
    use mom6_hip_capi
    use mom6_hip_common
    implicit none
    private
    public :: barotropic_CS_hip, barotropic_init_hip, btstep_hip, barotropic_end_hip

    type :: barotropic_CS_hip
        type(c_ptr) :: handle = c_null_ptr
        logical :: initialized = .false.
    end type

contains

    !> periodic_x/periodic_y (default .false.) select the halo-refilled ranges of
    !! the reference's split API (btstep_cuda_c_do_step, :464-562).
    subroutine barotropic_init_hip(CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nstep, dt, bebt, &
                                   first_direction, Datu, Datv, gtot_E, gtot_W, gtot_N, gtot_S, f_4_u, &
                                   f_4_v, bt_rem_u, bt_rem_v, IareaT, IdxCu, IdyCv, periodic_x, periodic_y)
        type(barotropic_CS_hip), intent(inout) :: CS
        integer, intent(in) :: isd, ied, jsd, jed, isc, iec, jsc, jec, nstep, first_direction
        real(dp), intent(in) :: dt, bebt
        real(dp), intent(in), dimension(isd:ied, jsd:jed) :: Datu, Datv, gtot_E, gtot_W, gtot_N, gtot_S, &
                                                             bt_rem_u, bt_rem_v, IareaT, IdxCu, IdyCv
        real(dp), intent(in) :: f_4_u(4, isd:ied, jsd:jed), f_4_v(4, isd:ied, jsd:jed)
        logical, intent(in), optional :: periodic_x, periodic_y
        integer(c_int) :: px, py
        px = 0; py = 0
        if (present(periodic_x)) px = merge(1_c_int, 0_c_int, periodic_x)
        if (present(periodic_y)) py = merge(1_c_int, 0_c_int, periodic_y)
        call hip_check(mom6h_barotropic_init(CS%handle, isd, ied, jsd, jed, isc, iec, jsc, jec, nstep, dt, bebt, &
                                             first_direction, Datu, Datv, gtot_E, gtot_W, gtot_N, gtot_S, &
                                             f_4_u, f_4_v, bt_rem_u, bt_rem_v, IareaT, IdxCu, IdyCv, px, py), &
                       'barotropic_init_hip')
        CS%initialized = .true.
    end subroutine

    !> The inputs may alias the outputs (the driver's corrector call).
    subroutine btstep_hip(eta_in_p, ubt_in_p, vbt_in_p, u_av_p, v_av_p, eta_av_p, CS, stream)
        type(c_ptr), intent(in) :: eta_in_p, ubt_in_p, vbt_in_p, u_av_p, v_av_p, eta_av_p
        type(barotropic_CS_hip), intent(inout) :: CS
        type(c_ptr), intent(in), optional :: stream
        type(c_ptr) :: st
        st = c_null_ptr
        if (present(stream)) st = stream
        call hip_check(mom6h_btstep(CS%handle, eta_in_p, ubt_in_p, vbt_in_p, u_av_p, v_av_p, eta_av_p, st), &
                       'btstep_hip')
    end subroutine

    subroutine barotropic_end_hip(CS)
        type(barotropic_CS_hip), intent(inout) :: CS
        if (.not. CS%initialized) return
        call hip_check(mom6h_barotropic_end(CS%handle), 'barotropic_end_hip')
        CS%handle = c_null_ptr
        CS%initialized = .false.
    end subroutine
end module mom6_barotropic_hip

!> vert_visc_init_cuda_c / vert_visc_cra_cuda_c / vert_visc_end_cuda_c
!! (src/cuda_c/mom6_vert_visc_cuda_c.F90:89-210), with the Rayleigh drag of the
!! serial module (bottom-layer coefficients Ray_u(:,:,nz), Ray_v(:,:,nz)).
module mom6_vert_visc_hip
    use iso_c_binding
    use iso_fortran_env, only: dp => real64
    use mom6_hip_capi
    use mom6_hip_common
    implicit none
    private
    public :: vert_visc_CS_hip, vert_visc_init_hip, vert_visc_cra_hip, vert_visc_end_hip
    public :: vert_visc_keep_coefficients_hip, vert_visc_cra_ray3d_hip

    type :: vert_visc_CS_hip
        type(c_ptr) :: handle = c_null_ptr
        logical :: initialized = .false.
    end type

contains

    subroutine vert_visc_init_hip(CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, mask2dCu, mask2dCv, Kv, Kv_ml, &
                                  Kv_extra_bbl, Hmix, Hbbl)
        type(vert_visc_CS_hip), intent(inout) :: CS
        integer, intent(in) :: isd, ied, jsd, jed, isc, iec, jsc, jec, nk
        real(dp), intent(in) :: mask2dCu(isd:ied, jsd:jed), mask2dCv(isd:ied, jsd:jed)
        real(dp), intent(in) :: Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl
        call hip_check(mom6h_vert_visc_init(CS%handle, isd, ied, jsd, jed, isc, iec, jsc, jec, nk, mask2dCu, &
                                            mask2dCv, Kv, Kv_ml, Kv_extra_bbl, Hmix, Hbbl), 'vert_visc_init_hip')
        CS%initialized = .true.
    end subroutine

    !> u_p, v_p updated in place; ray_u_p/ray_v_p: 2D bottom-layer Rayleigh
    !! coefficients (c_null_ptr = none).
    subroutine vert_visc_cra_hip(u_p, v_p, h_p, dt, CS, taux_p, tauy_p, ray_u_p, ray_v_p, stream)
        type(c_ptr), intent(in) :: u_p, v_p, h_p
        real(dp), intent(in) :: dt
        type(vert_visc_CS_hip), intent(inout) :: CS
        type(c_ptr), intent(in) :: taux_p, tauy_p
        type(c_ptr), intent(in), optional :: ray_u_p, ray_v_p, stream
        type(c_ptr) :: ru, rv, st
        ru = c_null_ptr; rv = c_null_ptr; st = c_null_ptr
        if (present(ray_u_p)) ru = ray_u_p
        if (present(ray_v_p)) rv = ray_v_p
        if (present(stream)) st = stream
        call hip_check(mom6h_vert_visc(CS%handle, u_p, v_p, h_p, dt, taux_p, tauy_p, ru, rv, st), 'vert_visc_cra_hip')
    end subroutine

    !> vert_visc_cra_hip with the serial module's general Rayleigh drag: ray_u3d_p,
    !! ray_v3d_p are 3D (visc%Ray_u, visc%Ray_v of every level, mom6_types.F90:120-125)
    subroutine vert_visc_cra_ray3d_hip(u_p, v_p, h_p, dt, CS, taux_p, tauy_p, ray_u3d_p, ray_v3d_p, stream)
        type(c_ptr), intent(in) :: u_p, v_p, h_p, taux_p, tauy_p, ray_u3d_p, ray_v3d_p
        real(dp), intent(in) :: dt
        type(vert_visc_CS_hip), intent(inout) :: CS
        type(c_ptr), intent(in), optional :: stream
        type(c_ptr) :: st
        st = c_null_ptr
        if (present(stream)) st = stream
        call hip_check(mom6h_vert_visc_ray3d(CS%handle, u_p, v_p, h_p, dt, taux_p, tauy_p, ray_u3d_p, ray_v3d_p, &
                                             st), 'vert_visc_cra_ray3d_hip')
    end subroutine

    !> keep = .false.: the coupling coefficients and effective thicknesses
    !! (the control structure's a_u, a_v, h_u, h_v) are not stored, for callers
    !! that never read them
    subroutine vert_visc_keep_coefficients_hip(CS, keep)
        type(vert_visc_CS_hip), intent(inout) :: CS
        logical, intent(in) :: keep
        integer(c_int) :: k
        k = 0
        if (keep) k = 1
        call hip_check(mom6h_vert_visc_keep_coefficients(CS%handle, k), 'vert_visc_keep_coefficients_hip')
    end subroutine

    subroutine vert_visc_end_hip(CS)
        type(vert_visc_CS_hip), intent(inout) :: CS
        if (.not. CS%initialized) return
        call hip_check(mom6h_vert_visc_end(CS%handle), 'vert_visc_end_hip')
        CS%handle = c_null_ptr
        CS%initialized = .false.
    end subroutine
end module mom6_vert_visc_hip

!> hor_visc_init_cuda_c / hor_visc_cuda_c / hor_visc_end_cuda_c
!! (src/cuda_c/mom6_hor_visc_cuda_c.F90:117-298), plus the serial module's
!! biharmonic option (hor_visc_init biharmonic=, Kh4).
module mom6_hor_visc_hip
    use iso_c_binding
    use iso_fortran_env, only: dp => real64
    use mom6_hip_capi
    use mom6_hip_common
    implicit none
    private
    public :: hor_visc_CS_hip, hor_visc_init_hip, hor_visc_hip, hor_visc_end_hip, hor_visc_set_closures_hip

    type :: hor_visc_CS_hip
        type(c_ptr) :: handle = c_null_ptr
        logical :: initialized = .false.
    end type

contains

    subroutine hor_visc_init_hip(CS, isd, ied, jsd, jed, isc, iec, jsc, jec, nz, Kh, h_neglect, DY_dxT, DX_dyT, &
                                 DY_dxBu, DX_dyBu, IdyCu, IdxCu, IdyCv, IdxCv, IareaCu, IareaCv, mask2dT, &
                                 mask2dBu, reduction_xx, reduction_xy, dy2h, dx2h, dy2q, dx2q, Ah, Idx2dyCu, &
                                 Idxdy2u, Idx2dyCv, Idxdy2v, laplacian, no_slip)
        type(hor_visc_CS_hip), intent(inout) :: CS
        integer, intent(in) :: isd, ied, jsd, jed, isc, iec, jsc, jec, nz
        real(dp), intent(in) :: Kh, h_neglect
        real(dp), intent(in), dimension(isd:ied, jsd:jed) :: DY_dxT, DX_dyT, DY_dxBu, DX_dyBu, IdyCu, IdxCu, IdyCv, &
                                                             IdxCv, IareaCu, IareaCv, mask2dT, mask2dBu, &
                                                             reduction_xx, reduction_xy, dy2h, dx2h, dy2q, dx2q
        real(dp), intent(in), optional :: Ah
        real(dp), intent(in), dimension(isd:ied, jsd:jed), optional :: Idx2dyCu, Idxdy2u, Idx2dyCv, Idxdy2v
        logical, intent(in), optional :: laplacian, no_slip
        integer(c_int) :: lap, bih, ns
        call hip_check(mom6h_hor_visc_init(CS%handle, isd, ied, jsd, jed, isc, iec, jsc, jec, nz, Kh, h_neglect, &
                                           DY_dxT, DX_dyT, DY_dxBu, DX_dyBu, IdyCu, IdxCu, IdyCv, IdxCv, IareaCu, &
                                           IareaCv, mask2dT, mask2dBu, reduction_xx, reduction_xy, dy2h, dx2h, &
                                           dy2q, dx2q), 'hor_visc_init_hip')
        CS%initialized = .true.
        lap = 1; bih = 0; ns = 0
        if (present(laplacian)) lap = merge(1_c_int, 0_c_int, laplacian)
        if (present(no_slip)) ns = merge(1_c_int, 0_c_int, no_slip)
        if (present(Ah)) then
            if (.not. (present(Idx2dyCu) .and. present(Idxdy2u) .and. present(Idx2dyCv) .and. present(Idxdy2v))) &
                error stop 'hor_visc_init_hip: Ah needs Idx2dyCu, Idxdy2u, Idx2dyCv, Idxdy2v'
            bih = 1
            call hip_check(mom6h_hor_visc_set_options(CS%handle, lap, bih, Ah, ns, Idx2dyCu, Idxdy2u, Idx2dyCv, &
                                                      Idxdy2v), 'hor_visc_init_hip')
        else if (lap == 0 .or. ns /= 0) then
            call hip_check(mom6h_hor_visc_set_options(CS%handle, lap, bih, 0.0_dp, ns, IdyCu, IdyCu, IdyCu, IdyCu), &
                           'hor_visc_init_hip')
        end if
    end subroutine

    subroutine hor_visc_hip(u_p, v_p, h_p, diffu_p, diffv_p, CS, stream)
        type(c_ptr), intent(in) :: u_p, v_p, h_p, diffu_p, diffv_p
        type(hor_visc_CS_hip), intent(inout) :: CS
        type(c_ptr), intent(in), optional :: stream
        type(c_ptr) :: st
        st = c_null_ptr
        if (present(stream)) st = stream
        call hip_check(mom6h_hor_visc(CS%handle, u_p, v_p, h_p, diffu_p, diffv_p, st), 'hor_visc_hip')
    end subroutine

    !> hor_visc_init's Smagorinsky=, Leith= and better_bound= options
    !! (mom6_hor_visc_serial.F90:160-193): call after hor_visc_init_hip and
    !! before the first hor_visc_hip; the q-point metrics are read by Leith only
    subroutine hor_visc_set_closures_hip(CS, isd, ied, jsd, jed, smagorinsky, leith, better_bound, areaT, dxBu, &
                                         dyBu, IdxBu, IdyBu)
        type(hor_visc_CS_hip), intent(inout) :: CS
        integer, intent(in) :: isd, ied, jsd, jed
        logical, intent(in) :: smagorinsky, leith, better_bound
        real(dp), intent(in), dimension(isd:ied, jsd:jed) :: areaT, dxBu, dyBu, IdxBu, IdyBu
        call hip_check(mom6h_hor_visc_set_closures(CS%handle, merge(1_c_int, 0_c_int, smagorinsky), &
                                                   merge(1_c_int, 0_c_int, leith), &
                                                   merge(1_c_int, 0_c_int, better_bound), areaT, dxBu, dyBu, &
                                                   IdxBu, IdyBu), 'hor_visc_set_closures_hip')
    end subroutine

    subroutine hor_visc_end_hip(CS)
        type(hor_visc_CS_hip), intent(inout) :: CS
        if (.not. CS%initialized) return
        call hip_check(mom6h_hor_visc_end(CS%handle), 'hor_visc_end_hip')
        CS%handle = c_null_ptr
        CS%initialized = .false.
    end subroutine
end module mom6_hor_visc_hip
