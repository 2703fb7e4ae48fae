! mrlbm_iface.f90 -- iso_c_binding interface to the C-ABI of include/mrlbm.h.
! The derived types mirror the C structs field by field (interoperable layout).
module mrlbm_iface
    use iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: MRLBM_FP64 = 0, MRLBM_FP32_DEV = 1
    integer(c_int32_t), parameter :: MRLBM_MASS_STRONG = 0, MRLBM_MASS_EQ = 1, MRLBM_MASS_WEAK = 2
    integer(c_int32_t), parameter :: MRLBM_INLET_UNIFORM = 0, MRLBM_INLET_PARABOLIC = 1
    integer(c_int), parameter :: MRLBM_E_NONFINITE = -4

    type, bind(C) :: mrlbm_params
        integer(c_int32_t) :: nx, ny
        integer(c_int32_t) :: x_periodic, y_periodic
        integer(c_int32_t) :: channel_with_cylinder
        real(c_double) :: x_c, y_c, r_c
        real(c_double) :: Re, uo, omega
        integer(c_int32_t) :: inlet_profile, mass_bc, rotated_coordinate
        integer(c_int32_t) :: vel_interp, mom_interp
        integer(c_int32_t) :: incomp_bc, incomp_bulk
        integer(c_int32_t) :: precision, label_policy, strip_rows, kernel, block_order
        integer(c_int32_t) :: time_blocking, steps_per_pass, quad_rows
        integer(c_int32_t) :: reserved(2)
    end type

    type, bind(C) :: mrlbm_step_opts
        integer(c_int32_t) :: post_process, statsbegin, statsend, iplot
        integer(c_int32_t) :: record_surface_p, stats_interval, nprobe
        integer(c_int32_t) :: probe_i(4), probe_j(4)
        integer(c_int32_t) :: reserved(6)
    end type

    type, bind(C) :: mrlbm_info
        integer(c_int32_t) :: nx, ny, nbct, precision
        real(c_double) :: r_cyl, nu, omega, x_c, y_c, r_c
        integer(c_int64_t) :: bytes_per_step
        integer(c_int32_t) :: box(4), n_stencil, levels, pair_rows, quad_rows
    end type

    type, bind(C) :: mrlbm_adjoint_opts
        integer(c_int32_t) :: iter_first, nsteps, window, checkpoint_interval
        integer(c_int32_t) :: serial_recompute
        integer(c_int32_t) :: reserved(3)
    end type

    type, bind(C) :: mrlbm_adjoint_result
        real(c_double) :: J, dJ_duo, forward_ms, adjoint_ms
        real(c_double) :: reserved(4)
    end type

    interface
        integer(c_int) function mrlbm_create(p, h) bind(C, name='mrlbm_create')
            import :: c_int, c_ptr, mrlbm_params
            type(mrlbm_params), intent(in) :: p
            type(c_ptr), intent(out) :: h
        end function
        integer(c_int) function mrlbm_destroy(h) bind(C, name='mrlbm_destroy')
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
        type(c_ptr) function mrlbm_last_error(h) bind(C, name='mrlbm_last_error')
            import :: c_ptr
            type(c_ptr), value :: h
        end function
        subroutine mrlbm_default_params(p) bind(C, name='mrlbm_default_params')
            import :: mrlbm_params
            type(mrlbm_params), intent(out) :: p
        end subroutine
        subroutine mrlbm_default_step_opts(o) bind(C, name='mrlbm_default_step_opts')
            import :: mrlbm_step_opts
            type(mrlbm_step_opts), intent(out) :: o
        end subroutine
        integer(c_int) function mrlbm_get_info(h, info) bind(C, name='mrlbm_get_info')
            import :: c_int, c_ptr, mrlbm_info
            type(c_ptr), value :: h
            type(mrlbm_info), intent(out) :: info
        end function
        integer(c_int) function mrlbm_get_boundary(h, bi, bj, label, theta, idx) bind(C, name='mrlbm_get_boundary')
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: bi(*), bj(*), label(*), idx(*)
            real(c_double), intent(out) :: theta(*)
        end function
        integer(c_int) function mrlbm_get_node_class(h, cls) bind(C, name='mrlbm_get_node_class')
            import :: c_int, c_ptr, c_int8_t
            type(c_ptr), value :: h
            integer(c_int8_t), intent(out) :: cls(*)
        end function
        integer(c_int) function mrlbm_set_populations(h, f) bind(C, name='mrlbm_set_populations')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: f(*)
        end function
        integer(c_int) function mrlbm_step(h, iter_first, nsteps, opts) bind(C, name='mrlbm_step')
            import :: c_int, c_ptr, c_int32_t, mrlbm_step_opts
            type(c_ptr), value :: h
            integer(c_int32_t), value :: iter_first, nsteps
            type(mrlbm_step_opts), intent(in) :: opts
        end function
        integer(c_int) function mrlbm_get_fields(h, rho, ux, uy) bind(C, name='mrlbm_get_fields')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: rho(*), ux(*), uy(*)
        end function
        integer(c_int) function mrlbm_get_populations(h, f) bind(C, name='mrlbm_get_populations')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: f(*)
        end function
        ! same populations without any allocation (callers that cannot allocate device memory)
        integer(c_int) function mrlbm_get_populations_noalloc(h, f) bind(C, name='mrlbm_get_populations_noalloc')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: f(*)
        end function
        integer(c_int) function mrlbm_get_forces(h, iters, fd, fl, nmax, cnt) bind(C, name='mrlbm_get_forces')
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: iters(*)
            real(c_double), intent(out) :: fd(*), fl(*)
            integer(c_int32_t), value :: nmax
            integer(c_int32_t), intent(out) :: cnt
        end function
        integer(c_int) function mrlbm_get_surface_p_series(h, iters, ps, pfit, nmax, cnt) &
                bind(C, name='mrlbm_get_surface_p_series')
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: iters(*)
            real(c_double), intent(out) :: ps(*), pfit(*)
            integer(c_int32_t), value :: nmax
            integer(c_int32_t), intent(out) :: cnt
        end function
        integer(c_int) function mrlbm_get_probes(h, iters, vals, nmax, cnt) bind(C, name='mrlbm_get_probes')
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: h
            integer(c_int32_t), intent(out) :: iters(*)
            real(c_double), intent(out) :: vals(*)
            integer(c_int32_t), value :: nmax
            integer(c_int32_t), intent(out) :: cnt
        end function
        integer(c_int) function mrlbm_get_rho_inf_mean(h, v) bind(C, name='mrlbm_get_rho_inf_mean')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function mrlbm_set_brinkman(h, chi) bind(C, name='mrlbm_set_brinkman')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(in) :: chi(*)
        end function
        subroutine mrlbm_default_adjoint_opts(o) bind(C, name='mrlbm_default_adjoint_opts')
            import :: mrlbm_adjoint_opts
            type(mrlbm_adjoint_opts), intent(out) :: o
        end subroutine
        integer(c_int) function mrlbm_adjoint_drag(h, o, dchi, res) bind(C, name='mrlbm_adjoint_drag')
            import :: c_int, c_ptr, c_double, mrlbm_adjoint_opts, mrlbm_adjoint_result
            type(c_ptr), value :: h
            type(mrlbm_adjoint_opts), intent(in) :: o
            real(c_double), intent(out) :: dchi(*)
            type(mrlbm_adjoint_result), intent(out) :: res
        end function
        integer(c_int) function mrlbm_max_ux(h, v) bind(C, name='mrlbm_max_ux')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: h
            real(c_double), intent(out) :: v
        end function
        integer(c_int) function mrlbm_check_finite(h) bind(C, name='mrlbm_check_finite')
            import :: c_int, c_ptr
            type(c_ptr), value :: h
        end function
    end interface

contains

    ! Message of the last failure as a Fortran string.
    function mrlbm_error_string(h) result(s)
        type(c_ptr), intent(in) :: h
        character(len=512) :: s
        character(kind=c_char), pointer :: cs(:)
        type(c_ptr) :: p
        integer :: i
        s = ' '
        p = mrlbm_last_error(h)
        if (.not. c_associated(p)) return
        call c_f_pointer(p, cs, [512])
        do i = 1, 512
            if (cs(i) == c_null_char) exit
            s(i:i) = cs(i)
        end do
    end function

end module mrlbm_iface
