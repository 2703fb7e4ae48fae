! mrlbm_driver.f90 -- drop-in driver for the reference time loop (sakthi-lbm/mrlbm_2d,
! program lbm_2d in 2d_mrlbm.f90).  It reads the same input.dat namelists, writes the same
! output files (output/, restart/, data_mean/, data_probe/, simulation_parameters.dat,
! rho_mean_inlet.dat) and replaces the body of the main loop (2d_mrlbm.f90:229-555) by
! calls into the gfx950 HIP library through iso_c_binding (mrlbm_iface).
!
! Differences from the reference, by design:
!  * the MLUPS log uses the wall clock (the reference's cpu_time sums all OpenMP threads);
!  * the initial plot3d file name carries no trailing NUL character;
!  * f_eq in restart.dat is written as zeros (the reference never assigns it);
!  * an optional &Accelerator group selects the precision form and inlet profile.
program mrlbm_gpu
    use iso_c_binding
    use mrlbm_iface
    implicit none

    real(8), parameter :: PI = 4.0d0*atan(1.0d0)
    integer :: num_threads, lattice_type
    integer :: iplot, max_iter, isave, irestart, statsbegin, statsend, iplotstats
    integer :: restart_step = 1
    integer :: nx, ny, ncy, L_front, L_back, L_top, L_bot
    integer :: nfront, nback, ntop, nbot, nbct
    real(8) :: Re, uo, rho_infty
    logical :: x_periodic, y_periodic, channel_with_cylinder, incomp_bc, incomp_bulk, channel_with_square
    logical :: vel_interp, mom_interp, rotated_coordinate, post_process
    logical :: strong_mass_bc, eq_mass_bc, weak_mass_bc
    integer :: precision = 0, inlet_profile = 0, strip_rows = 0, kernel = 0, time_blocking = -1
    integer :: steps_per_pass = 0
    integer :: quad_rows = 0
    logical :: record_surface_p = .true.

    integer :: i_probe(4), j_probe(4)
    real(8) :: r_cyl, nu, omega, delt_phy, L_lat, L_phy, nu_phy
    real(8), allocatable :: rho(:,:), ux(:,:), uy(:,:), f(:,:,:)
    integer(c_int32_t), allocatable :: bou_i(:), bou_j(:), label_bc(:), idx(:)
    real(8), allocatable :: theta_c(:), thetas(:)
    type(mrlbm_params) :: par
    type(mrlbm_step_opts) :: opts
    type(mrlbm_info) :: info
    type(c_ptr) :: h
    integer(c_int) :: rc
    integer :: iter, e, k
    integer(8) :: clk0, clk1, clkrate
    real(8) :: wall, mlups, maxux, rho_inf_mean
    character(len=100) :: filename_bin
    character(len=9) :: plabel
    integer, parameter :: unit_p = 301, unit_f = 302
    integer, parameter :: CHUNK = 256

    call read_input_file()

    if (channel_with_cylinder) then          ! 2d_mrlbm.f90:86-92
        nfront = L_front*ncy
        nback = L_back*ncy
        ntop = L_top*ncy
        nbot = L_bot*ncy
        nx = nfront + ncy + nback
        ny = ntop + ncy + nbot
    else
        nfront = 0; nback = 0; ntop = 0; nbot = 0
    end if

    if (irestart == 1) call check_restart_file()   ! before any device work

    call mrlbm_default_params(par)
    par%nx = nx
    par%ny = ny
    par%x_periodic = merge(1, 0, x_periodic)
    par%y_periodic = merge(1, 0, y_periodic)
    par%channel_with_cylinder = merge(1, 0, channel_with_cylinder)
    par%x_c = real(nfront + (ncy/2) + 1)     ! 2d_mrlbm.f90:946-948
    par%y_c = real(nbot + ncy/2)
    par%r_c = real(ncy/2)
    par%Re = Re
    par%uo = uo
    par%omega = 0.0d0
    par%inlet_profile = inlet_profile
    par%mass_bc = MRLBM_MASS_STRONG
    if (eq_mass_bc) par%mass_bc = MRLBM_MASS_EQ
    if (weak_mass_bc) par%mass_bc = MRLBM_MASS_WEAK
    par%rotated_coordinate = merge(1, 0, rotated_coordinate)
    par%vel_interp = merge(1, 0, vel_interp)
    par%mom_interp = merge(1, 0, mom_interp)
    par%incomp_bc = merge(1, 0, incomp_bc)
    par%incomp_bulk = merge(1, 0, incomp_bulk)
    par%precision = precision
    par%strip_rows = strip_rows
    par%kernel = kernel
    par%time_blocking = time_blocking
    par%steps_per_pass = steps_per_pass
    par%quad_rows = quad_rows
    rc = mrlbm_create(par, h)
    if (rc /= 0) then
        write(*, *) 'mrlbm_create failed: ', trim(mrlbm_error_string(c_null_ptr))
        stop 1
    end if
    rc = mrlbm_get_info(h, info)
    nbct = info%nbct
    r_cyl = info%r_cyl
    nu = info%nu
    omega = info%omega

    allocate(rho(nx, ny), ux(nx, ny), uy(nx, ny), f(nx, ny, 0:8))
    allocate(bou_i(max(nbct, 1)), bou_j(max(nbct, 1)), label_bc(max(nbct, 1)), idx(max(nbct, 1)))
    allocate(theta_c(max(nbct, 1)), thetas(max(nbct, 1)))
    if (nbct > 0) then
        rc = mrlbm_get_boundary(h, bou_i, bou_j, label_bc, theta_c, idx)
        do k = 1, nbct
            thetas(k) = theta_c(idx(k))
        end do
        open(unit=404, file='data_mean/theta.dat')      ! 2d_mrlbm.f90:1158-1162
        do k = 1, nbct
            write(404, *) thetas(k)
        end do
        close(404)
        open(unit=10, file='data_geo/bound.dat', status='replace')
        do k = 1, nbct
            write(10, *) real(bou_i(k), 8), real(bou_j(k), 8)
        end do
        close(10)
    end if
    call write_grid_plot3d()

    ! probes (2d_mrlbm.f90:104-118)
    i_probe(1) = (nfront + ncy + 1) + (ncy/2);   j_probe(1) = (nbot + (ncy/2) + 1)
    i_probe(2) = (nfront + ncy + 1) + (ncy);     j_probe(2) = (nbot + (ncy/2) + 1)
    i_probe(3) = (nfront + ncy + 1) + (3*ncy/2); j_probe(3) = (nbot + (ncy/2) + 1)
    i_probe(4) = (nfront + ncy + 1) + (2*ncy);   j_probe(4) = (nbot + (ncy/2) + 1)
    open(unit=100, file="data_probe/probe_loc.dat")
        do k = 1, 4
            write(plabel, '(a6,i1,a2)') 'probe ', k, ': '   ! one string, as the reference's "probe 1: "
            write(100, *) plabel, i_probe(k), j_probe(k)
        end do
    close(100)

    ! physical time step for master.p3d (2d_mrlbm.f90:137-150)
    L_lat = ny - 1
    if (channel_with_cylinder) L_lat = 2.0d0*r_cyl
    L_phy = 1.0d0
    nu_phy = 0.01d0
    delt_phy = ((L_phy/L_lat)**2)*nu/nu_phy
    call create_master_p3d_file()

    open(unit=10, file='simulation_parameters.dat')      ! 2d_mrlbm.f90:157-168
        write(10,*) 'nx,ny', nx, ny
        write(10,*) 'No. of points on diameter', ncy
        write(10,*) 'centre_cylin', nfront + (ncy/2), nbot + (ncy/2) - 1
        write(10,*) 'Maximum_radius', r_cyl
        write(10,*) 'No. of points on cylinder:', nbct
        write(10,*) 'Re:', Re
        write(10,*) 'uo:', uo
        write(10,*) 'nu:', nu
        write(10,*) 'tau:', 1.0d0/omega
        write(10,*) 'dt_phy:', delt_phy
    close(10)

    if (irestart == 0) then
        rho = 1.0d0
        ux = 0.0d0
        uy = 0.0d0
        write(filename_bin, '(a4,I8,a4)') 'grid', 10000000, '.bin'
        call write_function_plot3d(filename_bin)
    else
        call read_restart_file(iter)
        restart_step = iter
        write(*, *) 'Restarting step = ', restart_step
        rc = mrlbm_set_populations(h, f)
        call check(rc)
    end if

    open(unit=unit_p, file='data_mean/p_surface_time.dat', status='replace'); close(unit_p)
    open(unit=unit_p, file='data_mean/p_surface_smooth_time.dat', status='replace'); close(unit_p)
    open(unit=unit_f, file='data_mean/forces_time.dat', status='replace'); close(unit_f)
    open(unit=310, file='rho_mean_inlet.dat', status='replace'); close(310)
    if (post_process) then
        open(unit=101, file="data_probe/p_probe.dat", status='replace')
        open(unit=102, file="data_probe/ux_probe.dat", status='replace')
        open(unit=103, file="data_probe/uy_probe.dat", status='replace')
    end if

    call mrlbm_default_step_opts(opts)
    opts%post_process = merge(1, 0, post_process)
    opts%statsbegin = statsbegin
    opts%statsend = statsend
    opts%iplot = merge(iplot, 0, channel_with_cylinder .or. channel_with_square)
    opts%record_surface_p = merge(1, 0, record_surface_p .and. post_process)
    opts%nprobe = 0
    if (post_process .and. channel_with_cylinder) then
        opts%nprobe = 4
        opts%probe_i = i_probe
        opts%probe_j = j_probe
    end if

    call system_clock(clk0, clkrate)
    iter = restart_step + 1
    do while (iter <= max_iter)
        e = next_event(iter)
        rc = mrlbm_step(h, iter, e - iter + 1, opts)
        call check(rc)
        call write_records()
        ! The reference never checks (NaN propagates into every file silently): at each
        ! log, plot and restart step the state is scanned on the device, and the run stops
        ! with exit status 3 before any field file of a non-finite state is written.
        if (mod(e, 100) == 0 .or. mod(e, max(isave, 1)) == 0 .or. &
            (post_process .and. mod(e, max(iplot, 1)) == 0) .or. e == max_iter) then
            rc = mrlbm_check_finite(h)
            if (rc == MRLBM_E_NONFINITE) then
                write(*, *) 'STEP: ', e, '     non-finite state (NaN/Inf): stopping'
                stop 3
            end if
            call check(rc)
        end if
        if (post_process) then
            if (e >= statsbegin .and. e <= statsend .and. mod(e, 100) == 0) then
                rc = mrlbm_get_rho_inf_mean(h, rho_inf_mean)
                open(unit=310, file='rho_mean_inlet.dat', access='append')
                    write(310,*) e, rho_inf_mean
                close(310)
            end if
            if (mod(e, iplot) == 0) then
                rc = mrlbm_get_fields(h, rho, ux, uy)
                call check(rc)
                write(filename_bin, '(a4,I8,a4)') 'grid', 10000000 + e, '.bin'
                call write_function_plot3d(filename_bin)
            end if
        end if
        if (mod(e, 100) == 0) then
            call system_clock(clk1)
            wall = real(clk1 - clk0, 8)/real(clkrate, 8)
            clk0 = clk1
            mlups = (real(nx, 8)*real(ny, 8)*100.0d0)/(1.0d6*wall)
            rc = mrlbm_max_ux(h, maxux)
            write(*, *) "STEP: ", e, "     ", "MAX_ux: ", sngl(maxux), "     ", "MLUPS: ", sngl(mlups)
        end if
        if (mod(e, isave) == 0) call write_restart_file(e)
        iter = e + 1
    end do
    if (post_process) then
        close(101); close(102); close(103)
    end if
    if (post_process .and. channel_with_cylinder) call write_input_stat()
    rc = mrlbm_destroy(h)

contains

    ! The post-processor's input for this run, in the input_stat.dat format
    ! (2d_statistics.f90:29-32): boundary-node count, cylinder radius, inlet velocity and the
    ! running mean of the inlet density.
    subroutine write_input_stat()
        type(mrlbm_info) :: inf
        rc = mrlbm_get_info(h, inf)
        call check(rc)
        rc = mrlbm_get_rho_inf_mean(h, rho_inf_mean)
        call check(rc)
        open(unit=311, file='data_mean/input_stat.dat', status='replace')
            write(311, '(a)') '&parameters'
            write(311, '(a,i0)') '    nbct = ', inf%nbct
            write(311, '(a,es25.17)') '    r_cy = ', inf%r_cyl
            write(311, '(a,es25.17)') '    uo = ', uo
            write(311, '(a,es25.17)') '    rho_infty = ', rho_inf_mean
            write(311, '(a)') '/'
        close(311)
    end subroutine

    subroutine check(code)
        integer(c_int), intent(in) :: code
        if (code /= 0) then
            write(*, *) 'mrlbm error ', code, ': ', trim(mrlbm_error_string(h))
            stop 2
        end if
    end subroutine

    ! Last iteration of the next chunk: stop at every iteration where the reference writes
    ! something (logs every 100, plots every iplot, restarts every isave) and at max_iter.
    integer function next_event(it)
        integer, intent(in) :: it
        integer :: cand
        cand = min(max_iter, it + CHUNK - 1)
        cand = min(cand, ((it + 99)/100)*100)
        if (post_process .and. iplot > 0) cand = min(cand, ((it + iplot - 1)/iplot)*iplot)
        if (isave > 0) cand = min(cand, ((it + isave - 1)/isave)*isave)
        next_event = max(cand, it)
    end function

    subroutine write_records()
        integer(c_int32_t) :: its(CHUNK), cnt, cntp
        real(8) :: fd(CHUNK), fl(CHUNK), pv(12*CHUNK)
        real(8), allocatable :: ps(:), pf(:)
        integer :: s, k2
        if (.not. post_process) return
        rc = mrlbm_get_forces(h, its, fd, fl, CHUNK, cnt)
        if (cnt > 0 .and. channel_with_cylinder) then
            open(unit=unit_f, file='data_mean/forces_time.dat', status='unknown', position='append', action='write')
            do s = 1, cnt
                write(unit_f,*) its(s), fd(s), fl(s)
            end do
            close(unit_f)
        end if
        rc = mrlbm_get_probes(h, its, pv, CHUNK, cntp)
        do s = 1, cntp
            write(101,*) its(s), (pv(12*(s-1) + 3*(k2-1) + 1)/3.0d0, k2 = 1, 4)
            write(102,*) its(s), (pv(12*(s-1) + 3*(k2-1) + 2)/3.0d0, k2 = 1, 4)
            write(103,*) its(s), (pv(12*(s-1) + 3*(k2-1) + 3)/3.0d0, k2 = 1, 4)
        end do
        if (record_surface_p .and. nbct > 0) then
            allocate(ps(nbct*CHUNK), pf(nbct*CHUNK))
            rc = mrlbm_get_surface_p_series(h, its, ps, pf, CHUNK, cnt)
            if (cnt > 0) then
                open(unit=unit_p, file='data_mean/p_surface_time.dat', status='unknown', position='append', &
                     action='write', form='unformatted')
                do s = 1, cnt
                    write(unit_p) its(s), (ps(nbct*(s-1) + k2), k2 = 1, nbct)
                end do
                close(unit_p)
                open(unit=unit_p, file='data_mean/p_surface_smooth_time.dat', status='unknown', position='append', &
                     action='write', form='unformatted')
                do s = 1, cnt
                    write(unit_p) its(s), (pf(nbct*(s-1) + k2), k2 = 1, nbct)
                end do
                close(unit_p)
            end if
            deallocate(ps, pf)
        end if
    end subroutine

    subroutine write_grid_plot3d()      ! 2d_mrlbm.f90:2133-2157
        integer :: i, j
        open(unit=100, form='unformatted', file='output/grid.x')
            write(100) 1
            write(100) nx, ny
            write(100) ((sngl(real(i - 1, 8)), i = 1, nx), j = 1, ny), ((sngl(real(j - 1, 8)), i = 1, nx), j = 1, ny)
        close(100)
    end subroutine

    subroutine write_function_plot3d(fname)      ! 2d_mrlbm.f90:2160-2184
        character(len=*), intent(in) :: fname
        integer :: i, j
        open(unit=200, form='unformatted', file='output/'//trim(fname))
            write(200) 1
            write(200) nx, ny, 3
            write(200) ((sngl(rho(i, j)), i = 1, nx), j = 1, ny), ((sngl(ux(i, j)), i = 1, nx), j = 1, ny), &
                       ((sngl(uy(i, j)), i = 1, nx), j = 1, ny)
        close(200)
    end subroutine

    subroutine write_restart_file(soln_time)      ! 2d_mrlbm.f90:2186-2202
        integer, intent(in) :: soln_time
        real(8), allocatable :: f_eq(:,:,:)
        write(*, *) "Writing restart file at", soln_time, "iteration"
        rc = mrlbm_get_populations(h, f)
        call check(rc)
        rc = mrlbm_get_fields(h, rho, ux, uy)
        call check(rc)
        allocate(f_eq(nx, ny, 0:8))
        f_eq = 0.0d0
        open(unit=10000, file='restart/restart.dat', status="unknown", form="unformatted")
            write(10000) soln_time, f, f_eq, ux, uy, rho
        close(10000)
        deallocate(f_eq)
    end subroutine

    subroutine read_restart_file(soln_time)      ! 2d_mrlbm.f90:2204-2220
        integer, intent(out) :: soln_time
        real(8), allocatable :: f_eq(:,:,:)
        integer :: ios
        allocate(f_eq(nx, ny, 0:8))
        open(unit=10000, file='restart/restart.dat', status="old", form="unformatted", iostat=ios)
        if (ios == 0) read(10000, iostat=ios) soln_time, f, f_eq, ux, uy, rho
        close(10000)
        deallocate(f_eq)
        if (ios /= 0) then
            write(*, '(a)') "Error while reading restart/restart.dat: truncated, or written for another grid"
            stop 1
        end if
    end subroutine

    ! The reference opens restart.dat and reads one record (2d_mrlbm.f90:2208-2216) with no
    ! check; a short file then fails inside the Fortran runtime.  Here the file must exist and
    ! hold at least the record's payload (int32 soln_time and 21 nx*ny arrays of real(8):
    ! f, f_eq, ux, uy, rho) before the device is touched.
    subroutine check_restart_file()
        logical :: there
        integer(8) :: fsize, need
        inquire(file='restart/restart.dat', exist=there, size=fsize)
        need = 4_8 + 8_8*21_8*int(nx, 8)*int(ny, 8)
        if (.not. there) then
            write(*, '(a)') "Error: irestart = 1 but restart/restart.dat does not exist"
            stop 1
        end if
        if (fsize >= 0 .and. fsize < need + 8_8) then
            write(*, '(a,i0,a,i0,a,i0,a,i0)') "Error: restart/restart.dat holds ", fsize, " bytes; the ", nx, &
                " x ", ny, " record needs ", need + 8_8
            stop 1
        end if
    end subroutine

    subroutine create_master_p3d_file()      ! 2d_mrlbm.f90:2334-2367
        integer(8) :: it8, iend
        real(8) :: phy_time
        character(len=100) :: fb
        iend = 10000_8*int(max(iplot, 1), 8)
        open(unit=10, file='output/master.p3d')
            write(10,*) "{"
            write(10,*)
            write(10,*) '"auto-detect-format"', ": true,"
            write(10,*)
            write(10,*) '"filenames"', ": ["
            write(10,*)
            do it8 = 0_8, iend, int(max(iplot, 1), 8)
                phy_time = it8*delt_phy
                write(fb, '(a4,I8,a4)') 'grid', 10000000 + it8, '.bin'
                write(10,*) '{ "time" :', phy_time, ',"xyz" : "grid.x", "function" :', ' "', trim(fb), '" },'
            end do
            write(10,*)
            write(10,*) "]"
            write(10,*) "}"
        close(10)
    end subroutine

    ! read_input_file (2d_mrlbm.f90:2222-2275).  The file is read into memory and a '/'
    ! appended, so a last group without its closing '/' (as shipped) is accepted.
    subroutine read_input_file()
        character(len=512) :: lines(2000)
        integer :: n, ios, u
        namelist/DomainSize/nx, ny, lattice_type
        namelist/Cylinder/ncy, L_front, L_back, L_top, L_bot
        namelist/Numbers/Re
        namelist/Parallel/num_threads
        namelist/Controls/uo, rho_infty, iplot, max_iter, isave, irestart, statsbegin, statsend, iplotstats
        namelist/LogicalControls/post_process, x_periodic, y_periodic, channel_with_cylinder, channel_with_square, &
            & incomp_bc, incomp_bulk, vel_interp, mom_interp, rotated_coordinate, strong_mass_bc, eq_mass_bc, weak_mass_bc
        namelist/Accelerator/precision, inlet_profile, strip_rows, kernel, record_surface_p, time_blocking, &
            & steps_per_pass, quad_rows

        open(newunit=u, file="input.dat", status="old", iostat=ios)
        if (ios /= 0) then
            write(*, '(a)') "Error while reading input.dat file..."
            stop 1
        end if
        write(*, *) 'Reading input.dat file ...'
        n = 0
        do
            read(u, '(A)', iostat=ios) lines(n + 1)
            if (ios /= 0) exit
            n = n + 1
            if (n == size(lines) - 1) exit
        end do
        close(u)
        lines(n + 1) = '/'
        read(lines(1:n + 1), nml=DomainSize, iostat=ios); if (ios /= 0) call input_error()
        read(lines(1:n + 1), nml=Cylinder, iostat=ios); if (ios /= 0) call input_error()
        read(lines(1:n + 1), nml=Numbers, iostat=ios); if (ios /= 0) call input_error()
        read(lines(1:n + 1), nml=Parallel, iostat=ios); if (ios /= 0) call input_error()
        read(lines(1:n + 1), nml=Controls, iostat=ios); if (ios /= 0) call input_error()
        read(lines(1:n + 1), nml=LogicalControls, iostat=ios); if (ios /= 0) call input_error()
        read(lines(1:n + 1), nml=Accelerator, iostat=ios)    ! optional group
        write(*, DomainSize)
        write(*, Cylinder)
        write(*, Numbers)
        write(*, Parallel)
        write(*, Controls)
        write(*, LogicalControls)
        write(*, '(a40, 5x, g15.5)') 'Reynolds number (Re) is ', Re
    end subroutine

    subroutine input_error()
        write(*, '(a)') "Error while reading input.dat file..."
        stop 1
    end subroutine

end program mrlbm_gpu
