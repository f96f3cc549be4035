module rt_setup
    ! The initial interface as a signed distance; a module function, so the
    ! procedure pointer needs no trampoline (the reference's program uses an internal
    ! function, which under flang links only with -Wl,-z,execstack).
    use precision_mod, only : dp
    use global_mod, only : pi
    implicit none
    real(dp) :: Lx = 1.0_dp, Ly = 4.0_dp
contains
    function wave(x, y) result(d)
        real(dp), intent(in) :: x, y
        real(dp) :: d
        d = y - 0.05_dp*cos(2.0_dp*pi*x/Lx) - Ly/2.0_dp
    end function wave
end module rt_setup

program rayleigh_taylor_2d

    ! Two-phase Rayleigh-Taylor instability written against FEN's MF interface
    ! (multiphase_mod's rho_0 ... mu_1, volume_of_fluid_mod's distance pointer,
    ! init_solver, set_timestep, advance_solution), the setup of the
    ! reference's test/small_test/multiphase/Rayleigh-Taylor: 128 x 512 on
    ! [0,1] x [0,4], periodic in x, walls in y, heavy fluid (1.225) above the
    ! light one (0.1694) across y = 2 + 0.05 cos(2 pi x), mu = 3.13e-3 in both,
    ! gravity -9.80665 in y, dt from set_timestep with U = 1.
    ! Usage: rayleigh_taylor_2d [nsteps] [Nx]   (default: until t = 1)
    !   -> u.raw v.raw p.raw vof.raw (interiors) and the solver status lines;
    !   data/vof_<n> at t = 0 and every 0.1 from t = 0.7 when data/ exists.

    use IO_mod
    use precision_mod       , only : dp
    use global_mod          , only : gravity
    use rt_setup            , only : Lx, Ly, wave
    use grid_mod
    use multiphase_mod      , only : rho_0, rho_1, mu_0, mu_1
    use volume_of_fluid_mod , only : distance, vof
    use solver_mod
    use navier_stokes_mod   , only : set_timestep, g, v, p

    implicit none

    integer           :: Nx, Ny, Nz, step, nsteps, ne
    real(dp)          :: Lz, time, dt, Tmax, time_event, origin(3)
    type(grid)        :: comp_grid
    type(bc_type)     :: bc(4)
    character(len=32) :: arg
    character(len=7)  :: ss
    logical           :: have_data

    nsteps = -1
    Nx = 128
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) nsteps
    end if
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) Nx
    end if
    Ny = 4*Nx
    Nz = 1
    Lx = 1.0_dp
    Ly = 4.0_dp
    Lz = Lx*float(Nz)/float(Nx)
    origin = [0.0_dp, 0.0_dp, 0.0_dp]
    bc(1)%s = 'Periodic'
    bc(2)%s = 'Periodic'
    bc(3)%s = 'Wall'
    bc(4)%s = 'Wall'
    call comp_grid%setup(Nx, Ny, Nz, Lx, Ly, Lz, origin, 1, 1, bc)

    rho_0 = 0.1694_dp
    rho_1 = 1.2250_dp
    mu_0 = 3.13d-3
    mu_1 = mu_0
    g(2) = -gravity
    distance => wave

    call init_solver(comp_grid)
    step = 0
    time = 0.0_dp
    call set_timestep(comp_grid, dt, 1.0_dp)
    Tmax = 1.0_dp

    inquire(file = 'data/.', exist = have_data)
    if (have_data) call vof%write('data/vof_0')
    time_event = 0.7_dp
    ne = 1

    do
        if (nsteps >= 0) then
            if (step >= nsteps) exit
        else
            if (time >= Tmax) exit
        end if
        step = step + 1
        time = time + dt
        call advance_solution(comp_grid, step, dt)
        call print_solver_status(stdout, step, time, dt)
        if (have_data .and. time >= time_event) then
            write(ss, '(I0.1)') ne
            call vof%write('data/vof_'//trim(ss))
            time_event = time_event + 0.1_dp
            ne = ne + 1
        end if
    end do

    call v%x%write('u.raw')
    call v%y%write('v.raw')
    call p%write('p.raw')
    call vof%write('vof.raw')

    call destroy_solver()
    call comp_grid%destroy()

end program rayleigh_taylor_2d
