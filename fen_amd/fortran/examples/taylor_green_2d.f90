program taylor_green_2d

    ! 2-D Taylor-Green vortex written against FEN's interface (grid%setup,
    ! init_solver, advance_solution, print_solver_status, scalar%write), the
    ! pattern of the reference's test/small_test/navier_stokes/taylor_green_vortex.
    ! Doubly periodic [0,2pi]^2, u = sin x cos y, v = -cos x sin y,
    ! p = -(cos 2x + cos 2y)/4, nu = 0.01, dt = 1e-3.
    ! Usage: taylor_green_2d [N] [nsteps]   -> u.raw v.raw p.raw, and
    ! save_fields' data/vx_<nsteps>.raw, vy_, p_ when a data/ folder exists.

    use IO_mod
    use precision_mod     , only : dp
    use global_mod        , only : pi
    use grid_mod
    use solver_mod        , only : init_solver, advance_solution, destroy_solver, print_solver_status, &
                                   save_fields
    use navier_stokes_mod , only : v, p, viscosity, dt_o

    implicit none

    integer          :: Nx, Ny, Nz, step, nsteps
    real(dp)         :: Lx, Ly, Lz, dt, time, origin(3)
    type(grid)       :: comp_grid
    type(bc_type)    :: bc(4)
    character(len=32):: arg
    logical          :: have_data

    Nx = 128
    nsteps = 1000
    if (command_argument_count() >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) Nx
    end if
    if (command_argument_count() >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) nsteps
    end if

    Lx = 2.0_dp*pi
    Ly = Lx
    Ny = Nx
    Nz = 1
    Lz = Lx*float(Nz)/float(Nx)
    origin = [0.0_dp, 0.0_dp, 0.0_dp]
    bc(1)%s = 'Periodic'
    bc(2)%s = 'Periodic'
    bc(3)%s = 'Periodic'
    bc(4)%s = 'Periodic'
    call comp_grid%setup(Nx, Ny, Nz, Lx, Ly, Lz, origin, 1, 1, bc)

    call init_solver(comp_grid)
    viscosity = 0.01_dp
    call init_fields

    dt = 1.0e-3_dp
    dt_o = dt
    time = 0.0_dp
    do step = 1, nsteps
        time = time + dt
        call advance_solution(comp_grid, step, dt)
        if (mod(step, 100) == 0 .or. step == nsteps) call print_solver_status(stdout, step, time, dt)
    end do

    call v%x%write('u.raw')
    call v%y%write('v.raw')
    call p%write('p.raw')
    inquire(file = 'data/.', exist = have_data)
    if (have_data) call save_fields(nsteps)

    call destroy_solver()
    call comp_grid%destroy

contains

    subroutine init_fields
        integer  :: i, j, k
        real(dp) :: x, y
        do k = comp_grid%lo(3), comp_grid%hi(3)
            do j = comp_grid%lo(2), comp_grid%hi(2)
                do i = comp_grid%lo(1), comp_grid%hi(1)
                    x = float(i)*comp_grid%delta
                    y = (float(j) - 0.5_dp)*comp_grid%delta
                    v%x%f(i,j,k) = sin(x)*cos(y)
                    x = (float(i) - 0.5_dp)*comp_grid%delta
                    y = float(j)*comp_grid%delta
                    v%y%f(i,j,k) = -cos(x)*sin(y)
                    x = (float(i) - 0.5_dp)*comp_grid%delta
                    y = (float(j) - 0.5_dp)*comp_grid%delta
                    p%f(i,j,k) = -0.25_dp*(cos(2.0_dp*x) + cos(2.0_dp*y))
                end do
            end do
        end do
        call p%update_ghost_nodes
        call v%update_ghost_nodes
    end subroutine init_fields

end program taylor_green_2d
