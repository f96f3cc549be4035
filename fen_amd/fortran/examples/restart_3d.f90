program restart_3d

    ! Restart determinism through FEN's interface, the pattern of the
    ! reference's test/small_test/io (test_NS.f90 + run.sh): case 1 runs 10
    ! steps and saves the state after steps 5 and 10; case 2 loads step 5,
    ! runs steps 6..10 and saves as step 20.  The two files must be identical.
    ! Initial fields are read from init_{u,v,w,p}.raw (16^3 interiors).
    ! Usage: restart_3d 1|2

    use precision_mod     , only : dp
    use global_mod        , only : pi
    use grid_mod
    use solver_mod
    use navier_stokes_mod , only : v, p, set_timestep

    implicit none

    integer          :: Nx, Ny, Nz, step, i
    real(dp)         :: Lx, dt, time
    type(grid)       :: comp_grid
    character(len=1) :: case

    Nx = 16
    Ny = 16
    Nz = 16
    Lx = 2.0_dp*pi
    call comp_grid%setup(Nx, Ny, Nz, Lx, Lx, Lx, [0.0_dp, 0.0_dp, 0.0_dp], 1, 1)
    call init_solver(comp_grid)
    step = 0
    time = 0.0_dp
    call v%x%read('init_u.raw')
    call v%y%read('init_v.raw')
    call v%z%read('init_w.raw')
    call p%read('init_p.raw')
    call p%update_ghost_nodes()
    call v%update_ghost_nodes()
    call set_timestep(comp_grid, dt, 2.0_dp)

    call get_command_argument(1, case)
    select case (case)
    case ('1')
        do i = 1, 10
            step = step + 1
            call advance_solution(comp_grid, step, dt)
            if (i == 5) call save_state(step)
        end do
        call save_state(step)
    case ('2')
        step = 5
        call load_state(step)
        do i = 6, 10
            step = step + 1
            time = time + dt
            call advance_solution(comp_grid, step, dt)
        end do
        call save_state(step*2)
    end select

    call destroy_solver()
    call comp_grid%destroy()

end program restart_3d
