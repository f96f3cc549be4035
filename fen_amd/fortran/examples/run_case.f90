program run_case_file

    ! Runs the case a case file describes (format in include/fen.h): set-up,
    ! time loop with the status line, save_state / save_fields at the listed
    ! steps, restart from a saved state.  The same source builds for 2-D,
    ! 3-D and the two-phase build (-DMF), as FEN's own programs do.
    ! Usage: run_case <case.json>

    use precision_mod, only : dp
    use grid_mod     , only : grid
    use solver_mod   , only : destroy_solver
    use case_mod     , only : fen_case, read_case, setup_case, run_case

    implicit none

    type(fen_case)      :: cs
    type(grid)          :: comp_grid
    real(dp)            :: dt
    integer             :: step0
    character(len=1024) :: path

    if (command_argument_count() /= 1) then
        write(*, '(A)') 'usage: run_case <case.json>'
        error stop 2
    end if
    call get_command_argument(1, path)
    call read_case(trim(path), cs)
    call setup_case(cs, comp_grid, dt, step0)
    call run_case(cs, comp_grid, dt, step0)
    call destroy_solver()
    call comp_grid%destroy()

end program run_case_file
