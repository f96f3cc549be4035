!> Fortran interfaces to libfen's C-ABI (include/fen.h).  Every function
!> returns an integer status, FEN_OK (0) on success.
module fen_capi

    use iso_c_binding, only : c_ptr, c_int, c_double, c_size_t, c_char, c_null_ptr

    implicit none

    integer(c_int), parameter :: FEN_OK = 0
    integer(c_int), parameter :: FEN_P = 0, FEN_PHI = 1, FEN_U = 2, FEN_V = 3, FEN_W = 4
    integer(c_int), parameter :: FEN_DVO_X = 5, FEN_DVO_Y = 6, FEN_DVO_Z = 7
    integer(c_int), parameter :: FEN_S_X = 8, FEN_S_Y = 9, FEN_S_Z = 10
    integer(c_int), parameter :: FEN_VOF = 11, FEN_RHO = 12, FEN_MU = 13
    integer(c_int), parameter :: FEN_BC_PERIODIC = 0, FEN_BC_WALL = 1, FEN_BC_INFLOW = 2, FEN_BC_OUTFLOW = 3
    integer(c_int), parameter :: FEN_FP64 = 0, FEN_FP32_REFINE = 1
    integer, parameter :: FEN_CASE_MAXLIST = 32, FEN_CASE_PATHLEN = 256

    !> A case file's contents (struct fen_case of include/fen.h, member for member).
    type, bind(C) :: fen_case
        integer(c_int) :: dim, n(3), bc(6), constant_cfl, precision, restart, steps, status_every
        integer(c_int) :: n_save_state, save_state(FEN_CASE_MAXLIST)
        integer(c_int) :: n_save_fields, save_fields(FEN_CASE_MAXLIST)
        integer(c_int) :: write_final, multiphase
        real(c_double) :: origin(3), L(3)
        real(c_double) :: density, viscosity, g(3), cfl
        real(c_double) :: dt, U
        real(c_double) :: rho0, rho1, mu0, mu1, sigma
        real(c_double) :: iface(3)
        character(kind=c_char) :: name(64)
        character(kind=c_char) :: data_dir(FEN_CASE_PATHLEN)
        character(kind=c_char) :: restart_dir(FEN_CASE_PATHLEN)
        character(kind=c_char) :: init(FEN_CASE_PATHLEN, 5)   ! u, v, w, p, vof
    end type fen_case

    interface

        integer(c_int) function fen_create(ctx, dim, nx, ny, nz, lx, ly, lz, x0, bc, precision) &
                bind(C, name = 'fen_create')
            import :: c_ptr, c_int, c_double
            type(c_ptr), intent(out) :: ctx
            integer(c_int), value :: dim, nx, ny, nz, precision
            real(c_double), value :: lx, ly, lz
            real(c_double), intent(in) :: x0(3)
            integer(c_int), intent(in) :: bc(6)
        end function

        integer(c_int) function fen_destroy(ctx) bind(C, name = 'fen_destroy')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function

        type(c_ptr) function fen_last_error(ctx) bind(C, name = 'fen_last_error')
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function fen_set_params(ctx, density, viscosity, g, cfl, constant_cfl) &
                bind(C, name = 'fen_set_params')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: density, viscosity, cfl
            real(c_double), intent(in) :: g(3)
            integer(c_int), value :: constant_cfl
        end function

        integer(c_int) function fen_set_timestep(ctx, U, dt) bind(C, name = 'fen_set_timestep')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: U
            real(c_double), intent(out) :: dt
        end function

        integer(c_int) function fen_set_dt_o(ctx, dt_o) bind(C, name = 'fen_set_dt_o')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt_o
        end function

        integer(c_int) function fen_get_scalars(ctx, maxdiv, maxcfl, dt_o, dt_visc, dt_conv) &
                bind(C, name = 'fen_get_scalars')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            ! an absent argument is passed as NULL (not read back)
            real(c_double), intent(out), optional :: maxdiv, maxcfl, dt_o, dt_visc, dt_conv
        end function

        integer(c_int) function fen_upload(ctx, field, host, count) bind(C, name = 'fen_upload')
            import :: c_ptr, c_int, c_double, c_size_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: field
            real(c_double), intent(in) :: host(*)
            integer(c_size_t), value :: count
        end function

        integer(c_int) function fen_download(ctx, field, host, count) bind(C, name = 'fen_download')
            import :: c_ptr, c_int, c_double, c_size_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: field
            real(c_double), intent(out) :: host(*)
            integer(c_size_t), value :: count
        end function

        integer(c_int) function fen_set_bc_plane(ctx, field, face, plane, count) bind(C, name = 'fen_set_bc_plane')
            import :: c_ptr, c_int, c_double, c_size_t
            type(c_ptr), value :: ctx
            integer(c_int), value :: field, face
            real(c_double), intent(in) :: plane(*)
            integer(c_size_t), value :: count
        end function

        integer(c_int) function fen_set_bc_type(ctx, field, face, type) bind(C, name = 'fen_set_bc_type')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: field, face, type
        end function

        integer(c_int) function fen_update_ghost_nodes(ctx, field) bind(C, name = 'fen_update_ghost_nodes')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: field
        end function

        integer(c_int) function fen_ns_step(ctx, step, dt, maxdiv, maxcfl) bind(C, name = 'fen_ns_step')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: step
            real(c_double), intent(inout) :: dt
            ! absent: the step is queued without waiting for its checks
            real(c_double), intent(out), optional :: maxdiv, maxcfl
        end function

        integer(c_int) function fen_ns_run(ctx, step0, nsteps, dt) bind(C, name = 'fen_ns_run')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: step0, nsteps
            real(c_double), intent(inout) :: dt
        end function

        integer(c_int) function fen_poisson_solve(ctx, f, gl) bind(C, name = 'fen_poisson_solve')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(inout) :: f(*)
            integer(c_int), value :: gl
        end function

        integer(c_int) function fen_set_multiphase(ctx, rho0, rho1, mu0, mu1, sigma, beta) &
                bind(C, name = 'fen_set_multiphase')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: rho0, rho1, mu0, mu1, sigma, beta
        end function

        integer(c_int) function fen_update_material_properties(ctx) bind(C, name = 'fen_update_material_properties')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function fen_advect_vof(ctx, dt) bind(C, name = 'fen_advect_vof')
            import :: c_ptr, c_int, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function

        integer(c_int) function fen_set_vof_x_first(ctx, x_first) bind(C, name = 'fen_set_vof_x_first')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), value :: x_first
        end function

        integer(c_int) function fen_get_vof_x_first(ctx, x_first) bind(C, name = 'fen_get_vof_x_first')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
            integer(c_int), intent(out) :: x_first
        end function

        integer(c_int) function fen_synchronize(ctx) bind(C, name = 'fen_synchronize')
            import :: c_ptr, c_int
            type(c_ptr), value :: ctx
        end function

        integer(c_int) function fen_save_state(ctx, path) bind(C, name = 'fen_save_state')
            import :: c_ptr, c_int, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: path(*)
        end function

        integer(c_int) function fen_load_state(ctx, path) bind(C, name = 'fen_load_state')
            import :: c_ptr, c_int, c_char
            type(c_ptr), value :: ctx
            character(kind=c_char), intent(in) :: path(*)
        end function

        integer(c_int) function fen_case_read(path, out, err, errlen) bind(C, name = 'fen_case_read')
            import :: c_int, c_char, c_size_t, fen_case
            character(kind=c_char), intent(in) :: path(*)
            type(fen_case), intent(out) :: out
            character(kind=c_char), intent(out) :: err(*)
            integer(c_size_t), value :: errlen
        end function

    end interface

contains

    !> Aborts with the library's message when a call failed (the reference
    !> prints to stderr; an unusable device state cannot continue).
    subroutine fen_check(ctx, rc, where)
        use iso_c_binding, only : c_f_pointer, c_associated
        use, intrinsic :: iso_fortran_env, only : error_unit
        type(c_ptr), intent(in) :: ctx
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        type(c_ptr) :: p
        integer :: n
        if (rc == FEN_OK) return
        p = fen_last_error(ctx)
        write(error_unit, '(A,A,I0)') where, ' failed with status ', rc
        if (c_associated(p)) then
            call c_f_pointer(p, msg, [512])
            n = 0
            do while (n < 512)
                if (msg(n + 1) == char(0)) exit
                n = n + 1
            end do
            if (n > 0) write(error_unit, '(512A)') msg(1:n)
        end if
        error stop 1
    end subroutine fen_check

end module fen_capi
