! fen_modules.F90 -- FEN's Fortran interface for the fractional-step path,
! implemented over libfen (HIP, MI355X).
!
! The modules keep the reference's names and signatures so that FEN user
! programs compile against them: precision_mod, global_mod, IO_mod, grid_mod,
! scalar_mod, vector_mod, halo_mod, fields_mod, poisson_mod,
! navier_stokes_mod, solver_mod, and with -DMF (two-phase build, 2-D only,
! as FEN's volume of fluid) volume_of_fluid_mod and multiphase_mod.  Build
! with -DDIM=2 or -DDIM=3 as FEN is.
!
! Residency: the navier_stokes_mod fields live in HBM during the run.  Their
! host arrays (p%f, v%x%f, ...) are mirrors: everything the host holds is
! uploaded by the first navier_stokes_solver call after the fields were
! allocated, loaded or marked with fen_host_modified(); after a step the
! host mirrors are refreshed lazily, by the type-bound routines that read
! them (write, max_value, integral, update_ghost_nodes), by save_state, or by
! an explicit call fen_sync_host().  User code that reads or edits the
! arrays directly between steps calls fen_sync_host() before and
! fen_host_modified() after.  One process drives one GPU: MPI is not used,
! prow and pcol are accepted and ignored.

!===================================================================================================
module precision_mod
    ! Kinds (reference src/precision.f90).
    implicit none
    public
    integer, parameter :: sp = selected_real_kind(6, 37)
    integer, parameter :: dp = selected_real_kind(15, 307)
end module precision_mod

!===================================================================================================
module global_mod
    ! Constants (reference src/global.f90).
    use precision_mod, only : dp
    implicit none
#if DIM==3
    integer, parameter :: Ndim = 3
#else
    integer, parameter :: Ndim = 2
#endif
    real(dp), parameter :: pi = acos(-1.0_dp)
    real(dp), parameter :: small = 1.0e-14_dp
    real(dp), parameter :: gravity = 9.80665_dp
    integer :: myrank = 0
    integer :: ierror = 0
end module global_mod

!===================================================================================================
module IO_mod
    ! Error messages go to stderr and the run continues (reference src/IO.f90:12-30).
    use, intrinsic :: iso_fortran_env, only : stdout => output_unit, stderr => error_unit
    implicit none
contains
    subroutine print_error_message(message)
        character(len=*), intent(in) :: message
        write(stderr, '(A)') message
    end subroutine print_error_message
end module IO_mod

!===================================================================================================
module grid_mod
    ! Uniform Cartesian grid (reference src/grid.f90:22-265).
    use precision_mod, only : dp
    use global_mod, only : Ndim
    implicit none
    private
    public :: bc_type, grid

    type bc_type
        character(:), allocatable :: s
    end type bc_type

    type grid
        integer :: Nx, Ny, Nz
        integer, dimension(3) :: lo, hi, lo_y, hi_y, lo_z, hi_z
        real(dp) :: Lx, Ly, Lz
        real(dp) :: origin(3)
        real(dp) :: delta
        real(dp), dimension(:), allocatable :: x, y, z
        type(bc_type) :: boundary_conditions(Ndim*2)
        logical :: periodic_bc(3)
        integer :: rank = 0, nranks = 1, prow = 1, pcol = 1
        character(len=99) :: name = 'grid'
    contains
        procedure, pass(self) :: setup
        procedure, pass(self) :: print_json
        procedure, pass(self) :: destroy
    end type grid

contains

    subroutine setup(self, Nx, Ny, Nz, Lx, Ly, Lz, x0, prow, pcol, bc)
        use IO_mod, only : print_error_message
        class(grid), intent(inout) :: self
        integer, intent(in) :: Nx, Ny, Nz, prow, pcol
        real(dp), intent(in) :: Lx, Ly, Lz, x0(3)
        type(bc_type), intent(in), optional :: bc(:)
        integer :: i, f

        do f = 1, Ndim*2
            self%boundary_conditions(f)%s = 'Periodic'
        end do
        if (present(bc)) self%boundary_conditions = bc
        self%periodic_bc = [.false., .false., .true.]
        self%periodic_bc(1) = self%boundary_conditions(1)%s == 'Periodic' .and. &
                              self%boundary_conditions(2)%s == 'Periodic'
        self%periodic_bc(2) = self%boundary_conditions(3)%s == 'Periodic' .and. &
                              self%boundary_conditions(4)%s == 'Periodic'
#if DIM==3
        self%periodic_bc(3) = self%boundary_conditions(5)%s == 'Periodic' .and. &
                              self%boundary_conditions(6)%s == 'Periodic'
#endif
        self%Nx = Nx
        self%Ny = Ny
        self%Nz = Nz
        self%Lx = Lx
        self%Ly = Ly
        self%Lz = Lz
        self%origin = x0
        self%delta = Lx/float(Nx)
        if (Lx/float(Nx) .ne. Ly/float(Ny)) &
            call print_error_message('The grid spacing must be equal in all directions')
        allocate(self%x(0:Nx+1), self%y(0:Ny+1), self%z(0:Nz+1))
        do i = 0, Nx + 1
            self%x(i) = x0(1) + (i - 0.5_dp)*self%delta
        end do
        do i = 0, Ny + 1
            self%y(i) = x0(2) + (i - 0.5_dp)*self%delta
        end do
        do i = 0, Nz + 1
            self%z(i) = x0(3) + (i - 0.5_dp)*self%delta
        end do
        ! one process owns the whole grid
        self%lo = [1, 1, 1]
        self%hi = [Nx, Ny, Nz]
        self%lo_y = self%lo
        self%hi_y = self%hi
        self%lo_z = self%lo
        self%hi_z = self%hi
        self%rank = 0
        self%nranks = 1
        self%prow = 1
        self%pcol = 1
        if (prow*pcol < 0) call print_error_message('invalid process grid')
        call self%print_json
    end subroutine setup

    !> grid.json, the layout post-processing scripts read (reference src/grid.f90:233-265).
    subroutine print_json(self)
        class(grid), intent(in) :: self
        integer :: out_id
        open(newunit = out_id, file = trim(self%name)//'.json')
        write(out_id, '(A1)') '{'
        write(out_id, '(4x,A9)') '"Grid": {'
        write(out_id, '(8x,A6,1x,I7,A1)') '"Nx": ', self%Nx, ','
        write(out_id, '(8x,A6,1x,I7,A1)') '"Ny": ', self%Ny, ','
        write(out_id, '(8x,A6,1x,I7,A1)') '"Nz": ', self%Nz, ','
        write(out_id, '(8x,A11,1x,E16.8,A1,E16.8,A1,E16.8,A2)') '"origin": [', &
            self%origin(1), ',', self%origin(2), ',', self%origin(3), '],'
        write(out_id, '(8x,A6,1x,E16.8,A1)') '"Lx": ', self%Lx, ','
        write(out_id, '(8x,A6,1x,E16.8,A1)') '"Ly": ', self%Ly, ','
        write(out_id, '(8x,A6,1x,E16.8)') '"Lz": ', self%Lz
        write(out_id, '(4x,A3)') '}'
        write(out_id, '(A1)') '}'
        close(out_id)
    end subroutine print_json

    subroutine destroy(self)
        class(grid), intent(inout) :: self
        integer :: f
        if (allocated(self%x)) deallocate(self%x, self%y, self%z)
        do f = 1, Ndim*2
            if (allocated(self%boundary_conditions(f)%s)) deallocate(self%boundary_conditions(f)%s)
        end do
    end subroutine destroy

end module grid_mod

!===================================================================================================
module scalar_mod
    ! Scalar field with ghost nodes and boundary conditions (reference src/scalar.f90).
    use precision_mod, only : dp
    use grid_mod, only : grid
    implicit none
    private
    public :: scalar, field_bc, scalar_sync_hook

    ! bc types: -1 halo, 0 periodic, 1 Dirichlet, 2 Neumann
    type field_bc
        integer :: type_left = 0, type_right = 0, type_bottom = 0, type_top = 0
        integer :: type_front = 0, type_back = 0
        real(dp), allocatable :: left(:,:), right(:,:), bottom(:,:), top(:,:)
        real(dp), allocatable :: front(:,:), back(:,:)
    end type field_bc

    type scalar
        real(dp), allocatable :: f(:,:,:)
        integer :: gl = 0
        type(grid), pointer :: G => Null()
        type(field_bc) :: bc
        character(1) :: c = 'c'
        character(len=99) :: name = 'unset'
        integer :: fid = -1   ! device field id when this is a navier_stokes_mod field
    contains
        procedure, pass(self) :: allocate
        procedure, pass(self) :: setToValue
        procedure, pass(self) :: update_ghost_nodes
        procedure, pass(self) :: max_value
        procedure, pass(self) :: integral
        procedure, pass(self) :: read
        procedure, pass(self) :: write
        procedure, pass(self) :: destroy
    end type scalar

    abstract interface
        subroutine sync_i()
        end subroutine sync_i
    end interface

    !> Refreshes the host mirrors of device-resident fields (set by navier_stokes_mod).
    procedure(sync_i), pointer :: scalar_sync_hook => Null()

contains

    subroutine sync_if_resident(self)
        class(scalar), intent(in) :: self
        if (self%fid >= 0 .and. associated(scalar_sync_hook)) call scalar_sync_hook()
    end subroutine sync_if_resident

    subroutine allocate(self, G, l)
        class(scalar), intent(inout) :: self
        type(grid), intent(in), target :: G
        integer, intent(in), optional :: l
        self%G => G
        if (present(l)) self%gl = l
        allocate(self%f(G%lo(1)-self%gl:G%hi(1)+self%gl, G%lo(2)-self%gl:G%hi(2)+self%gl, &
                        G%lo(3)-self%gl:G%hi(3)+self%gl))
        self%f = 0.0_dp
        if (self%gl > 0) then
            allocate(self%bc%left(G%lo(2)-self%gl:G%hi(2)+self%gl, G%lo(3)-self%gl:G%hi(3)+self%gl))
            allocate(self%bc%right(G%lo(2)-self%gl:G%hi(2)+self%gl, G%lo(3)-self%gl:G%hi(3)+self%gl))
            allocate(self%bc%bottom(G%lo(1)-self%gl:G%hi(1)+self%gl, G%lo(3)-self%gl:G%hi(3)+self%gl))
            allocate(self%bc%top(G%lo(1)-self%gl:G%hi(1)+self%gl, G%lo(3)-self%gl:G%hi(3)+self%gl))
            allocate(self%bc%front(G%lo(1)-self%gl:G%hi(1)+self%gl, G%lo(2)-self%gl:G%hi(2)+self%gl))
            allocate(self%bc%back(G%lo(1)-self%gl:G%hi(1)+self%gl, G%lo(2)-self%gl:G%hi(2)+self%gl))
            self%bc%left = 0.0_dp
            self%bc%right = 0.0_dp
            self%bc%bottom = 0.0_dp
            self%bc%top = 0.0_dp
            self%bc%front = 0.0_dp
            self%bc%back = 0.0_dp
        end if
    end subroutine allocate

    subroutine setToValue(self, val)
        class(scalar), intent(inout) :: self
        real(dp), intent(in) :: val
        self%f = val
    end subroutine setToValue

    !> Ghost fill on the host array, in the reference's order left, right,
    !> bottom, top, front, back over full planes (src/scalar.f90:254-388).
    !> The device keeps the same rules for the resident fields.
    subroutine update_ghost_nodes(self)
        use IO_mod, only : print_error_message
        class(scalar), intent(inout) :: self
        integer :: lo(3), hi(3)
        logical :: tang
        call sync_if_resident(self)
        if (self%gl < 1) then
            call print_error_message('Cannot update ghost nodes on scalar '//trim(self%name))
            return
        end if
        lo = self%G%lo
        hi = self%G%hi
        ! x faces
        tang = self%c /= 'x'
        select case (self%bc%type_left)
        case (0); self%f(lo(1)-1,:,:) = self%f(hi(1),:,:)
        case (1)
            if (tang) then
                self%f(lo(1)-1,:,:) = 2.0_dp*self%bc%left(:,:) - self%f(lo(1),:,:)
            else
                self%f(lo(1)-1,:,:) = self%bc%left
            end if
        case (2); self%f(lo(1)-1,:,:) = self%f(lo(1),:,:)
        end select
        select case (self%bc%type_right)
        case (0); self%f(hi(1)+1,:,:) = self%f(lo(1),:,:)
        case (1)
            if (tang) then
                self%f(hi(1)+1,:,:) = 2.0_dp*self%bc%right(:,:) - self%f(hi(1),:,:)
            else
                self%f(hi(1),:,:) = self%bc%right
                self%f(hi(1)+1,:,:) = self%bc%right
            end if
        case (2); self%f(hi(1)+1,:,:) = self%f(hi(1),:,:)
        end select
        ! y faces
        tang = self%c /= 'y'
        select case (self%bc%type_bottom)
        case (0); self%f(:,lo(2)-1,:) = self%f(:,hi(2),:)
        case (1)
            if (tang) then
                self%f(:,lo(2)-1,:) = 2.0_dp*self%bc%bottom(:,:) - self%f(:,lo(2),:)
            else
                self%f(:,lo(2)-1,:) = self%bc%bottom
            end if
        case (2); self%f(:,lo(2)-1,:) = self%f(:,lo(2),:)
        end select
        select case (self%bc%type_top)
        case (0); self%f(:,hi(2)+1,:) = self%f(:,lo(2),:)
        case (1)
            if (tang) then
                self%f(:,hi(2)+1,:) = 2.0_dp*self%bc%top(:,:) - self%f(:,hi(2),:)
            else
                self%f(:,hi(2),:) = self%bc%top
                self%f(:,hi(2)+1,:) = self%bc%top
            end if
        case (2); self%f(:,hi(2)+1,:) = self%f(:,hi(2),:)
        end select
#if DIM==3
        ! z faces
        tang = self%c /= 'z'
        select case (self%bc%type_front)
        case (0); self%f(:,:,lo(3)-1) = self%f(:,:,hi(3))
        case (1)
            if (tang) then
                self%f(:,:,lo(3)-1) = 2.0_dp*self%bc%front(:,:) - self%f(:,:,lo(3))
            else
                self%f(:,:,lo(3)-1) = self%bc%front
            end if
        case (2); self%f(:,:,lo(3)-1) = self%f(:,:,lo(3))
        end select
        select case (self%bc%type_back)
        case (0); self%f(:,:,hi(3)+1) = self%f(:,:,lo(3))
        case (1)
            if (tang) then
                self%f(:,:,hi(3)+1) = 2.0_dp*self%bc%back(:,:) - self%f(:,:,hi(3))
            else
                self%f(:,:,hi(3)) = self%bc%back
                self%f(:,:,hi(3)+1) = self%bc%back
            end if
        case (2); self%f(:,:,hi(3)+1) = self%f(:,:,hi(3))
        end select
#endif
    end subroutine update_ghost_nodes

    !> Signed maximum over the interior (reference src/scalar.f90:179-197).
    real(dp) function max_value(self)
        class(scalar), intent(in) :: self
        call sync_if_resident(self)
        max_value = maxval(self%f(self%G%lo(1):self%G%hi(1), self%G%lo(2):self%G%hi(2), &
                                  self%G%lo(3):self%G%hi(3)))
    end function max_value

    real(dp) function integral(self)
        class(scalar), intent(in) :: self
        call sync_if_resident(self)
        integral = sum(self%f(self%G%lo(1):self%G%hi(1), self%G%lo(2):self%G%hi(2), &
                              self%G%lo(3):self%G%hi(3)))
        integral = integral*self%G%delta**3
    end function integral

    !> Raw float64 interior in Fortran order (the reference's one-rank file layout).
    subroutine read(self, filename)
        class(scalar), intent(inout) :: self
        character(len=*), intent(in) :: filename
        integer :: u
        open(newunit = u, file = filename, access = 'stream', form = 'unformatted', status = 'old')
        read(u) self%f(self%G%lo(1):self%G%hi(1), self%G%lo(2):self%G%hi(2), self%G%lo(3):self%G%hi(3))
        close(u)
    end subroutine read

    subroutine write(self, filename)
        class(scalar), intent(in) :: self
        character(len=*), intent(in) :: filename
        integer :: u
        call sync_if_resident(self)
        open(newunit = u, file = filename, access = 'stream', form = 'unformatted', status = 'replace')
        write(u) self%f(self%G%lo(1):self%G%hi(1), self%G%lo(2):self%G%hi(2), self%G%lo(3):self%G%hi(3))
        close(u)
    end subroutine write

    subroutine destroy(self)
        class(scalar), intent(inout) :: self
        if (allocated(self%f)) deallocate(self%f)
        if (allocated(self%bc%left)) deallocate(self%bc%left, self%bc%right, self%bc%bottom, self%bc%top, &
                                                self%bc%front, self%bc%back)
        self%G => Null()
        self%fid = -1
    end subroutine destroy

end module scalar_mod

!===================================================================================================
module vector_mod
    ! Staggered vector: components on x, y and z faces (reference src/vector.f90).
    use precision_mod, only : dp
    use grid_mod, only : grid
    use scalar_mod, only : scalar
    implicit none
    private
    public :: vector

    type vector
        type(grid), pointer :: G => Null()
        type(scalar) :: x
        type(scalar) :: y
        type(scalar) :: z
        character(len=99) :: name = 'unset'
    contains
        procedure, pass(self) :: allocate
        procedure, pass(self) :: setToValue
        procedure, pass(self) :: update_ghost_nodes
        procedure, pass(self) :: destroy
    end type vector

contains

    subroutine allocate(self, G, l)
        class(vector), intent(inout) :: self
        type(grid), intent(in), target :: G
        integer, intent(in), optional :: l
        if (present(l)) then
            self%x%gl = l
            self%y%gl = l
            self%z%gl = l
        end if
        self%G => G
        call self%x%allocate(G, self%x%gl)
        call self%y%allocate(G, self%y%gl)
#if DIM==3
        call self%z%allocate(G, self%z%gl)
#endif
        self%x%c = 'x'
        self%y%c = 'y'
        self%z%c = 'z'
    end subroutine allocate

    subroutine setToValue(self, val)
        class(vector), intent(inout) :: self
        real(dp), intent(in) :: val
        self%x%f = val
        self%y%f = val
#if DIM==3
        self%z%f = val
#endif
    end subroutine setToValue

    subroutine update_ghost_nodes(self)
        class(vector), intent(inout) :: self
        call self%x%update_ghost_nodes()
        call self%y%update_ghost_nodes()
#if DIM==3
        call self%z%update_ghost_nodes()
#endif
    end subroutine update_ghost_nodes

    subroutine destroy(self)
        class(vector), intent(inout) :: self
        call self%x%destroy()
        call self%y%destroy()
#if DIM==3
        call self%z%destroy()
#endif
        self%G => Null()
    end subroutine destroy

end module vector_mod

!===================================================================================================
module halo_mod
    ! update_halos(f, G, l) (reference src/halo.f90:12-43).  One process owns
    ! the whole grid, so the halo exchange is the periodic wrap in y and z of
    ! the rows i in [lo, hi]; the ghost fill then overwrites every halo plane.
    implicit none
    private
    public :: update_halos
contains
    subroutine update_halos(f, G, l)
        use precision_mod, only : dp
        use grid_mod, only : grid
        integer, intent(in) :: l
        type(grid), intent(in) :: G
        real(dp), intent(inout) :: f(G%lo(1)-l:G%hi(1)+l, G%lo(2)-l:G%hi(2)+l, G%lo(3)-l:G%hi(3)+l)
        integer :: lo(3), hi(3), s
        lo = G%lo
        hi = G%hi
        do s = 1, l
            f(lo(1):hi(1), lo(2)-s, lo(3):hi(3)) = f(lo(1):hi(1), hi(2)-s+1, lo(3):hi(3))
            f(lo(1):hi(1), hi(2)+s, lo(3):hi(3)) = f(lo(1):hi(1), lo(2)+s-1, lo(3):hi(3))
        end do
#if DIM==3
        do s = 1, l
            f(lo(1):hi(1), :, lo(3)-s) = f(lo(1):hi(1), :, hi(3)-s+1)
            f(lo(1):hi(1), :, hi(3)+s) = f(lo(1):hi(1), :, lo(3)+s-1)
        end do
#endif
    end subroutine update_halos
end module halo_mod

!===================================================================================================
module fields_mod
    ! Host operators used by user programs (reference src/fields.f90:31-64,120-153).
    use precision_mod, only : dp
    use scalar_mod
    use vector_mod
    implicit none
    interface gradient
        module procedure gradient_of_scalar
    end interface gradient
    interface divergence
        module procedure divergence_of_vector
    end interface divergence
contains
    subroutine gradient_of_scalar(s, grad_s)
        type(scalar), intent(in) :: s
        type(vector), intent(inout) :: grad_s
        integer :: i, j, k
        real(dp) :: idelta
        idelta = 1.0_dp/s%G%delta
        do k = s%G%lo(3), s%G%hi(3)
            do j = s%G%lo(2), s%G%hi(2)
                do i = s%G%lo(1), s%G%hi(1)
                    grad_s%x%f(i,j,k) = (s%f(i+1,j,k) - s%f(i,j,k))*idelta
                    grad_s%y%f(i,j,k) = (s%f(i,j+1,k) - s%f(i,j,k))*idelta
#if DIM==3
                    grad_s%z%f(i,j,k) = (s%f(i,j,k+1) - s%f(i,j,k))*idelta
#endif
                end do
            end do
        end do
    end subroutine gradient_of_scalar

    subroutine divergence_of_vector(v, div_v)
        type(vector), intent(in) :: v
        type(scalar), intent(inout) :: div_v
        integer :: i, j, k
        real(dp) :: idelta
        idelta = 1.0_dp/v%G%delta
        do k = v%G%lo(3), v%G%hi(3)
            do j = v%G%lo(2), v%G%hi(2)
                do i = v%G%lo(1), v%G%hi(1)
                    div_v%f(i,j,k) = (v%x%f(i,j,k) - v%x%f(i-1,j,k))*idelta + &
                                     (v%y%f(i,j,k) - v%y%f(i,j-1,k))*idelta
#if DIM==3
                    div_v%f(i,j,k) = div_v%f(i,j,k) + (v%z%f(i,j,k) - v%z%f(i,j,k-1))*idelta
#endif
                end do
            end do
        end do
    end subroutine divergence_of_vector

    !> Face value averaged to cell centres, 0.5*(f(i) + f(i-1)) along `face`
    !> (reference src/fields.f90:210-250).
    subroutine face_to_center(sf, sc, face)
        type(scalar), intent(in) :: sf
        type(scalar), intent(inout) :: sc
        character(len=1), intent(in) :: face
        integer :: i, j, k, di, dj, dk
        di = 0; dj = 0; dk = 0
        select case(face)
            case('x'); di = 1
            case('y'); dj = 1
            case('z'); dk = 1
        end select
        do k = sc%G%lo(3), sc%G%hi(3)
            do j = sc%G%lo(2), sc%G%hi(2)
                do i = sc%G%lo(1), sc%G%hi(1)
                    sc%f(i,j,k) = 0.5_dp*(sf%f(i,j,k) + sf%f(i-di,j-dj,k-dk))
                end do
            end do
        end do
    end subroutine face_to_center
end module fields_mod

!===================================================================================================
module fen_bc_codes
    ! Physical boundary strings -> C-ABI codes.
    use fen_capi
    use grid_mod, only : grid
    use global_mod, only : Ndim
    implicit none
contains
    subroutine physical_bc_codes(G, codes)
        type(grid), intent(in) :: G
        integer(c_int), intent(out) :: codes(6)
        integer :: f
        codes = FEN_BC_PERIODIC
        do f = 1, Ndim*2
            select case (G%boundary_conditions(f)%s)
            case ('Periodic'); codes(f) = FEN_BC_PERIODIC
            case ('Inflow');   codes(f) = FEN_BC_INFLOW
            case ('Outflow');  codes(f) = FEN_BC_OUTFLOW
            case default;      codes(f) = FEN_BC_WALL
            end select
        end do
    end subroutine physical_bc_codes
end module fen_bc_codes

!===================================================================================================
module poisson_mod
    ! Fast direct Poisson solver (reference src/poisson.f90:21-113,1456-1481).
    ! solve_poisson solves lap(phi) = f in place on f's interior; the solve
    ! runs on the GPU.
    use precision_mod, only : dp
    use scalar_mod
    use fen_capi
    use iso_c_binding, only : c_ptr, c_null_ptr, c_associated, c_int, c_double
    implicit none
    private
    public :: init_Poisson_Solver, solve_Poisson, destroy_Poisson_solver, poisson_shared_ctx

    interface
        subroutine poisson_Solver(f)
            use scalar_mod
            type(scalar), intent(inout) :: f
        end subroutine poisson_Solver
    end interface

    procedure(poisson_Solver), pointer :: solve_poisson => Null()

    !> Context of navier_stokes_mod, shared by its pressure solve when set.
    type(c_ptr) :: poisson_shared_ctx = c_null_ptr
    type(c_ptr) :: ctx = c_null_ptr
    logical :: owns_ctx = .false.

contains

    subroutine init_poisson_solver(phi)
        use global_mod, only : Ndim
        use fen_bc_codes, only : physical_bc_codes
        type(scalar), intent(in) :: phi
        integer(c_int) :: codes(6), rc
        if (c_associated(poisson_shared_ctx)) then
            ctx = poisson_shared_ctx
            owns_ctx = .false.
        else
            call physical_bc_codes(phi%G, codes)
            rc = fen_create(ctx, Ndim, phi%G%Nx, phi%G%Ny, phi%G%Nz, phi%G%Lx, phi%G%Ly, phi%G%Lz, &
                            phi%G%origin, codes, FEN_FP64)
            if (rc /= FEN_OK) then
                print *, 'Unable to find the proped poisson solver with the selected boundary conditions. Aborting ...'
                stop
            end if
            owns_ctx = .true.
        end if
        solve_poisson => poisson_solver_gpu
    end subroutine init_poisson_solver

    subroutine poisson_solver_gpu(f)
        type(scalar), intent(inout) :: f
        call fen_check(ctx, fen_poisson_solve(ctx, f%f, int(f%gl, c_int)), 'solve_poisson')
    end subroutine poisson_solver_gpu

    subroutine destroy_Poisson_solver(phi)
        type(scalar), intent(in) :: phi
        integer(c_int) :: rc
        if (owns_ctx .and. c_associated(ctx)) rc = fen_destroy(ctx)
        ctx = c_null_ptr
        owns_ctx = .false.
        solve_poisson => Null()
    end subroutine destroy_Poisson_solver

end module poisson_mod

!===================================================================================================
module navier_stokes_mod
    ! Fractional-step solver (reference src/navier_stokes.f90).  Module state and
    ! entry points keep the reference's names; the step runs on the GPU.
    use precision_mod, only : dp
    use grid_mod, only : grid
    use scalar_mod
    use vector_mod
    use fen_capi
    use iso_c_binding, only : c_ptr, c_null_ptr, c_associated, c_int, c_double, c_size_t
    implicit none

    real(dp) :: density = 1.0_dp, viscosity = 1.0_dp
    real(dp) :: g(3) = [0.0_dp, 0.0_dp, 0.0_dp]
    real(dp) :: maxdiv = 0.0_dp, maxCFL = 0.0_dp, CFL = 1.0_dp
    real(dp) :: dt_visc, dt_conv
    real(dp) :: dt_o

    type(scalar) :: p, phi, rho, mu
    type(vector) :: v, dv, dv_o, grad_p, S, rhof
#ifdef MF
    !> Volume fraction of phase 1 (re-exported by volume_of_fluid_mod).
    type(scalar) :: vof
    !> Order of the vof's directional sweeps, x first when .true.
    !> (src/volume_of_fluid.f90:30; re-exported by volume_of_fluid_mod).  Each
    !> advection flips it; the device's copy is set from it before each step
    !> and read back after.
    logical :: x_first = .true.
    logical, private :: multiphase = .false.
#endif

    logical :: constant_CFL = .false.
    logical :: constant_viscosity = .true.

    !> Arithmetic of the GPU step: FEN_FP64 (default) or FEN_FP32_REFINE.
    integer(c_int) :: fen_precision = FEN_FP64

    !> .true.: navier_stokes_solver queues the step without waiting for its
    !> checks; maxdiv and maxCFL are then fetched by print_solver_status or
    !> fen_sync_checks() (no counterpart in the reference, whose values are
    !> always current).  .false. (default): both are current after each step.
    logical :: lazy_checks = .false.

    type(c_ptr) :: fen_ctx = c_null_ptr
    logical, private :: host_modified = .true.   ! host mirrors hold data the device lacks
    logical, private :: device_ahead = .false.   ! device holds data the host mirrors lack
    logical, private :: checks_stale = .false.   ! maxdiv, maxCFL not fetched since the last step
    logical, private :: source_resident = .false.   ! S has been uploaded once

contains

    !> Allocates the fields, maps the physical boundary conditions onto them
    !> (src/navier_stokes.f90:752-1067) and creates the GPU context.
    subroutine allocate_navier_stokes_fields(comp_grid)
        use global_mod, only : Ndim
        use fen_bc_codes, only : physical_bc_codes
        use poisson_mod, only : poisson_shared_ctx
        use IO_mod, only : print_error_message
        type(grid), intent(in), target :: comp_grid
        integer(c_int) :: codes(6), rc

        call p%allocate(comp_grid, 1)
        call phi%allocate(comp_grid, 1)
        call rho%allocate(comp_grid, 1)
        call rhof%allocate(comp_grid)
        call mu%allocate(comp_grid, 1)
        call v%allocate(comp_grid, 1)
        call dv%allocate(comp_grid)
        call dv_o%allocate(comp_grid)
        call grad_p%allocate(comp_grid)
        call S%allocate(comp_grid)
        rho%f = density
        mu%f = viscosity
#ifdef MF
        call vof%allocate(comp_grid, 1)
#endif

        call physical_bc_codes(comp_grid, codes)
        call map_face(1, codes(1))
        call map_face(2, codes(2))
        call map_face(3, codes(3))
        call map_face(4, codes(4))
#if DIM==3
        call map_face(5, codes(5))
        call map_face(6, codes(6))
#endif
        rc = fen_create(fen_ctx, Ndim, comp_grid%Nx, comp_grid%Ny, comp_grid%Nz, comp_grid%Lx, comp_grid%Ly, &
                        comp_grid%Lz, comp_grid%origin, codes, fen_precision)
        if (rc /= FEN_OK) then
            call print_error_message('Unable to find the proped poisson solver with the selected boundary conditions.')
            error stop 1
        end if
        p%fid = FEN_P
        phi%fid = FEN_PHI
        v%x%fid = FEN_U
        v%y%fid = FEN_V
#if DIM==3
        v%z%fid = FEN_W
#endif
        dv_o%x%fid = FEN_DVO_X
        dv_o%y%fid = FEN_DVO_Y
        dv_o%z%fid = FEN_DVO_Z
        poisson_shared_ctx = fen_ctx
        scalar_sync_hook => fen_sync_host
        host_modified = .true.
        device_ahead = .false.
    end subroutine allocate_navier_stokes_fields

    !> bc types of every field on one face (src/navier_stokes.f90:779-1017).
    subroutine map_face(face, code)
        integer, intent(in) :: face
        integer(c_int), intent(in) :: code
        integer :: tp, tu, tv, tw
        select case (code)
        case (FEN_BC_PERIODIC)
            tp = 0; tu = 0; tv = 0; tw = 0
        case (FEN_BC_WALL)
            tp = 2; tu = 1; tv = 1; tw = 1
        case (FEN_BC_INFLOW)
            tp = 2; tu = 1; tv = 1; tw = 1
            if (face <= 2) then
                tv = 2; tw = 2
            end if
        case default   ! Outflow
            tp = 1; tu = 2; tv = 2; tw = 2
        end select
        call set_type(p, face, tp)
        call set_type(phi, face, tp)
        call set_type(rho, face, 2)
        call set_type(mu, face, 2)
        if (code == FEN_BC_PERIODIC) then
            call set_type(rho, face, 0)
            call set_type(mu, face, 0)
        end if
#ifdef MF
        ! Periodic or Neumann (src/volume_of_fluid.f90:69-219)
        if (code == FEN_BC_PERIODIC) then
            call set_type(vof, face, 0)
        else
            call set_type(vof, face, 2)
        end if
#endif
        call set_type(v%x, face, tu)
        call set_type(v%y, face, tv)
        call set_type(v%z, face, tw)
    end subroutine map_face

    subroutine set_type(s, face, t)
        type(scalar), intent(inout) :: s
        integer, intent(in) :: face, t
        select case (face)
        case (1); s%bc%type_left = t
        case (2); s%bc%type_right = t
        case (3); s%bc%type_bottom = t
        case (4); s%bc%type_top = t
        case (5); s%bc%type_front = t
        case (6); s%bc%type_back = t
        end select
    end subroutine set_type

    subroutine push_params()
        integer(c_int) :: ccfl
        ccfl = 0
        if (constant_CFL) ccfl = 1
        call fen_check(fen_ctx, fen_set_params(fen_ctx, density, viscosity, g, CFL, ccfl), 'fen_set_params')
    end subroutine push_params

    subroutine upload(s)
        type(scalar), intent(in) :: s
        call fen_check(fen_ctx, fen_upload(fen_ctx, int(s%fid, c_int), s%f, int(size(s%f), c_size_t)), 'fen_upload')
    end subroutine upload

    subroutine download(s)
        type(scalar), intent(inout) :: s
        call fen_check(fen_ctx, fen_download(fen_ctx, int(s%fid, c_int), s%f, int(size(s%f), c_size_t)), &
                       'fen_download')
    end subroutine download

    subroutine upload_bc(s)
        type(scalar), intent(in) :: s
        integer(c_int) :: fid
        fid = int(s%fid, c_int)
        call fen_check(fen_ctx, fen_set_bc_type(fen_ctx, fid, 0, int(s%bc%type_left, c_int)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_type(fen_ctx, fid, 1, int(s%bc%type_right, c_int)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_type(fen_ctx, fid, 2, int(s%bc%type_bottom, c_int)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_type(fen_ctx, fid, 3, int(s%bc%type_top, c_int)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_plane(fen_ctx, fid, 0, s%bc%left, int(size(s%bc%left), c_size_t)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_plane(fen_ctx, fid, 1, s%bc%right, int(size(s%bc%right), c_size_t)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_plane(fen_ctx, fid, 2, s%bc%bottom, int(size(s%bc%bottom), c_size_t)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_plane(fen_ctx, fid, 3, s%bc%top, int(size(s%bc%top), c_size_t)), 'bc')
#if DIM==3
        call fen_check(fen_ctx, fen_set_bc_type(fen_ctx, fid, 4, int(s%bc%type_front, c_int)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_type(fen_ctx, fid, 5, int(s%bc%type_back, c_int)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_plane(fen_ctx, fid, 4, s%bc%front, int(size(s%bc%front), c_size_t)), 'bc')
        call fen_check(fen_ctx, fen_set_bc_plane(fen_ctx, fid, 5, s%bc%back, int(size(s%bc%back), c_size_t)), 'bc')
#endif
    end subroutine upload_bc

    !> Host mirrors -> HBM (fields, boundary types and values, source term).
    subroutine fen_sync_device()
        integer(c_int) :: fsrc
        logical :: nonzero
        call upload_bc(p)
        call upload_bc(phi)
        call upload_bc(v%x)
        call upload_bc(v%y)
        call upload(p)
        call upload(phi)
        call upload(v%x)
        call upload(v%y)
        call upload(dv_o%x)
        call upload(dv_o%y)
#if DIM==3
        call upload_bc(v%z)
        call upload(v%z)
        call upload(dv_o%z)
#endif
#ifdef MF
        if (multiphase) then
            call upload(vof)
            call fen_check(fen_ctx, fen_update_ghost_nodes(fen_ctx, FEN_VOF), 'vof ghost nodes')
        end if
#endif
        ! S is uploaded once it is non-zero in any component, and then every
        ! time, so that a source switched off reaches the device as zeros
        nonzero = any(S%x%f /= 0.0_dp) .or. any(S%y%f /= 0.0_dp)
#if DIM==3
        nonzero = nonzero .or. any(S%z%f /= 0.0_dp)
#endif
        if (nonzero .or. source_resident) then
            source_resident = .true.
            fsrc = FEN_S_X
            call fen_check(fen_ctx, fen_upload(fen_ctx, fsrc, S%x%f, int(size(S%x%f), c_size_t)), 'S')
            fsrc = FEN_S_Y
            call fen_check(fen_ctx, fen_upload(fen_ctx, fsrc, S%y%f, int(size(S%y%f), c_size_t)), 'S')
#if DIM==3
            fsrc = FEN_S_Z
            call fen_check(fen_ctx, fen_upload(fen_ctx, fsrc, S%z%f, int(size(S%z%f), c_size_t)), 'S')
#endif
        end if
        host_modified = .false.
    end subroutine fen_sync_device

    !> HBM -> host mirrors (p, phi, v, dv_o).
    subroutine fen_sync_host()
        if (.not. device_ahead) return
        device_ahead = .false.
        call download(p)
        call download(phi)
        call download(v%x)
        call download(v%y)
        call download(dv_o%x)
        call download(dv_o%y)
#if DIM==3
        call download(v%z)
        call download(dv_o%z)
#endif
#ifdef MF
        if (multiphase) then
            call download(vof)
            call download(rho)
            call download(mu)
        end if
#endif
    end subroutine fen_sync_host

    !> Uploads the host mirrors if they hold data the device lacks.
    subroutine fen_push_host()
        if (host_modified) call fen_sync_device()
    end subroutine fen_push_host

    !> Declares that a library call changed resident fields.
    subroutine fen_device_modified()
        device_ahead = .true.
    end subroutine fen_device_modified

#ifdef MF
    !> Makes the context a two-phase one (fen_set_multiphase) and the vof,
    !> rho and mu mirrors resident fields.
    subroutine enable_multiphase(rho_0, rho_1, mu_0, mu_1, sigma, beta)
        real(dp), intent(in) :: rho_0, rho_1, mu_0, mu_1, sigma, beta
        call fen_check(fen_ctx, fen_set_multiphase(fen_ctx, rho_0, rho_1, mu_0, mu_1, sigma, beta), &
                       'fen_set_multiphase')
        vof%fid = FEN_VOF
        rho%fid = FEN_RHO
        mu%fid = FEN_MU
        multiphase = .true.
        host_modified = .true.
        x_first = .true.
    end subroutine enable_multiphase

    !> x_first to the device before an advection, back after it.
    subroutine push_x_first()
        integer(c_int) :: xf
        xf = 0
        if (x_first) xf = 1
        call fen_check(fen_ctx, fen_set_vof_x_first(fen_ctx, xf), 'fen_set_vof_x_first')
    end subroutine push_x_first

    subroutine pull_x_first()
        integer(c_int) :: xf
        call fen_check(fen_ctx, fen_get_vof_x_first(fen_ctx, xf), 'fen_get_vof_x_first')
        x_first = xf /= 0
    end subroutine pull_x_first
#endif

    !> Declares that user code changed host mirrors (fields, bc, S) mid-run.
    subroutine fen_host_modified()
        call fen_sync_host()
        host_modified = .true.
    end subroutine fen_host_modified

    !> One time step (src/navier_stokes.f90:50-136).
    subroutine navier_stokes_solver(comp_grid, step, dt)
        type(grid), intent(in) :: comp_grid
        integer, intent(in) :: step
        real(dp), intent(inout) :: dt
        real(c_double) :: d, md, mc, dto, dtv, dtc
        if (host_modified) call fen_sync_device()
        call push_params()
        call fen_check(fen_ctx, fen_set_dt_o(fen_ctx, dt_o), 'fen_set_dt_o')
        d = dt
#ifdef MF
        if (multiphase) call push_x_first()
#endif
        if (lazy_checks) then
            call fen_check(fen_ctx, fen_ns_step(fen_ctx, int(step, c_int), d), 'navier_stokes_solver')
            checks_stale = .true.
        else
            call fen_check(fen_ctx, fen_ns_step(fen_ctx, int(step, c_int), d, md, mc), 'navier_stokes_solver')
            maxdiv = md
            maxCFL = mc
            checks_stale = .false.
        end if
        dt = d
#ifdef MF
        if (multiphase) call pull_x_first()
#endif
        call fen_check(fen_ctx, fen_get_scalars(fen_ctx, dt_o=dto, dt_visc=dtv, dt_conv=dtc), 'fen_get_scalars')
        dt_o = dto
        if (constant_CFL) then
            dt_conv = dtc
        end if
        device_ahead = .true.
        if (comp_grid%Nx < 0) print *, 'invalid grid'
    end subroutine navier_stokes_solver

    !> dt = min(CFL*delta/U, dt_visc) and dt_o = dt (src/navier_stokes.f90:623-666).
    subroutine set_timestep(comp_grid, dt, U)
        type(grid), intent(in) :: comp_grid
        real(dp), intent(in) :: U
        real(dp), intent(inout) :: dt
        real(c_double) :: d, dto, dtv, dtc
        call push_params()
        call fen_check(fen_ctx, fen_set_timestep(fen_ctx, U, d), 'set_timestep')
        call fen_check(fen_ctx, fen_get_scalars(fen_ctx, dt_o=dto, dt_visc=dtv, dt_conv=dtc), 'fen_get_scalars')
        dt = d
        dt_o = dto
        dt_visc = dtv
        dt_conv = dtc
        if (comp_grid%Nx < 0) print *, 'invalid grid'
    end subroutine set_timestep

    !> Fetches maxdiv and maxCFL of the last step (waits for it).
    subroutine fen_sync_checks()
        real(c_double) :: md, mc
        if (.not. checks_stale) return
        call fen_check(fen_ctx, fen_get_scalars(fen_ctx, maxdiv=md, maxcfl=mc), 'fen_get_scalars')
        maxdiv = md
        maxCFL = mc
        checks_stale = .false.
    end subroutine fen_sync_checks

    subroutine print_navier_stokes_solver_status(log_id, step, time, dt)
        use global_mod, only : myrank
        integer, intent(in) :: log_id, step
        real(dp), intent(in) :: time, dt
        call fen_sync_checks()
        if (myrank == 0) write(log_id, 10) 'step: ', step, 'time: ', time, 'dt: ', dt, &
             'maxdiv: ', maxdiv, 'maxCFL: ', maxCFL
10      format(A6,I7,1x,A6,E13.6,1x,A4,E13.6,1x,A8,E13.6,1x,A9,1x,E13.6)
    end subroutine print_navier_stokes_solver_status

    subroutine destroy_navier_stokes_solver()
        use poisson_mod, only : poisson_shared_ctx
        integer(c_int) :: rc
        call p%destroy()
        call phi%destroy()
        call rho%destroy()
        call mu%destroy()
        call v%destroy()
        call dv%destroy()
        call dv_o%destroy()
        call grad_p%destroy()
        call S%destroy()
        call rhof%destroy()
#ifdef MF
        call vof%destroy()
        multiphase = .false.
#endif
        if (c_associated(fen_ctx)) rc = fen_destroy(fen_ctx)
        fen_ctx = c_null_ptr
        source_resident = .false.
        checks_stale = .false.
        poisson_shared_ctx = c_null_ptr
        scalar_sync_hook => Null()
    end subroutine destroy_navier_stokes_solver

end module navier_stokes_mod

#ifdef MF
!===================================================================================================
module volume_of_fluid_mod
    ! MTHINC volume of fluid (reference src/volume_of_fluid.f90).  The vof
    ! field is navier_stokes_mod's resident one; reconstruction and advection
    ! run on the GPU inside the step, or alone through advect_vof.
    use precision_mod, only : dp
    use scalar_mod
    use vector_mod
    use navier_stokes_mod, only : vof, x_first
    implicit none
    private
    public :: vof, x_first, beta, distance, distance_from_interface, get_vof_from_distance, advect_vof

    !> Interface sharpness of the MTHINC profile (src/volume_of_fluid.f90:24).
    real(dp) :: beta = 1.0_dp

    abstract interface
        function distance_from_interface(x, y) result(dist)
            use precision_mod, only : dp
            real(dp), intent(in) :: x, y
            real(dp) :: dist
        end function distance_from_interface
    end interface

    !> Signed distance from the interface, set by the user program.
    procedure(distance_from_interface), pointer :: distance => Null()

contains

    !> vof = cell average of 0.5 (1 + tanh(beta d / delta)) by the 2 x 2 Gauss
    !> rule, on the host mirror, then the ghost nodes
    !> (src/volume_of_fluid.f90:665-706).
    subroutine get_vof_from_distance()
        use IO_mod, only : print_error_message
        use navier_stokes_mod, only : fen_host_modified
        ! Gauss points at 0.5 (1 -+ 1/sqrt(3)) of the cell, averaged pairwise in x, then in y
        real(dp), parameter :: gm = 0.5_dp*(1.0_dp - 1.0_dp/sqrt(3.0_dp)) - 0.5_dp
        real(dp), parameter :: gq = 0.5_dp*(1.0_dp + 1.0_dp/sqrt(3.0_dp)) - 0.5_dp
        real(dp) :: dl, xm, xp, ym, yp
        integer :: i, j, k
        if (.not. associated(distance)) then
            call print_error_message('ERROR: distance function not defined.')
            return
        end if
        call fen_host_modified()
        dl = vof%G%delta
        do k = vof%G%lo(3), vof%G%hi(3)
            do j = vof%G%lo(2), vof%G%hi(2)
                ym = vof%G%y(j) + dl*gm
                yp = vof%G%y(j) + dl*gq
                do i = vof%G%lo(1), vof%G%hi(1)
                    xm = vof%G%x(i) + dl*gm
                    xp = vof%G%x(i) + dl*gq
                    vof%f(i,j,k) = 0.5_dp*(0.5_dp*(prof(xm, ym) + prof(xp, ym)) + 0.5_dp*(prof(xm, yp) + prof(xp, yp)))
                end do
            end do
        end do
        call vof%update_ghost_nodes()
    contains
        real(dp) function prof(x, y)
            real(dp), intent(in) :: x, y
            prof = 0.5_dp*(1.0_dp + tanh(beta*distance(x, y)/dl))
        end function prof
    end subroutine get_vof_from_distance

    !> Advects the vof with the resident velocity over dt
    !> (src/volume_of_fluid.f90:420-544); v must be navier_stokes_mod's v.
    subroutine advect_vof(v, dt)
        use fen_capi
        use navier_stokes_mod, only : fen_ctx, fen_push_host, fen_device_modified, push_x_first, pull_x_first
        type(vector), intent(in) :: v
        real(dp), intent(in) :: dt
        if (.not. associated(v%G)) return
        call fen_push_host()
        call push_x_first()
        call fen_check(fen_ctx, fen_advect_vof(fen_ctx, dt), 'advect_vof')
        call pull_x_first()
        call fen_device_modified()
    end subroutine advect_vof

end module volume_of_fluid_mod

!===================================================================================================
module multiphase_mod
    ! Two-fluid material properties (reference src/multiphase.f90).
    use precision_mod, only : dp
    use scalar_mod
    use vector_mod
    use navier_stokes_mod, only : rho, mu
    implicit none

    !> Densities and viscosities of phase 0 (vof = 0) and phase 1 (vof = 1).
    real(dp) :: rho_0 = 1.0_dp, rho_1 = 1.0_dp, mu_0 = 1.0_dp, mu_1 = 1.0_dp
    !> Surface tension coefficient.
    real(dp) :: sigma = 0.0_dp
    !> Minimum density and its inverse (Dodd & Ferrante splitting).
    real(dp) :: rhomin, irhomin

    interface
        subroutine interface_advection(v, dt)
            use precision_mod, only : dp
            use vector_mod
            real(dp), intent(in) :: dt
            type(vector), intent(in) :: v
        end subroutine interface_advection
    end interface

    procedure(interface_advection), pointer :: advect_interface => Null()

contains

    !> rho and mu from the vof, on the GPU (src/multiphase.f90:126-140).
    subroutine update_material_properties(rho, mu, vof)
        use fen_capi
        use navier_stokes_mod, only : fen_ctx, fen_push_host, fen_device_modified
        type(scalar), intent(inout) :: rho, mu
        type(scalar), intent(in) :: vof
        if (rho%fid /= FEN_RHO .or. mu%fid /= FEN_MU .or. vof%fid /= FEN_VOF) return
        call fen_push_host()
        call fen_check(fen_ctx, fen_update_material_properties(fen_ctx), 'update_material_properties')
        call fen_device_modified()
    end subroutine update_material_properties

end module multiphase_mod
#endif

!===================================================================================================
module solver_mod
    ! Solver driver (reference src/solver.f90).
    use precision_mod, only : dp
    use grid_mod, only : grid
    implicit none

    interface
        subroutine advance_solver(comp_grid, step, dt)
            use precision_mod, only : dp
            use grid_mod, only : grid
            type(grid), intent(in) :: comp_grid
            integer, intent(in) :: step
            real(dp), intent(inout) :: dt
        end subroutine advance_solver
        subroutine solver_status(log_id, step, time, dt)
            use precision_mod, only : dp
            integer, intent(in) :: log_id, step
            real(dp), intent(in) :: time, dt
        end subroutine solver_status
    end interface

    procedure(advance_solver), pointer :: advance_solution => Null()
    procedure(solver_status), pointer :: print_solver_status => Null()

contains

    !> Allocates the fields and the GPU context; in the two-phase build also
    !> the vof from the user's distance function and rho, mu from it
    !> (src/solver.f90:32-97).
    subroutine init_solver(comp_grid)
        use poisson_mod, only : init_poisson_solver
        use navier_stokes_mod, only : allocate_navier_stokes_fields, phi, navier_stokes_solver, &
                                      print_navier_stokes_solver_status
#ifdef MF
        use navier_stokes_mod, only : constant_viscosity, rho, mu, enable_multiphase, fen_host_modified
        use multiphase_mod, only : rho_0, rho_1, mu_0, mu_1, sigma, rhomin, irhomin, &
                                   update_material_properties, advect_interface
        use volume_of_fluid_mod, only : vof, beta, get_vof_from_distance, advect_vof
#endif
        type(grid), intent(in), target :: comp_grid
        call allocate_navier_stokes_fields(comp_grid)
        call init_Poisson_Solver(phi)
        advance_solution => navier_stokes_solver
        print_solver_status => print_navier_stokes_solver_status
#ifdef MF
        constant_viscosity = .false.
        call enable_multiphase(rho_0, rho_1, mu_0, mu_1, sigma, beta)
        call get_vof_from_distance
        call update_material_properties(rho, mu, vof)
        ! refresh the mirrors; fields the program sets after init_solver are uploaded by the first step
        call fen_host_modified()
        rhomin = min(rho_0, rho_1)
        irhomin = 1.0_dp/rhomin
        advect_interface => advect_vof
#endif
    end subroutine init_solver

    !> Cell-centred fields data/vx_<step>.raw, vy_, [vz_,] p_ (raw float64
    !> interiors; reference src/solver.f90:103-156).  dir replaces data/
    !> (no counterpart in the reference, whose folder is fixed).
    subroutine save_fields(step, dir)
        use scalar_mod, only : scalar
        use fields_mod, only : face_to_center
        use navier_stokes_mod, only : v, p, fen_sync_host
#ifdef MF
        use navier_stokes_mod, only : vof
#endif
        integer, intent(in) :: step
        character(len=*), intent(in), optional :: dir
        type(scalar) :: temp
        character(len=7) :: sn
        character(len=:), allocatable :: d
        d = 'data'
        if (present(dir)) d = trim(dir)
        call fen_sync_host()
        call temp%allocate(v%G, v%x%gl)
        write(sn, '(I0.7)') step
        call face_to_center(v%x, temp, 'x')
        call temp%write(d//'/vx_'//sn//'.raw')
        call face_to_center(v%y, temp, 'y')
        call temp%write(d//'/vy_'//sn//'.raw')
#if DIM==3
        call face_to_center(v%z, temp, 'z')
        call temp%write(d//'/vz_'//sn//'.raw')
#endif
        call p%write(d//'/p_'//sn//'.raw')
#ifdef MF
        call vof%write(d//'/vof_'//sn//'.raw')
#endif
        call temp%destroy()
    end subroutine save_fields

    !> Restart file data/state_<step>.raw: p, u, v, dv_o%x, dv_o%y [, w, dv_o%z]
    !> [, vof, rho, mu, p_o in the two-phase build], each the raw float64
    !> interior in Fortran order (src/solver.f90:160-329), written and read by
    !> the library from the resident fields (fen_save_state, fen_load_state).
    !> dir replaces data/ (no counterpart in the reference).
    subroutine save_state(step, dir)
        use iso_c_binding, only : c_null_char
        use fen_capi, only : fen_save_state, fen_check
        use navier_stokes_mod, only : fen_ctx, fen_push_host
        integer, intent(in) :: step
        character(len=*), intent(in), optional :: dir
        character(len=7) :: sn
        character(len=:), allocatable :: d
        d = 'data'
        if (present(dir)) d = trim(dir)
        write(sn, '(I0.7)') step
        call fen_push_host()
        call fen_check(fen_ctx, fen_save_state(fen_ctx, d//'/state_'//sn//'.raw'//c_null_char), 'save_state')
    end subroutine save_state

    subroutine load_state(step, dir)
        use iso_c_binding, only : c_null_char
        use fen_capi, only : fen_load_state, fen_check
        use navier_stokes_mod, only : fen_ctx, fen_push_host, fen_device_modified
        integer, intent(in) :: step
        character(len=*), intent(in), optional :: dir
        character(len=7) :: sn
        character(len=:), allocatable :: d
        d = 'data'
        if (present(dir)) d = trim(dir)
        write(sn, '(I0.7)') step
        ! what the host holds goes first, so the loaded state is not overwritten by it
        call fen_push_host()
        call fen_check(fen_ctx, fen_load_state(fen_ctx, d//'/state_'//sn//'.raw'//c_null_char), 'load_state')
        ! the ghost fills of p and v (and the two-phase fields) ran on the device
        call fen_device_modified()
    end subroutine load_state

    subroutine destroy_solver
        use poisson_mod, only : destroy_poisson_solver
        use navier_stokes_mod, only : phi, destroy_navier_stokes_solver
        call destroy_poisson_solver(phi)
        call destroy_navier_stokes_solver
        advance_solution => Null()
        print_solver_status => Null()
    end subroutine destroy_solver

end module solver_mod

!===================================================================================================
module case_mod
    ! Set-up and time loop of a run from a case file (a JSON description read
    ! by libfen's fen_case_read; format in include/fen.h).  setup_case does
    ! what a FEN program's own set-up code does: grid%setup, init_solver,
    ! navier_stokes_mod's module variables, the initial fields or a restart,
    ! set_timestep.  run_case is the program's time loop: advance_solution,
    ! print_solver_status, save_state and save_fields at the listed steps.
    ! No counterpart in the reference, which has no input file.
    use iso_c_binding, only : c_char, c_null_char, c_size_t, c_int
    use precision_mod, only : dp
    use fen_capi, only : fen_case, fen_case_read, FEN_OK, FEN_FP64, FEN_BC_PERIODIC, FEN_BC_WALL, &
                         FEN_BC_INFLOW, FEN_BC_OUTFLOW
    implicit none
    private
    public :: fen_case, read_case, setup_case, run_case, case_text

    !> The case being set up (the two-phase interface function reads it).
    type(fen_case), save :: current

contains

    !> A NUL-terminated C text as a Fortran string.
    function case_text(chars) result(s)
        character(kind=c_char), intent(in) :: chars(:)
        character(len=:), allocatable :: s
        integer :: n, i
        n = 0
        do while (n < size(chars))
            if (chars(n + 1) == c_null_char) exit
            n = n + 1
        end do
        allocate(character(len=n) :: s)
        do i = 1, n
            s(i:i) = chars(i)
        end do
    end function case_text

    !> Reads a case file; stops with the reader's message if it is not valid.
    subroutine read_case(path, cs)
        use IO_mod, only : print_error_message
        character(len=*), intent(in) :: path
        type(fen_case), intent(out) :: cs
        character(kind=c_char) :: err(512)
        integer(c_int) :: rc
        rc = fen_case_read(trim(path)//c_null_char, cs, err, int(size(err), c_size_t))
        if (rc /= FEN_OK) then
            call print_error_message(case_text(err))
            error stop 1
        end if
    end subroutine read_case

    function bc_name(code) result(s)
        integer(c_int), intent(in) :: code
        character(len=:), allocatable :: s
        select case (code)
        case (FEN_BC_WALL)
            s = 'Wall'
        case (FEN_BC_INFLOW)
            s = 'Inflow'
        case (FEN_BC_OUTFLOW)
            s = 'Outflow'
        case default
            s = 'Periodic'
        end select
    end function bc_name

#ifdef MF
    !> Signed distance y - level - amplitude cos(2 pi (x - x0) / wavelength)
    !> of the case's interface (positive in the phase above it).
    function case_interface(x, y) result(d)
        use global_mod, only : pi
        real(dp), intent(in) :: x, y
        real(dp) :: d
        d = y - current%iface(2)*cos(2.0_dp*pi*(x - current%origin(1))/current%iface(3)) - current%iface(1)
    end function case_interface
#endif

    !> Sets a run up as the case describes.  On return the fields hold the
    !> initial state (or the restart state of step step0), dt is the time
    !> step and dt_o is set.
    subroutine setup_case(cs, comp_grid, dt, step0)
        use IO_mod, only : print_error_message
        use global_mod, only : Ndim
        use grid_mod, only : grid, bc_type
        use solver_mod, only : init_solver, load_state
        use navier_stokes_mod, only : density, viscosity, g, CFL, constant_CFL, fen_precision, v, p, &
                                      set_timestep, dt_o, fen_host_modified
#ifdef MF
        use navier_stokes_mod, only : rho, mu
        use multiphase_mod, only : rho_0, rho_1, mu_0, mu_1, sigma, update_material_properties
        use volume_of_fluid_mod, only : distance, vof, x_first
#endif
        type(fen_case), intent(in) :: cs
        type(grid), intent(inout), target :: comp_grid
        real(dp), intent(out) :: dt
        integer, intent(out) :: step0
        type(bc_type) :: bc(Ndim*2)
        integer :: f

        current = cs
        if (cs%dim /= Ndim) then
            call print_error_message('case file: its grid is not of this build''s dimension (DIM)')
            error stop 1
        end if
#ifdef MF
        if (cs%multiphase == 0) then
            call print_error_message('case file: the two-phase build needs a "Multiphase" object')
            error stop 1
        end if
#else
        if (cs%multiphase /= 0) then
            call print_error_message('case file: "Multiphase" needs the two-phase build (-DMF)')
            error stop 1
        end if
#endif
        do f = 1, Ndim*2
            bc(f)%s = bc_name(cs%bc(f))
        end do
        comp_grid%name = case_text(cs%name)
        call comp_grid%setup(cs%n(1), cs%n(2), cs%n(3), cs%L(1), cs%L(2), cs%L(3), cs%origin, 1, 1, bc)

        density = cs%density
        viscosity = cs%viscosity
        g = cs%g
        CFL = cs%cfl
        constant_CFL = cs%constant_cfl /= 0
        fen_precision = cs%precision
#ifdef MF
        rho_0 = cs%rho0
        rho_1 = cs%rho1
        mu_0 = cs%mu0
        mu_1 = cs%mu1
        sigma = cs%sigma
        distance => case_interface
#endif
        call init_solver(comp_grid)

        step0 = 0
        if (cs%restart >= 0) then
            call load_state(cs%restart, case_text(cs%restart_dir))
            step0 = cs%restart
#ifdef MF
            ! each advection flips the sweep order, which starts x first
            x_first = mod(step0, 2) == 0
#endif
        else
            if (len(case_text(cs%init(:, 1))) > 0) call v%x%read(case_text(cs%init(:, 1)))
            if (len(case_text(cs%init(:, 2))) > 0) call v%y%read(case_text(cs%init(:, 2)))
#if DIM==3
            if (len(case_text(cs%init(:, 3))) > 0) call v%z%read(case_text(cs%init(:, 3)))
#endif
            if (len(case_text(cs%init(:, 4))) > 0) call p%read(case_text(cs%init(:, 4)))
            call p%update_ghost_nodes()
            call v%update_ghost_nodes()
            call fen_host_modified()
        end if
#ifdef MF
        if (cs%restart < 0 .and. len(case_text(cs%init(:, 5))) > 0) then
            call vof%read(case_text(cs%init(:, 5)))
            call vof%update_ghost_nodes()
            call fen_host_modified()
            call update_material_properties(rho, mu, vof)
        end if
#endif
        if (cs%U > 0.0_dp) then
            call set_timestep(comp_grid, dt, cs%U)
        else
            dt = cs%dt
            dt_o = dt
        end if
    end subroutine setup_case

    !> The time loop: steps step0+1 .. step0+steps with the status line every
    !> status_every steps, save_state and save_fields at the listed steps
    !> (into the case's data_dir), and the final interiors (u.raw, v.raw,
    !> [w.raw,] p.raw [, vof.raw]) when write_final is set.  time starts at
    !> step0 * dt.
    subroutine run_case(cs, comp_grid, dt, step0)
        use IO_mod, only : stdout
        use grid_mod, only : grid
        use solver_mod, only : advance_solution, print_solver_status, save_state, save_fields
        use navier_stokes_mod, only : v, p
#ifdef MF
        use volume_of_fluid_mod, only : vof
#endif
        type(fen_case), intent(in) :: cs
        type(grid), intent(in) :: comp_grid
        real(dp), intent(inout) :: dt
        integer, intent(in) :: step0
        character(len=:), allocatable :: dir
        integer :: step
        real(dp) :: time

        dir = case_text(cs%data_dir)
        time = step0*dt
        do step = step0 + 1, step0 + cs%steps
            time = time + dt
            call advance_solution(comp_grid, step, dt)
            if (cs%status_every > 0) then
                if (mod(step - step0, cs%status_every) == 0) call print_solver_status(stdout, step, time, dt)
            end if
            if (any(cs%save_state(1:cs%n_save_state) == step)) call save_state(step, dir)
            if (any(cs%save_fields(1:cs%n_save_fields) == step)) call save_fields(step, dir)
        end do
        if (cs%write_final /= 0) then
            call v%x%write('u.raw')
            call v%y%write('v.raw')
#if DIM==3
            call v%z%write('w.raw')
#endif
            call p%write('p.raw')
#ifdef MF
            call vof%write('vof.raw')
#endif
        end if
    end subroutine run_case

end module case_mod
