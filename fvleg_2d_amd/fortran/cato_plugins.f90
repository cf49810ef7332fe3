!> The reference's plugin interfaces for the per-timestep hot path, backed by
!> the HIP library through the C ABI.
!>
!>   flux_solver_t%solve / %initialize      src/lib/flux_solvers/flux_solver.f90:40-96
!>   muscl_interpolation_t%interpolate_edge_values
!>                                          src/lib/spatial_reconstruction/muscl_interpolation.f90:19-53
!>   boundary_condition_t%apply             src/lib/boundary_conditions/boundary_conditions.f90:35-70
!>   fluid_t%integrate / %get_timestep      src/lib/fluid/mod_fluid.f90:411-562
!>
!> The abstract types restate those interfaces with the field data as plain
!> halo'd arrays (field_2d_t%data, lower bounds 1-nh); the hip_* types are
!> their implementations.  All plugin objects of one solve share one device
!> context (hip_solve_t), which owns the resident state.  Errors the
!> reference reports with `error stop` stop the program the same way.
module cato_hip_plugins
  use, intrinsic :: iso_c_binding
  use cato_hip_api
  use cato_deck, only: cato_input_t
  implicit none
  private

  integer, parameter, public :: rk = c_double
  integer, parameter, public :: ALL_OK = 0, NEG_DENSITY = 1, NEG_PRESSURE = 2, NANS_FOUND = 3  ! error.f90:6-9

  public :: hip_grid_t, hip_solve_t, flux_solver_t, hip_flux_solver_t, muscl_interpolation_t, hip_muscl_t, &
            boundary_condition_t, hip_boundary_conditions_t, hip_fluid_t, config_from_input

  !> The part of grid_block_2d_t the path reads: cell bounds and node coordinates incl. ghost nodes
  type :: hip_grid_t
    integer :: ni = 0, nj = 0, n_halo = 0
    integer :: ilo = 1, ihi = 0, jlo = 1, jhi = 0
    real(rk), allocatable :: node_x(:, :), node_y(:, :)  !< (ilo-nh:ihi+nh+1, jlo-nh:jhi+nh+1)
  end type

  !> One device-resident solve (cato_ctx)
  type :: hip_solve_t
    type(c_ptr) :: ctx = c_null_ptr
    type(cato_config_t) :: cfg
    integer :: ni = 0, nj = 0, nh = 0
  contains
    procedure :: create => solve_create
    procedure :: destroy => solve_destroy
  end type

  ! ---------------------------------------------------------------- flux solver
  type, abstract :: flux_solver_t
    character(len=32) :: name = ''
    integer :: iteration = 0
    real(rk) :: time = 0.0_rk
    real(rk) :: dt = 0.0_rk
    integer, dimension(2) :: lbounds = 0  !< (i,j) lower cell bounds
    integer, dimension(2) :: ubounds = 0  !< (i,j) upper cell bounds
  contains
    procedure(flux_initialize_i), deferred :: initialize
    procedure(flux_solve_i), deferred :: solve
  end type

  abstract interface
    subroutine flux_initialize_i(self, input, time)
      import :: flux_solver_t, cato_input_t, rk
      class(flux_solver_t), intent(inout) :: self
      class(cato_input_t), intent(in) :: input
      real(rk), intent(in) :: time
    end subroutine
    subroutine flux_solve_i(self, dt, grid, rho, u, v, p, d_rho_dt, d_rho_u_dt, d_rho_v_dt, d_rho_E_dt)
      import :: flux_solver_t, hip_grid_t, rk
      class(flux_solver_t), intent(inout) :: self
      real(rk), intent(in) :: dt
      class(hip_grid_t), intent(in) :: grid
      real(rk), intent(inout) :: rho(:, :), u(:, :), v(:, :), p(:, :)
      real(rk), allocatable, intent(out) :: d_rho_dt(:, :), d_rho_u_dt(:, :), d_rho_v_dt(:, :), d_rho_E_dt(:, :)
    end subroutine
  end interface

  !> flux_solver_t backed by cato_residual (BC fill + reconstruction + edge flux + flux_split_edges)
  type, extends(flux_solver_t) :: hip_flux_solver_t
    type(hip_solve_t), pointer :: dev => null()
  contains
    procedure :: initialize => flux_initialize
    procedure :: solve => flux_solve
  end type

  ! ------------------------------------------------------------ reconstruction
  type, abstract :: muscl_interpolation_t
    integer :: order = 2
    character(len=32) :: name = ''
    character(len=32) :: limiter_name = ''
  contains
    procedure(interpolate_i), deferred :: interpolate_edge_values
  end type

  abstract interface
    subroutine interpolate_i(self, q, i_edges, j_edges)
      import :: muscl_interpolation_t, rk
      class(muscl_interpolation_t), intent(in) :: self
      real(rk), intent(in) :: q(:, :)
      real(rk), allocatable, intent(out) :: i_edges(:, :, :), j_edges(:, :, :)
    end subroutine
  end interface

  !> L/R edge states of one halo'd primitive field: ((L,R), 0:ni, 0:nj) for i- and j-edges
  type, extends(muscl_interpolation_t) :: hip_muscl_t
    type(hip_solve_t), pointer :: dev => null()
  contains
    procedure :: interpolate_edge_values => muscl_interpolate
  end type

  ! -------------------------------------------------------- boundary conditions
  type, abstract :: boundary_condition_t
    character(len=32) :: name = ''
    character(len=2) :: location = ''
    integer :: priority = 0
  contains
    procedure(bc_apply_i), deferred :: apply
  end type

  abstract interface
    subroutine bc_apply_i(self, rho, u, v, p)
      import :: boundary_condition_t, rk
      class(boundary_condition_t), intent(inout) :: self
      real(rk), intent(inout) :: rho(:, :), u(:, :), v(:, :), p(:, :)
    end subroutine
  end interface

  !> The four sides applied in the reference's priority order (flux_solver.f90:138-185)
  type, extends(boundary_condition_t) :: hip_boundary_conditions_t
    type(hip_solve_t), pointer :: dev => null()
  contains
    procedure :: apply => bc_apply
  end type

  ! --------------------------------------------------------------------- fluid
  !> fluid_t's time-integration surface: the state lives on the device
  type :: hip_fluid_t
    type(hip_solve_t), pointer :: dev => null()
    real(rk) :: time = 0.0_rk
    integer :: iteration = 0
  contains
    procedure :: initialize => fluid_initialize
    procedure :: initialize_conserved => fluid_initialize_conserved
    procedure :: get_timestep => fluid_get_timestep
    procedure :: integrate => fluid_integrate
    procedure :: set_energy_source => fluid_set_energy_source
    procedure :: convergence => fluid_convergence
    procedure :: get_conserved => fluid_get_conserved
    procedure :: get_primitives => fluid_get_primitives
    procedure :: finalize => fluid_finalize
  end type

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= CATO_OK) then
      write(*, '(a)') 'Error: '//what//': '//cato_error_message()
      error stop 1
    end if
  end subroutine

  integer function name_id(name, table, what)
    character(len=*), intent(in) :: name, what
    character(len=*), intent(in) :: table(:)
    integer :: k
    do k = 1, size(table)
      if (trim(name) == trim(table(k))) then
        name_id = k
        return
      end if
    end do
    write(*, '(a)') 'Error: unsupported '//what//" '"//trim(name)//"'"
    error stop 1
  end function

  !> Scheme names of the deck -> cato_config (mod_fluid.f90:223-238, muscl_factory.f90:69-90,
  !> bc_factory.f90:48-64, mod_fluid.f90:424-446, mod_input.f90:320-338)
  function config_from_input(input) result(cfg)
    class(cato_input_t), intent(in) :: input
    type(cato_config_t) :: cfg
    character(len=16), parameter :: fluxes(12) = [character(len=16) :: 'AUSMPW+', 'SLAU', 'SLAU2', 'SD-SLAU', &
                                                   'SD-SLAU2', 'Roe', 'S-Roe', 'SP-Roe', 'AUSM+', 'AUSM+-u', 'HLLC', &
                                                   'M-AUSMPW+']
    integer, parameter :: flux_ids(12) = [0, 1, 2, 3, 4, 5, 5, 5, 6, 6, 7, 8]
    character(len=16), parameter :: limiters(6) = [character(len=16) :: 'minmod', 'superbee', 'van_leer', &
                                                    'vanleer', 'MLP3', 'MLP5']
    integer, parameter :: limiter_ids(6) = [0, 1, 2, 2, 3, 4]
    character(len=16), parameter :: bcs(8) = [character(len=16) :: 'periodic', 'zero_gradient', 'symmetry', &
                                               'mirror', 'reflection', 'outflow', 'outlet', 'pressure_input']
    integer, parameter :: bc_ids(8) = [0, 1, 2, 2, 2, 3, 3, 4]
    character(len=16), parameter :: rks(9) = [character(len=16) :: 'ssp_rk2', 'ssp_rk3', 'ssp_rk33', 'rk(1,3)', &
                                               'rk(1,4)', 'rk(1,5)', 'rk(2,3)', 'rk(2,4)', 'rk(2,5)']
    integer, parameter :: rk_ids(9) = [0, 1, 1, 2, 3, 4, 5, 6, 7]
    if (trim(input%spatial_reconstruction) /= 'MUSCL') then
      write(*, '(a)') "Error: unsupported spatial_reconstruction '"//trim(input%spatial_reconstruction)//"'"
      error stop 1
    end if
    cfg%flux = flux_ids(name_id(input%flux_solver, fluxes, 'flux_solver'))
    cfg%limiter = limiter_ids(name_id(input%limiter, limiters, 'limiter'))
    cfg%integrator = rk_ids(name_id(input%integration_strategy, rks, 'integration_strategy'))
    cfg%precision = CATO_PRECISION_FP64
    cfg%bc_plus_x = bc_ids(name_id(input%plus_x, bcs, 'boundary condition'))
    cfg%bc_minus_x = bc_ids(name_id(input%minus_x, bcs, 'boundary condition'))
    cfg%bc_plus_y = bc_ids(name_id(input%plus_y, bcs, 'boundary condition'))
    cfg%bc_minus_y = bc_ids(name_id(input%minus_y, bcs, 'boundary condition'))
    cfg%n_ghost_layers = merge(3, 2, cfg%limiter == 4)
    cfg%apply_low_mach_fix = merge(1, 0, input%apply_low_mach_fix)
    cfg%gamma = input%polytropic_index
    cfg%cfl = input%cfl
    cfg%tile = -1
    cfg%bc_pressure = input%outflow_ambient_pressure
    if (trim(input%plus_x) == 'pressure_input') then
      ! pressure_input_bc_constructor (pressure_input_bc.f90:81-89): the constants of the constant mode
      cfg%bc_pressure = input%constant_bc_pressure_value
      cfg%bc_density = input%bc_density
    end if
  end function

  ! ------------------------------------------------------------ device solve
  subroutine solve_create(self, cfg, grid)
    class(hip_solve_t), intent(inout) :: self
    type(cato_config_t), intent(in) :: cfg
    type(hip_grid_t), target, intent(in) :: grid
    type(cato_geom_t) :: geom
    if (size(grid%node_x, 1) /= grid%ni + 2 * grid%n_halo + 1 .or. &
        size(grid%node_x, 2) /= grid%nj + 2 * grid%n_halo + 1) then
      write(*, '(a)') 'Error: node arrays do not match the grid size'
      error stop 1
    end if
    self%cfg = cfg
    self%ni = grid%ni
    self%nj = grid%nj
    self%nh = grid%n_halo
    geom%ni = grid%ni
    geom%nj = grid%nj
    geom%node_x = c_loc(grid%node_x)
    geom%node_y = c_loc(grid%node_y)
    call check(cato_ctx_create(self%cfg, geom, self%ctx), 'cato_ctx_create')
  end subroutine

  subroutine solve_destroy(self)
    class(hip_solve_t), intent(inout) :: self
    if (c_associated(self%ctx)) call check(cato_ctx_destroy(self%ctx), 'cato_ctx_destroy')
    self%ctx = c_null_ptr
  end subroutine

  subroutine require_halo(dev, a, what)
    type(hip_solve_t), intent(in) :: dev
    real(rk), intent(in) :: a(:, :)
    character(len=*), intent(in) :: what
    if (size(a, 1) /= dev%ni + 2 * dev%nh .or. size(a, 2) /= dev%nj + 2 * dev%nh) then
      write(*, '(a)') 'Error: '//what//' must carry n_ghost_layers halos on the solve grid'
      error stop 1
    end if
  end subroutine

  ! ------------------------------------------------------------- flux solver
  subroutine flux_initialize(self, input, time)
    class(hip_flux_solver_t), intent(inout) :: self
    class(cato_input_t), intent(in) :: input
    real(rk), intent(in) :: time
    self%name = input%flux_solver
    self%time = time
    if (.not. associated(self%dev)) then
      write(*, '(a)') 'Error: hip_flux_solver_t needs its device solve (dev) before initialize'
      error stop 1
    end if
    self%lbounds = [1, 1]
    self%ubounds = [self%dev%ni, self%dev%nj]
  end subroutine

  subroutine flux_solve(self, dt, grid, rho, u, v, p, d_rho_dt, d_rho_u_dt, d_rho_v_dt, d_rho_E_dt)
    class(hip_flux_solver_t), intent(inout) :: self
    real(rk), intent(in) :: dt
    class(hip_grid_t), intent(in) :: grid
    real(rk), intent(inout) :: rho(:, :), u(:, :), v(:, :), p(:, :)
    real(rk), allocatable, intent(out) :: d_rho_dt(:, :), d_rho_u_dt(:, :), d_rho_v_dt(:, :), d_rho_E_dt(:, :)
    integer :: nh
    if (dt < tiny(1.0_rk)) then  ! slau_solver.fypp:162-166
      write(*, '(a)') 'Error: The timestep dt is < tiny(1.0_rk)'
      error stop 1
    end if
    call require_halo(self%dev, rho, 'rho')
    nh = self%dev%nh
    allocate(d_rho_dt(grid%ilo - nh:grid%ihi + nh, grid%jlo - nh:grid%jhi + nh), source=0.0_rk)
    allocate(d_rho_u_dt, d_rho_v_dt, d_rho_E_dt, mold=d_rho_dt)
    call check(cato_residual(self%dev%ctx, rho, u, v, p, d_rho_dt, d_rho_u_dt, d_rho_v_dt, d_rho_E_dt, &
                             int(size(rho, 1), c_int)), 'flux_solver%solve')
    self%iteration = self%iteration + 1
    self%time = self%time + dt
    self%dt = dt
  end subroutine

  ! ---------------------------------------------------------- reconstruction
  subroutine muscl_interpolate(self, q, i_edges, j_edges)
    class(hip_muscl_t), intent(in) :: self
    real(rk), intent(in) :: q(:, :)
    real(rk), allocatable, intent(out) :: i_edges(:, :, :), j_edges(:, :, :)
    call require_halo(self%dev, q, 'q')
    allocate(i_edges(1:2, 0:self%dev%ni, 0:self%dev%nj), source=0.0_rk)
    allocate(j_edges(1:2, 0:self%dev%ni, 0:self%dev%nj), source=0.0_rk)
    call check(cato_interpolate_edges(self%dev%ctx, q, int(size(q, 1), c_int), i_edges, j_edges), &
               'interpolate_edge_values')
  end subroutine

  ! ----------------------------------------------------- boundary conditions
  subroutine bc_apply(self, rho, u, v, p)
    class(hip_boundary_conditions_t), intent(inout) :: self
    real(rk), intent(inout) :: rho(:, :), u(:, :), v(:, :), p(:, :)
    call require_halo(self%dev, rho, 'rho')
    call check(cato_apply_bc(self%dev%ctx, rho, u, v, p, int(size(rho, 1), c_int)), 'boundary_condition%apply')
  end subroutine

  ! ------------------------------------------------------------------- fluid
  !> fluid_t%initialize_from_hdf5 (mod_fluid.f90:256-381): primitives incl. ghosts from the deck
  subroutine fluid_initialize(self, dev, rho, u, v, p)
    class(hip_fluid_t), intent(inout) :: self
    type(hip_solve_t), target, intent(inout) :: dev
    real(rk), intent(in) :: rho(:, :), u(:, :), v(:, :), p(:, :)
    self%dev => dev
    call require_halo(dev, rho, 'density')
    call check(cato_set_primitive(dev%ctx, rho, u, v, p, int(size(rho, 1), c_int), int(dev%nh, c_int)), &
               'fluid%initialize')
    self%time = 0.0_rk
    self%iteration = 0
  end subroutine

  !> A checkpoint of the conserved state (interior arrays): the continuation of the run that wrote it,
  !> bit for bit, since the primitives are recovered exactly as every stage recovers them
  subroutine fluid_initialize_conserved(self, dev, rho, rho_u, rho_v, rho_E, time, iteration)
    class(hip_fluid_t), intent(inout) :: self
    type(hip_solve_t), target, intent(inout) :: dev
    real(rk), intent(in) :: rho(:, :), rho_u(:, :), rho_v(:, :), rho_E(:, :)
    real(rk), intent(in) :: time
    integer, intent(in) :: iteration
    self%dev => dev
    if (size(rho, 1) /= dev%ni .or. size(rho, 2) /= dev%nj) then
      write(*, '(a)') 'Error: the checkpoint does not match the grid of the deck'
      error stop 1
    end if
    call check(cato_set_conserved(dev%ctx, rho, rho_u, rho_v, rho_E, int(size(rho, 1), c_int), 0_c_int), &
               'fluid%initialize_conserved')
    self%time = time
    self%iteration = iteration
  end subroutine

  !> fluid_t%get_timestep (mod_fluid.f90:518-562)
  real(rk) function fluid_get_timestep(self) result(dt)
    class(hip_fluid_t), intent(inout) :: self
    call check(cato_compute_dt(self%dev%ctx, dt), 'fluid%get_timestep')
  end function

  !> fluid_t%integrate (mod_fluid.f90:411-450): one step of the configured scheme
  !> dt returns the step's dt: smaller than given when subcycling retried it (mod_fluid.f90:411-448)
  subroutine fluid_integrate(self, dt, error_code)
    class(hip_fluid_t), intent(inout) :: self
    real(rk), intent(inout) :: dt
    integer, intent(out) :: error_code
    integer(c_int) :: e
    real(c_double) :: used
    call check(cato_step(self%dev%ctx, real(dt, c_double), e), 'fluid%integrate')
    call check(cato_last_dt(self%dev%ctx, used), 'fluid%integrate')
    dt = used
    error_code = e
    self%time = self%time + dt
    self%iteration = self%iteration + 1
  end subroutine

  !> source_term of fluid_t%integrate (mod_fluid.f90:411-416, added to d(rho E)/dt by
  !> time_derivative, mod_fluid.f90:500-509): held for every following step; absent clears it
  subroutine fluid_set_energy_source(self, q_dot)
    class(hip_fluid_t), intent(inout) :: self
    real(rk), intent(in), target, contiguous, optional :: q_dot(:, :)
    if (present(q_dot)) then
      call require_halo(self%dev, q_dot, 'q_dot')
      call check(cato_set_energy_source(self%dev%ctx, c_loc(q_dot), int(size(q_dot, 1), c_int), &
                                        int(self%dev%nh, c_int)), 'fluid%set_energy_source')
    else
      call check(cato_set_energy_source(self%dev%ctx, c_null_ptr, 0_c_int, 0_c_int), 'fluid%set_energy_source')
    end if
  end subroutine

  !> calculate_convergence of the last step (mod_fluid.f90:1162-1205); needs cato_set_diagnostics
  subroutine fluid_convergence(self, norms)
    class(hip_fluid_t), intent(inout) :: self
    real(rk), intent(out) :: norms(4)
    call check(cato_get_convergence(self%dev%ctx, norms), 'fluid%convergence')
  end subroutine

  subroutine fluid_get_conserved(self, rho, rho_u, rho_v, rho_E)
    class(hip_fluid_t), intent(inout) :: self
    real(rk), intent(inout) :: rho(:, :), rho_u(:, :), rho_v(:, :), rho_E(:, :)
    call check(cato_get_conserved(self%dev%ctx, rho, rho_u, rho_v, rho_E, int(size(rho, 1), c_int), 0_c_int), &
               'fluid%get_conserved')
  end subroutine

  subroutine fluid_get_primitives(self, rho, u, v, p, cs)
    class(hip_fluid_t), intent(inout) :: self
    real(rk), intent(inout) :: rho(:, :), u(:, :), v(:, :), p(:, :), cs(:, :)
    call check(cato_get_primitive(self%dev%ctx, rho, u, v, p, cs, int(size(rho, 1), c_int), 0_c_int), &
               'fluid%get_primitives')
  end subroutine

  subroutine fluid_finalize(self)
    class(hip_fluid_t), intent(inout) :: self
    if (associated(self%dev)) call self%dev%destroy()
    self%dev => null()
  end subroutine

end module cato_hip_plugins
