!> ISO_C_BINDING interface of libcato_hip.so (include/cato_hip.h).
!>
!> Every entry point returns an integer status (CATO_OK = 0); the message of
!> the last failure is available through cato_error_message().  Arrays are
!> passed as Fortran (column-major, i fastest) arrays plus their leading
!> dimension and the number of ghost layers they carry, exactly as the
!> reference stores field_2d_t%data(ilo-nh:ihi+nh, jlo-nh:jhi+nh)
!> (src/lib/field/field_2d.f90:477-483).
module cato_hip_api
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: CATO_OK = 0
  integer(c_int), parameter, public :: CATO_ERR_INVALID = -1, CATO_ERR_UNSUPPORTED = -2, &
                                       CATO_ERR_HIP = -3, CATO_ERR_NO_DEVICE = -4
  integer(c_int), parameter, public :: CATO_PRECISION_FP64 = 0, CATO_PRECISION_FP32_FLUX = 1

  !> cato_config (include/cato_hip.h)
  type, bind(C), public :: cato_config_t
    integer(c_int) :: flux = 0
    integer(c_int) :: limiter = 0
    integer(c_int) :: integrator = 0
    integer(c_int) :: precision = 0
    integer(c_int) :: bc_plus_x = 0, bc_minus_x = 0, bc_plus_y = 0, bc_minus_y = 0
    integer(c_int) :: n_ghost_layers = 2
    integer(c_int) :: apply_low_mach_fix = 1
    real(c_double) :: gamma = 1.4_c_double
    real(c_double) :: cfl = 0.1_c_double
    integer(c_int) :: tile = -1
    real(c_double) :: bc_pressure = 0.0_c_double
    real(c_double) :: bc_density = 0.0_c_double
  end type

  !> cato_geom: interior cells and the node coordinates incl. ghost nodes
  type, bind(C), public :: cato_geom_t
    integer(c_int) :: ni = 0, nj = 0
    type(c_ptr) :: node_x = c_null_ptr
    type(c_ptr) :: node_y = c_null_ptr
  end type

  public :: cato_abi_version, cato_ctx_create, cato_ctx_destroy, cato_set_primitive, cato_set_conserved, &
            cato_get_conserved, cato_get_primitive, cato_set_energy_source, cato_set_diagnostics, &
            cato_get_convergence, cato_get_history, cato_compute_dt, cato_step, cato_run, cato_apply_bc, &
            cato_interpolate_edges, cato_edge_fluxes, cato_residual, cato_error_message, &
            cato_set_pressure_input, cato_set_time, cato_set_subcycle, cato_last_dt, &
            cato_set_loop_dt

  interface
    function cato_abi_version() bind(C, name="cato_abi_version") result(v)
      import :: c_int
      integer(c_int) :: v
    end function

    function cato_last_error_c() bind(C, name="cato_last_error") result(p)
      import :: c_ptr
      type(c_ptr) :: p
    end function

    function cato_ctx_create(cfg, geom, ctx) bind(C, name="cato_ctx_create") result(rc)
      import :: c_int, c_ptr, cato_config_t, cato_geom_t
      type(cato_config_t), intent(in) :: cfg
      type(cato_geom_t), intent(in) :: geom
      type(c_ptr), intent(out) :: ctx
      integer(c_int) :: rc
    end function

    function cato_ctx_destroy(ctx) bind(C, name="cato_ctx_destroy") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int) :: rc
    end function

    function cato_set_primitive(ctx, rho, u, v, p, ld, halo) bind(C, name="cato_set_primitive") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rho(*), u(*), v(*), p(*)
      integer(c_int), value :: ld, halo
      integer(c_int) :: rc
    end function

    function cato_set_conserved(ctx, rho, rho_u, rho_v, rho_E, ld, halo) bind(C, name="cato_set_conserved") &
      result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: rho(*), rho_u(*), rho_v(*), rho_E(*)
      integer(c_int), value :: ld, halo
      integer(c_int) :: rc
    end function

    function cato_set_diagnostics(ctx, enable) bind(C, name="cato_set_diagnostics") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      integer(c_int) :: rc
    end function

    function cato_get_convergence(ctx, norms) bind(C, name="cato_get_convergence") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: norms(4)
      integer(c_int) :: rc
    end function

    !> rows(7, max_rows): iteration, time, dt, norms(4) per step, oldest first
    function cato_get_history(ctx, rows, max_rows, n_rows) bind(C, name="cato_get_history") result(rc)
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: rows(*)
      integer(c_long_long), value :: max_rows
      integer(c_long_long), intent(out) :: n_rows
      integer(c_int) :: rc
    end function

    !> q_dot: c_loc of a contiguous (ld, *) field, or c_null_ptr to clear the source
    function cato_set_energy_source(ctx, q_dot, ld, halo) bind(C, name="cato_set_energy_source") result(rc)
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, q_dot
      integer(c_int), value :: ld, halo
      integer(c_int) :: rc
    end function

    !> pressure_input table rows (time, pressure, density), scaled; n = 0 returns to the constants
    function cato_set_pressure_input(ctx, time, pressure, density, n) bind(C, name="cato_set_pressure_input") &
      result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: time(*), pressure(*), density(*)
      integer(c_int), value :: n
      integer(c_int) :: rc
    end function

    !> [time] allow_subcycle / subcycle_error_tolerance (host-driven steps)
    function cato_set_subcycle(ctx, enable, tolerance) bind(C, name="cato_set_subcycle") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
      real(c_double), value :: tolerance
      integer(c_int) :: rc
    end function

    !> startup period / given first dt of cato_run (main.f90:140-158)
    function cato_set_loop_dt(ctx, initial_dt, enable_startup, startup_period, startup_dt, start_iteration, &
                              exact_landing) &
        bind(C, name="cato_set_loop_dt") result(rc)
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), value :: initial_dt
      integer(c_int), value :: enable_startup
      real(c_double), value :: startup_period, startup_dt
      integer(c_long_long), value :: start_iteration
      integer(c_int), value :: exact_landing
      integer(c_int) :: rc
    end function

    !> the dt the last cato_step used
    function cato_last_dt(ctx, dt) bind(C, name="cato_last_dt") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
      integer(c_int) :: rc
    end function

    !> fluid time (and the time the solve starts from)
    function cato_set_time(ctx, time) bind(C, name="cato_set_time") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: time
      integer(c_int) :: rc
    end function

    function cato_get_conserved(ctx, rho, rho_u, rho_v, rho_E, ld, halo) bind(C, name="cato_get_conserved") &
      result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: rho(*), rho_u(*), rho_v(*), rho_E(*)
      integer(c_int), value :: ld, halo
      integer(c_int) :: rc
    end function

    function cato_get_primitive(ctx, rho, u, v, p, cs, ld, halo) bind(C, name="cato_get_primitive") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: rho(*), u(*), v(*), p(*), cs(*)
      integer(c_int), value :: ld, halo
      integer(c_int) :: rc
    end function

    function cato_compute_dt(ctx, dt) bind(C, name="cato_compute_dt") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
      integer(c_int) :: rc
    end function

    function cato_step(ctx, dt, error_code) bind(C, name="cato_step") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int), intent(out) :: error_code
      integer(c_int) :: rc
    end function

    function cato_run(ctx, max_time, max_iterations, contour_interval_dt, constant_dt, iterations, time, &
                      error_code) bind(C, name="cato_run") result(rc)
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: ctx
      real(c_double), value :: max_time
      integer(c_long_long), value :: max_iterations
      real(c_double), value :: contour_interval_dt, constant_dt
      integer(c_long_long), intent(out) :: iterations
      real(c_double), intent(out) :: time
      integer(c_int), intent(out) :: error_code
      integer(c_int) :: rc
    end function

    function cato_apply_bc(ctx, rho, u, v, p, ld) bind(C, name="cato_apply_bc") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: rho(*), u(*), v(*), p(*)
      integer(c_int), value :: ld
      integer(c_int) :: rc
    end function

    function cato_interpolate_edges(ctx, q, ld, i_edges, j_edges) bind(C, name="cato_interpolate_edges") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: q(*)
      integer(c_int), value :: ld
      real(c_double), intent(inout) :: i_edges(*), j_edges(*)
      integer(c_int) :: rc
    end function

    function cato_edge_fluxes(ctx, edges, p, ld, iflux, jflux) bind(C, name="cato_edge_fluxes") result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: edges(*), p(*)
      integer(c_int), value :: ld
      real(c_double), intent(inout) :: iflux(*), jflux(*)
      integer(c_int) :: rc
    end function

    function cato_residual(ctx, rho, u, v, p, d_rho, d_rho_u, d_rho_v, d_rho_E, ld) bind(C, name="cato_residual") &
      result(rc)
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(inout) :: rho(*), u(*), v(*), p(*)
      real(c_double), intent(inout) :: d_rho(*), d_rho_u(*), d_rho_v(*), d_rho_E(*)
      integer(c_int), value :: ld
      integer(c_int) :: rc
    end function

    function c_strlen(s) bind(C, name="strlen") result(n)
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: n
    end function
  end interface

contains

  !> Message of the calling thread's last failed call (cato_last_error).
  function cato_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, k
    p = cato_last_error_c()
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
      msg(k:k) = chars(k)
    end do
  end function

end module cato_hip_api
