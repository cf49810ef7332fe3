!> Coarray-free CATO driver on the HIP hot path: reads the reference's
!> input.ini and pycato HDF5 deck, runs the time loop of src/main.f90:117-198
!> through the plugin objects, and writes the final state.
!>
!>   cato_hip_host input.ini [output.h5] [--device-loop] [--dry-run] [--contours DIR]
!>
!> --device-loop runs the same loop on the device (cato_run) instead of one
!> host call per step; the results are identical.  --dry-run reads and checks
!> the deck, prints the configuration and stops before touching a device.
!> --contours DIR writes DIR/step_NNNNNNN.h5 at the contour cadence of
!> main.f90 (iteration 0, every contour_interval_dt, and the last step), as the
!> reference's contour writer does (host loop only).  residual_hist.csv is
!> written in the working directory as fluid_t does (mod_fluid.f90:244-246,
!> 1207-1223), one row per step; --no-residual-hist skips it.
program cato_hip_host
  use, intrinsic :: iso_c_binding
  use cato_hip_api
  use cato_deck
  use cato_hip_plugins
  implicit none

  type(cato_input_t) :: input
  type(hip_grid_t), target :: grid
  type(hip_solve_t), target :: dev
  type(hip_fluid_t) :: fluid
  type(cato_config_t) :: cfg
  real(rk), allocatable :: x(:, :), y(:, :), rho(:, :), u(:, :), v(:, :), p(:, :)
  real(rk), allocatable :: r(:, :), ru(:, :), rv(:, :), rE(:, :), cs(:, :)
  real(rk), allocatable :: pi_t(:), pi_p(:), pi_r(:)
  character(len=512) :: ini_file, out_file, arg
  logical :: device_loop, dry_run, residual_hist, constant_dt, from_checkpoint
  real(rk) :: restart_time
  integer :: restart_iteration
  integer :: hist_unit
  character(len=512) :: contour_dir
  type(hip_boundary_conditions_t) :: bcs
  integer :: last_io_iteration
  real(rk) :: last_dt
  integer :: nh, nargs, k, error_code, iteration
  real(rk) :: time, new_time, delta_t, next_output_time, contour_interval_dt, master_time
  integer(c_long_long) :: iters
  integer(c_int) :: e

  nargs = command_argument_count()
  if (nargs < 1) then
    write(*, '(a)') 'usage: cato_hip_host input.ini [output.h5] [--device-loop]'
    error stop 1
  end if
  call get_command_argument(1, ini_file)
  out_file = 'final_state.h5'
  device_loop = .false.
  dry_run = .false.
  residual_hist = .true.
  contour_dir = ''
  k = 2
  do while (k <= nargs)
    call get_command_argument(k, arg)
    if (trim(arg) == '--device-loop') then
      device_loop = .true.
    else if (trim(arg) == '--dry-run') then
      dry_run = .true.
    else if (trim(arg) == '--no-residual-hist') then
      residual_hist = .false.
    else if (trim(arg) == '--contours') then
      k = k + 1
      call get_command_argument(k, contour_dir)
    else
      out_file = arg
    end if
    k = k + 1
  end do

  call read_input_ini(trim(ini_file), input)
  if (.not. input%read_from_file) then
    write(*, '(a)') 'Error: only initial conditions read from file are supported'
    error stop 1
  end if
  if (input%non_dimensionalize .and. (input%reference_length /= 1.0_rk .or. input%reference_density /= 1.0_rk &
                                      .or. input%reference_velocity /= 1.0_rk)) then
    write(*, '(a)') 'Error: reference_state scaling other than 1 is not supported by this host'
    error stop 1
  end if
  if (input%enable_source_terms) then
    write(*, '(a)') 'Error: [source_terms] from the deck are not supported; the library takes the step''s '// &
      'source field through fluid%set_energy_source'
    error stop 1
  end if
  cfg = config_from_input(input)

  ! grid_block_2d_t%initialize from the deck (grid_block_2d.f90:194-265)
  call read_initial_conditions(trim(input%initial_condition_file), nh, x, y, rho, u, v, p)
  if (nh /= cfg%n_ghost_layers) then
    write(*, '(a, i0, a, i0)') 'Error: deck has n_ghost_layers = ', nh, ', the scheme needs ', cfg%n_ghost_layers
    error stop 1
  end if
  grid%n_halo = nh
  grid%ni = size(rho, 1) - 2 * nh
  grid%nj = size(rho, 2) - 2 * nh
  grid%ihi = grid%ni
  grid%jhi = grid%nj
  allocate(grid%node_x(1 - nh:grid%ni + nh + 1, 1 - nh:grid%nj + nh + 1), source=x)
  allocate(grid%node_y(1 - nh:grid%ni + nh + 1, 1 - nh:grid%nj + nh + 1), source=y)

  if (dry_run) then
    write(*, '(a, i0, a, i0, a, i0)') 'deck: ni = ', grid%ni, ', nj = ', grid%nj, ', n_ghost_layers = ', nh
    write(*, '(a, 8(1x, i0))') 'config: flux limiter integrator bcs(+x -x +y -y) n_ghost =', cfg%flux, cfg%limiter, &
      cfg%integrator, cfg%bc_plus_x, cfg%bc_minus_x, cfg%bc_plus_y, cfg%bc_minus_y, cfg%n_ghost_layers
    write(*, '(a, 3(1x, es23.16))') 'gamma cfl max_time =', cfg%gamma, cfg%cfl, input%max_time
    write(*, '(a, 1x, es23.16)') 'contour_interval_dt =', input%contour_interval_dt
    write(*, '(a, 2(1x, es23.16))') 'bc_pressure bc_density =', cfg%bc_pressure, cfg%bc_density
    write(*, '(a, 2(1x, es23.16))') 'x(1-nh,1-nh) y(ni+nh+1,nj+nh+1) =', grid%node_x(1 - nh, 1 - nh), &
      grid%node_y(grid%ni + nh + 1, grid%nj + nh + 1)
    write(*, '(a, 2(1x, es23.16))') 'density(1,1) pressure(ni,nj) =', rho(1 + nh, 1 + nh), p(grid%ni + nh, grid%nj + nh)
    stop
  end if
  call dev%create(cfg, grid)
  if (input%allow_subcycle) then
    ! fluid_t%initialize (mod_fluid.f90:171-172); ssp_rk_2_2 / ssp_rk_3_3 retry a step with dt / 10
    if (device_loop .and. (cfg%integrator == 0 .or. cfg%integrator == 1)) then
      write(*, '(a)') 'Error: --device-loop does not support allow_subcycle (its retries are host-driven)'
      error stop 1
    end if
    call check_rc(cato_set_subcycle(dev%ctx, 1_c_int, input%subcycle_error_tolerance), 'cato_set_subcycle')
  end if
  if (trim(input%plus_x) == 'pressure_input' .and. .not. input%apply_constant_bc_pressure) then
    ! pressure_input_bc_constructor: the time-dependent input from its file
    call read_pressure_input(trim(input%bc_pressure_input_file), input%bc_pressure_scale_factor, pi_t, pi_p, pi_r)
    call check_rc(cato_set_pressure_input(dev%ctx, pi_t, pi_p, pi_r, int(size(pi_t), c_int)), &
                  'pressure input '//trim(input%bc_pressure_input_file))
  end if
  restart_time = 0.0_rk
  restart_iteration = 0
  if (input%restart_from_file) then
    ! main.f90:100-104 + master_puppeteer.f90:106-121: state, time and iteration from the restart file
    call read_restart(trim(input%restart_file), restart_time, restart_iteration, from_checkpoint, rho, u, v, p)
    if (from_checkpoint) then
      call fluid%initialize_conserved(dev, rho, u, v, p, restart_time, restart_iteration)
    else
      if (size(rho, 1) /= grid%ni + 2 * nh .or. size(rho, 2) /= grid%nj + 2 * nh) then
        write(*, '(a)') 'Error: the restart file does not match the grid of the deck'
        error stop 1
      end if
      call fluid%initialize(dev, rho, u, v, p)
      fluid%time = restart_time
      fluid%iteration = restart_iteration
    end if
    ! the fluid (and its flux solver) start from the restart time (master_puppeteer.f90:106-132)
    call check_rc(cato_set_time(dev%ctx, restart_time), 'cato_set_time')
    write(*, '(a, es22.15, a, i0)') 'Restart from '//trim(input%restart_file)//': time = ', restart_time, &
      ', iteration ', restart_iteration
  else
    call fluid%initialize(dev, rho, u, v, p)
  end if
  bcs%dev => dev
  bcs%name = 'deck boundary conditions'
  if (len_trim(contour_dir) > 0) then
    select case (trim(input%contour_format))
    case ('xdmf', 'hdf5', 'h5')
    case default
      write(*, '(a)') "Error: Unsupported I/O contour format '"//trim(input%contour_format)// &
        "' (must be .xdmf .h5 or .hdf5)"
      error stop 1
    end select
    if (device_loop) then
      write(*, '(a)') 'Error: --contours needs the host-side loop (drop --device-loop)'
      error stop 1
    end if
    call execute_command_line('mkdir -p '//trim(contour_dir))
  end if
  write(*, '(a, i0, a, i0, a)') 'CATO (HIP hot path): ', grid%ni, ' x ', grid%nj, ' cells, '// &
    trim(input%flux_solver)//' / '//trim(input%limiter)//' / '//trim(input%integration_strategy)

  contour_interval_dt = input%contour_interval_dt
  time = 0.0_rk
  iteration = 0
  if (residual_hist) then
    call check_rc(cato_set_diagnostics(dev%ctx, 1_c_int), 'cato_set_diagnostics')
    open(newunit=hist_unit, file='residual_hist.csv', status='replace')
    write(hist_unit, '(a)') 'iteration,time,dt,rho,rho_u,rho_v,rho_E'
  end if
  if (device_loop) then
    ! the device loop starts at the fluid time (the restart time) and counts its own iterations;
    ! main.f90:140-158: its startup period or given first dt, on the driver's iteration count
    call check_rc(cato_set_loop_dt(dev%ctx, merge(input%initial_delta_t, 0.0_rk, &
                                                  .not. input%use_constant_delta_t), &
                                   merge(1_c_int, 0_c_int, input%enable_startup_time), input%startup_period, &
                                   input%startup_delta_t, int(restart_iteration, c_long_long), 0_c_int), &
                  'cato_set_loop_dt')
    if (input%max_iterations - restart_iteration > 0) then
      call check_rc(cato_run(dev%ctx, input%max_time, int(input%max_iterations - restart_iteration, c_long_long), &
                             contour_interval_dt, merge(input%initial_delta_t, 0.0_rk, input%use_constant_delta_t), &
                             iters, time, e), 'cato_run')
    else
      ! main.f90:138: a restart already at max_iterations takes no step (cato_run reads a count <= 0 as
      ! no limit, so it is not called)
      iters = 0
      time = restart_time
      e = 0
    end if
    iteration = restart_iteration + int(iters)
    error_code = e
    if (residual_hist) call drain_history()
  else
    ! main.f90:100-198: contour at iteration 0, then the loop with the contour-interval clip of the driver time
    ! main.f90:100-107: a restart resumes the driver at the file's time and iteration, with no contour
    time = restart_time
    iteration = restart_iteration
    next_output_time = time + contour_interval_dt
    master_time = restart_time
    error_code = ALL_OK
    last_dt = 0.0_rk
    last_io_iteration = iteration
    if (len_trim(contour_dir) > 0 .and. .not. input%restart_from_file) call contour(0, master_time, last_dt)
    do while (time < input%max_time .and. iteration < input%max_iterations)
      ! main.f90:140-158: a startup period of constant dt, else a constant or given first dt, else CFL
      constant_dt = .false.
      if (input%enable_startup_time) then
        if (time <= input%startup_period) then
          constant_dt = .true.
          delta_t = input%startup_delta_t
        end if
      else if (input%use_constant_delta_t .or. (iteration == 0 .and. input%initial_delta_t > 0.0_rk)) then
        constant_dt = .true.
        delta_t = input%initial_delta_t
      end if
      if (.not. constant_dt) delta_t = fluid%get_timestep()
      call fluid%integrate(delta_t, error_code)
      if (residual_hist) call history_row(fluid%iteration, fluid%time, delta_t)
      master_time = master_time + delta_t  ! master_puppeteer%time
      last_dt = delta_t
      new_time = master_time
      if (error_code /= ALL_OK) then
        ! main.f90:170-175: the failing state is written to a contour before stopping
        if (len_trim(contour_dir) > 0) call contour(iteration, master_time, last_dt)
        last_io_iteration = iteration
        exit
      end if
      if (abs(time - next_output_time) < epsilon(1.0_rk)) then
        next_output_time = next_output_time + contour_interval_dt
        if (len_trim(contour_dir) > 0) call contour(iteration, master_time, last_dt)
        last_io_iteration = iteration
      end if
      if (new_time > next_output_time) then
        delta_t = abs(next_output_time - time)
        new_time = time + delta_t
      end if
      time = new_time
      iteration = iteration + 1
    end do
    if (len_trim(contour_dir) > 0 .and. iteration > last_io_iteration) call contour(iteration, master_time, last_dt)
  end if
  if (residual_hist) close(hist_unit)
  if (error_code /= ALL_OK) write(*, '(a, i0)') 'Something went wrong in the time integration, error code ', error_code

  allocate(r(grid%ni, grid%nj), ru(grid%ni, grid%nj), rv(grid%ni, grid%nj), rE(grid%ni, grid%nj))
  allocate(cs(grid%ni, grid%nj))
  deallocate(rho, u, v, p)
  allocate(rho(grid%ni, grid%nj), u(grid%ni, grid%nj), v(grid%ni, grid%nj), p(grid%ni, grid%nj))
  call fluid%get_conserved(r, ru, rv, rE)
  call fluid%get_primitives(rho, u, v, p, cs)
  call write_state(trim(out_file), time, iteration, r, ru, rv, rE, u, v, p)
  write(*, '(a, es22.15, a, i0, a)') 'Done: time = ', time, ', ', iteration, ' iterations -> '//trim(out_file)
  call fluid%finalize()
  if (error_code /= ALL_OK) error stop 1

contains

  !> contour_writer_t%write_contour: primitives of the current state with BC-filled ghost cells
  subroutine contour(it, t, dt_last)
    integer, intent(in) :: it
    real(rk), intent(in) :: t, dt_last
    real(rk), allocatable :: hr(:, :), hu(:, :), hv(:, :), hp(:, :), hc(:, :)
    real(rk), allocatable :: ir(:, :), iu(:, :), iv(:, :), ip(:, :), ic(:, :)
    integer :: n1, n2
    n1 = grid%ni + 2 * nh
    n2 = grid%nj + 2 * nh
    allocate(ir(grid%ni, grid%nj), iu(grid%ni, grid%nj), iv(grid%ni, grid%nj), ip(grid%ni, grid%nj))
    allocate(ic(grid%ni, grid%nj))
    call fluid%get_primitives(ir, iu, iv, ip, ic)
    allocate(hr(n1, n2), hu(n1, n2), hv(n1, n2), hp(n1, n2), hc(n1, n2), source=0.0_rk)
    hr(nh + 1:nh + grid%ni, nh + 1:nh + grid%nj) = ir
    hu(nh + 1:nh + grid%ni, nh + 1:nh + grid%nj) = iu
    hv(nh + 1:nh + grid%ni, nh + 1:nh + grid%nj) = iv
    hp(nh + 1:nh + grid%ni, nh + 1:nh + grid%nj) = ip
    call bcs%apply(hr, hu, hv, hp)
    where (hr > 0.0_rk .and. hp >= 0.0_rk) hc = sqrt(cfg%gamma * hp / hr)
    hc(nh + 1:nh + grid%ni, nh + 1:nh + grid%nj) = ic
    if (input%plot_volume) then
      call write_contour(trim(contour_dir), trim(input%title), it, t, dt_last, nh, x, y, hr, hu, hv, hp, hc, &
                         cell_volumes(), unit_system=input%unit_system, plot_64bit=input%plot_64bit)
    else
      call write_contour(trim(contour_dir), trim(input%title), it, t, dt_last, nh, x, y, hr, hu, hv, hp, hc, &
                         unit_system=input%unit_system, plot_64bit=input%plot_64bit)
    end if
    ! contour_writer_t%write_contour (contour_writer.f90:178-190): 'xdmf' adds the companion file
    if (trim(input%contour_format) == 'xdmf') &
      call write_xdmf(trim(contour_dir), it, t, n1, n2, input%unit_system, input%plot_volume)
  end subroutine

  !> quad_cell_t volumes of the halo'd cells (quad_cell.f90:78-91 round-off fixes, 160-175 the volume)
  function cell_volumes() result(vol)
    real(rk), allocatable :: vol(:, :)
    real(rk) :: qx(4), qy(4)
    integer :: i, j
    allocate(vol(size(x, 1) - 1, size(x, 2) - 1))
    do j = 1, size(vol, 2)
      do i = 1, size(vol, 1)
        qx = [x(i, j), x(i + 1, j), x(i + 1, j + 1), x(i, j + 1)]
        qy = [y(i, j), y(i + 1, j), y(i + 1, j + 1), y(i, j + 1)]
        if (abs(qy(2) - qy(1)) < epsilon(1.0_rk)) qy(2) = qy(1)
        if (abs(qy(4) - qy(3)) < epsilon(1.0_rk)) qy(4) = qy(3)
        if (abs(qx(3) - qx(2)) < epsilon(1.0_rk)) qx(3) = qx(2)
        if (abs(qx(4) - qx(1)) < epsilon(1.0_rk)) qx(4) = qx(1)
        vol(i, j) = (qx(2) - qx(1)) * (qy(4) - qy(1)) - (qx(4) - qx(1)) * (qy(2) - qy(1))
      end do
    end do
  end function

  !> write_residual_history (mod_fluid.f90:1207-1223) for the step just taken
  subroutine history_row(it, t, dt)
    integer, intent(in) :: it
    real(rk), intent(in) :: t, dt
    real(rk) :: norms(4)
    call fluid%convergence(norms)
    write(hist_unit, '(i0, ",", 6(es16.6, ","))') it, t, dt, norms
  end subroutine

  !> the rows the device-side loop collected, oldest first
  subroutine drain_history()
    real(c_double) :: rows(7, 256)
    integer(c_long_long) :: n
    integer :: r
    do
      call check_rc(cato_get_history(dev%ctx, rows, 256_c_long_long, n), 'cato_get_history')
      do r = 1, int(n)
        write(hist_unit, '(i0, ",", 6(es16.6, ","))') restart_iteration + nint(rows(1, r)), rows(2:7, r)
      end do
      if (n < 256) exit
    end do
  end subroutine

  subroutine check_rc(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= CATO_OK) then
      write(*, '(a)') 'Error: '//what//': '//cato_error_message()
      error stop 1
    end if
  end subroutine

end program cato_hip_host
