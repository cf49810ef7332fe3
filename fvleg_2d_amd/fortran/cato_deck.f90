!> The pycato input deck of the reference: the input.ini entries the hot path
!> needs (src/lib/io/mod_input.f90:200-410) and the HDF5 initial-condition
!> file (pycato/generate_initial_grids.py:603-672, read by
!> src/lib/grid/grid_block_2d.f90:194-265 and src/lib/fluid/mod_fluid.f90:256-381).
!>
!> HDF5 is reached through its C library (ISO_C_BINDING), so no Fortran HDF5
!> modules are needed.  Datasets are stored transposed (HDF5 dims (nj, ni)),
!> which is exactly a Fortran (ni, nj) array in memory.
module cato_deck
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer, parameter :: rk = c_double
  public :: cato_input_t, read_input_ini, read_initial_conditions, read_restart, write_state, write_contour, &
            read_pressure_input, write_xdmf

  !> Deck entries of the hot path (names as in input.ini)
  type :: cato_input_t
    character(len=256) :: title = ''
    ! [time]
    real(rk) :: max_time = 1.0_rk
    real(rk) :: initial_delta_t = 0.0_rk
    logical :: use_constant_delta_t = .false.
    real(rk) :: cfl = 0.1_rk
    character(len=32) :: integration_strategy = ''
    integer :: max_iterations = huge(1)
    logical :: enable_startup_time = .false.
    real(rk) :: startup_period = 0.0_rk, startup_delta_t = 0.0_rk
    logical :: allow_subcycle = .false.
    real(rk) :: subcycle_error_tolerance = 1e-3_rk
    ! [source_terms]
    logical :: enable_source_terms = .false.
    ! [restart]
    logical :: restart_from_file = .false.
    character(len=256) :: restart_file = ''
    ! [reference_state]
    logical :: non_dimensionalize = .true.
    real(rk) :: reference_length = 1.0_rk, reference_density = 1.0_rk, reference_velocity = 1.0_rk
    ! [physics]
    real(rk) :: polytropic_index = 5.0_rk / 3.0_rk
    ! [scheme]
    character(len=32) :: flux_solver = 'AUSMPW+'
    character(len=32) :: spatial_reconstruction = 'MUSCL'
    character(len=32) :: limiter = 'minmod'
    logical :: apply_low_mach_fix = .true.
    ! [initial_conditions]
    logical :: read_from_file = .true.
    character(len=256) :: initial_condition_file = ''
    ! [boundary_conditions]
    character(len=32) :: plus_x = 'periodic', minus_x = 'periodic', plus_y = 'periodic', minus_y = 'periodic'
    real(rk) :: outflow_ambient_pressure = 0.0_rk
    ! pressure_input (mod_input.f90:366-382)
    logical :: apply_constant_bc_pressure = .false.
    real(rk) :: constant_bc_pressure_value = 0.0_rk
    real(rk) :: bc_density = 0.0_rk
    character(len=256) :: bc_pressure_input_file = ''
    real(rk) :: bc_pressure_scale_factor = 1.0_rk
    ! [io]
    real(rk) :: contour_interval_dt = 0.1_rk
    logical :: plot_64bit = .true., plot_volume = .true.   ! mod_input.f90:432,443
    character(len=8) :: contour_format = 'xdmf'            ! [io] format (mod_input.f90:429-430)
    ! [general] units: the contour writer's unit system (units.f90:94-143)
    character(len=8) :: unit_system = 'cgs'
    ! [grid]
    integer :: n_ghost_layers = 0
  end type

  integer(c_int), parameter :: H5F_ACC_RDONLY = 0, H5F_ACC_TRUNC = 2
  integer(c_int64_t), parameter :: H5P_DEFAULT = 0, H5S_ALL = 0
  integer(c_int64_t), bind(C, name="H5T_NATIVE_DOUBLE_g") :: h5t_native_double
  integer(c_int64_t), bind(C, name="H5T_NATIVE_INT_g") :: h5t_native_int
  integer(c_int64_t), bind(C, name="H5T_C_S1_g") :: h5t_c_s1
  integer(c_int64_t), bind(C, name="H5T_NATIVE_FLOAT_g") :: h5t_native_float
  integer(c_int), parameter :: H5S_SCALAR = 0

  interface
    function H5open() bind(C, name="H5open") result(r)
      import :: c_int
      integer(c_int) :: r
    end function
    function H5Fopen(name, flags, fapl) bind(C, name="H5Fopen") result(h)
      import :: c_char, c_int, c_int64_t
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), value :: flags
      integer(c_int64_t), value :: fapl
      integer(c_int64_t) :: h
    end function
    function H5Fcreate(name, flags, fcpl, fapl) bind(C, name="H5Fcreate") result(h)
      import :: c_char, c_int, c_int64_t
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), value :: flags
      integer(c_int64_t), value :: fcpl, fapl
      integer(c_int64_t) :: h
    end function
    function H5Fclose(h) bind(C, name="H5Fclose") result(r)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h
      integer(c_int) :: r
    end function
    function H5Dopen2(loc, name, dapl) bind(C, name="H5Dopen2") result(h)
      import :: c_char, c_int64_t
      integer(c_int64_t), value :: loc, dapl
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t) :: h
    end function
    function H5Dclose(h) bind(C, name="H5Dclose") result(r)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: h
      integer(c_int) :: r
    end function
    function H5Dget_space(d) bind(C, name="H5Dget_space") result(h)
      import :: c_int64_t
      integer(c_int64_t), value :: d
      integer(c_int64_t) :: h
    end function
    function H5Sget_simple_extent_ndims(s) bind(C, name="H5Sget_simple_extent_ndims") result(n)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: s
      integer(c_int) :: n
    end function
    function H5Sget_simple_extent_dims(s, dims, maxdims) bind(C, name="H5Sget_simple_extent_dims") result(n)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: s
      integer(c_int64_t), intent(out) :: dims(*)
      type(c_ptr), value :: maxdims
      integer(c_int) :: n
    end function
    function H5Sclose(s) bind(C, name="H5Sclose") result(r)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: s
      integer(c_int) :: r
    end function
    function H5Screate_simple(rank, dims, maxdims) bind(C, name="H5Screate_simple") result(h)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int), value :: rank
      integer(c_int64_t), intent(in) :: dims(*)
      type(c_ptr), value :: maxdims
      integer(c_int64_t) :: h
    end function
    function H5Dread(d, mtype, mspace, fspace, plist, buf) bind(C, name="H5Dread") result(r)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: d, mtype, mspace, fspace, plist
      type(c_ptr), value :: buf
      integer(c_int) :: r
    end function
    function H5Dcreate2(loc, name, dtype, space, lcpl, dcpl, dapl) bind(C, name="H5Dcreate2") result(h)
      import :: c_char, c_int64_t
      integer(c_int64_t), value :: loc, dtype, space, lcpl, dcpl, dapl
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int64_t) :: h
    end function
    function H5Screate(kind) bind(C, name="H5Screate") result(h)
      import :: c_int, c_int64_t
      integer(c_int), value :: kind
      integer(c_int64_t) :: h
    end function
    function H5Tcopy(t) bind(C, name="H5Tcopy") result(h)
      import :: c_int64_t
      integer(c_int64_t), value :: t
      integer(c_int64_t) :: h
    end function
    function H5Tset_size(t, n) bind(C, name="H5Tset_size") result(r)
      import :: c_int, c_int64_t, c_size_t
      integer(c_int64_t), value :: t
      integer(c_size_t), value :: n
      integer(c_int) :: r
    end function
    function H5Tclose(t) bind(C, name="H5Tclose") result(r)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: t
      integer(c_int) :: r
    end function
    function H5Dwrite(d, mtype, mspace, fspace, plist, buf) bind(C, name="H5Dwrite") result(r)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: d, mtype, mspace, fspace, plist
      type(c_ptr), value :: buf
      integer(c_int) :: r
    end function
    function H5Lexists(loc, name, lapl) bind(C, name="H5Lexists") result(r)
      import :: c_int, c_int64_t, c_char
      integer(c_int64_t), value :: loc, lapl
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int) :: r
    end function
    function H5Aexists_by_name(loc, obj, attr, lapl) bind(C, name="H5Aexists_by_name") result(r)
      import :: c_int, c_int64_t, c_char
      integer(c_int64_t), value :: loc, lapl
      character(kind=c_char), intent(in) :: obj(*), attr(*)
      integer(c_int) :: r
    end function
    function H5Aopen_by_name(loc, obj, attr, aapl, lapl) bind(C, name="H5Aopen_by_name") result(h)
      import :: c_int64_t, c_char
      integer(c_int64_t), value :: loc, aapl, lapl
      character(kind=c_char), intent(in) :: obj(*), attr(*)
      integer(c_int64_t) :: h
    end function
    function H5Aget_type(a) bind(C, name="H5Aget_type") result(h)
      import :: c_int64_t
      integer(c_int64_t), value :: a
      integer(c_int64_t) :: h
    end function
    function H5Tget_size(t) bind(C, name="H5Tget_size") result(n)
      import :: c_int64_t, c_size_t
      integer(c_int64_t), value :: t
      integer(c_size_t) :: n
    end function
    function H5Tis_variable_str(t) bind(C, name="H5Tis_variable_str") result(r)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: t
      integer(c_int) :: r
    end function
    function H5Aread(a, mtype, buf) bind(C, name="H5Aread") result(r)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: a, mtype
      type(c_ptr), value :: buf
      integer(c_int) :: r
    end function
    function H5Aclose(a) bind(C, name="H5Aclose") result(r)
      import :: c_int, c_int64_t
      integer(c_int64_t), value :: a
      integer(c_int) :: r
    end function
    function H5Acreate_by_name(loc, obj, attr, atype, space, acpl, aapl, lapl) bind(C, name="H5Acreate_by_name") &
      result(h)
      import :: c_int64_t, c_char
      integer(c_int64_t), value :: loc, atype, space, acpl, aapl, lapl
      character(kind=c_char), intent(in) :: obj(*), attr(*)
      integer(c_int64_t) :: h
    end function
    function H5Awrite(a, mtype, buf) bind(C, name="H5Awrite") result(r)
      import :: c_int, c_int64_t, c_ptr
      integer(c_int64_t), value :: a, mtype
      type(c_ptr), value :: buf
      integer(c_int) :: r
    end function
  end interface

contains

  pure function lower(s) result(r)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: r
    integer :: k
    r = s
    do k = 1, len(s)
      if (r(k:k) >= 'A' .and. r(k:k) <= 'Z') r(k:k) = achar(iachar(r(k:k)) + 32)
    end do
  end function

  pure function unquote(s) result(r)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: r
    integer :: n
    r = adjustl(s)
    n = len_trim(r)
    if (n >= 2) then
      if ((r(1:1) == "'" .and. r(n:n) == "'") .or. (r(1:1) == '"' .and. r(n:n) == '"')) r = r(2:n - 1)
    end if
  end function

  logical function to_logical(s)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: t
    t = lower(adjustl(s))
    to_logical = (t(1:1) == 't' .or. t(1:2) == '.t' .or. t(1:1) == 'y' .or. t(1:1) == '1')
  end function

  real(rk) function to_real(s, key)
    character(len=*), intent(in) :: s, key
    integer :: ios
    read(s, *, iostat=ios) to_real
    if (ios /= 0) error stop 'input.ini: bad real value for '//trim(key)
  end function

  !> Read the [section] key = value pairs the hot path uses (mod_input.f90:200-410).
  !> pressure_input_bc_t%read_pressure_input (pressure_input_bc.f90:93-178): count the records (a first
  !> one starting with '#' is a header), then one list-directed read of time, pressure and density per
  !> row; pressure and density times the scale factor (code units: this host takes unit reference scales).
  subroutine read_pressure_input(path, scale, t, p, r)
    character(len=*), intent(in) :: path
    real(rk), intent(in) :: scale
    real(rk), allocatable, intent(out) :: t(:), p(:), r(:)
    character(len=300) :: buf
    logical :: header, exists
    integer :: u, ios, nlines, k
    inquire(file=path, exist=exists)
    if (.not. exists) then
      write(*, '(a)') "Error: Pressure input file '"//trim(path)//"' not found"
      error stop 1
    end if
    header = .false.
    nlines = 0
    open(newunit=u, file=path, status='old', action='read')
    do
      buf = ''
      read(u, *, iostat=ios) buf
      if (nlines == 0) then
        buf = adjustl(buf)
        if (buf(1:1) == '#') header = .true.
      end if
      if (ios /= 0) exit
      nlines = nlines + 1
    end do
    close(u)
    open(newunit=u, file=path, status='old', action='read')
    if (header) then
      read(u, *, iostat=ios)
      nlines = nlines - 1
    end if
    allocate(t(max(nlines, 0)), p(max(nlines, 0)), r(max(nlines, 0)))
    t = 0.0_rk
    p = 0.0_rk
    r = 0.0_rk
    do k = 1, nlines
      read(u, *, iostat=ios) t(k), p(k), r(k)
      if (ios /= 0) exit
    end do
    close(u)
    p = (p * 1.0_rk) * scale
    r = (r * 1.0_rk) * scale
  end subroutine

  subroutine read_input_ini(filename, inp)
    character(len=*), intent(in) :: filename
    type(cato_input_t), intent(out) :: inp
    character(len=1024) :: line
    character(len=64) :: section, key
    character(len=960) :: val
    integer :: u, ios, eq, c
    open(newunit=u, file=filename, status='old', action='read', iostat=ios)
    if (ios /= 0) error stop 'input.ini: cannot open '//trim(filename)
    section = ''
    do
      read(u, '(a)', iostat=ios) line
      if (ios /= 0) exit
      c = scan(line, ';#')
      if (c > 0) line = line(1:c - 1)
      line = adjustl(line)
      if (len_trim(line) == 0) cycle
      if (line(1:1) == '[') then
        section = lower(line(2:index(line, ']') - 1))
        cycle
      end if
      eq = index(line, '=')
      if (eq == 0) cycle
      key = lower(adjustl(line(1:eq - 1)))
      val = unquote(trim(line(eq + 1:)))
      select case (trim(section)//'.'//trim(key))
      case ('general.title'); inp%title = val
      case ('time.max_time'); inp%max_time = to_real(val, key)
      case ('time.initial_delta_t'); inp%initial_delta_t = to_real(val, key)
      case ('time.use_constant_delta_t'); inp%use_constant_delta_t = to_logical(val)
      case ('time.cfl'); inp%cfl = to_real(val, key)
      case ('time.integration_strategy'); inp%integration_strategy = val
      case ('time.max_iterations'); inp%max_iterations = int(to_real(val, key))
      case ('time.enable_startup_time'); inp%enable_startup_time = to_logical(val)
      case ('time.startup_period'); inp%startup_period = to_real(val, key)
      case ('time.startup_delta_t'); inp%startup_delta_t = to_real(val, key)
      case ('time.allow_subcycle'); inp%allow_subcycle = to_logical(val)
      case ('time.subcycle_error_tolerance'); inp%subcycle_error_tolerance = to_real(val, key)
      case ('source_terms.enable_source_terms'); inp%enable_source_terms = to_logical(val)
      case ('restart.restart_from_file'); inp%restart_from_file = to_logical(val)
      case ('restart.restart_file'); inp%restart_file = val
      case ('reference_state.non_dimensionalize'); inp%non_dimensionalize = to_logical(val)
      case ('reference_state.reference_length'); inp%reference_length = to_real(val, key)
      case ('reference_state.reference_density'); inp%reference_density = to_real(val, key)
      case ('reference_state.reference_velocity'); inp%reference_velocity = to_real(val, key)
      case ('physics.polytropic_index'); inp%polytropic_index = to_real(val, key)
      case ('scheme.flux_solver'); inp%flux_solver = val
      case ('scheme.spatial_reconstruction'); inp%spatial_reconstruction = val
      case ('scheme.limiter'); inp%limiter = val
      case ('scheme.apply_low_mach_fix'); inp%apply_low_mach_fix = to_logical(val)
      case ('initial_conditions.read_from_file'); inp%read_from_file = to_logical(val)
      case ('initial_conditions.initial_condition_file'); inp%initial_condition_file = val
      case ('boundary_conditions.plus_x'); inp%plus_x = val
      case ('boundary_conditions.minus_x'); inp%minus_x = val
      case ('boundary_conditions.plus_y'); inp%plus_y = val
      case ('boundary_conditions.minus_y'); inp%minus_y = val
      case ('boundary_conditions.outflow_ambient_pressure'); inp%outflow_ambient_pressure = to_real(val, key)
      case ('boundary_conditions.apply_constant_bc_pressure'); inp%apply_constant_bc_pressure = to_logical(val)
      case ('boundary_conditions.constant_bc_pressure_value'); inp%constant_bc_pressure_value = to_real(val, key)
      case ('boundary_conditions.bc_density'); inp%bc_density = to_real(val, key)
      case ('boundary_conditions.bc_pressure_input_file'); inp%bc_pressure_input_file = val
      case ('boundary_conditions.bc_pressure_scale_factor'); inp%bc_pressure_scale_factor = to_real(val, key)
      case ('io.contour_interval_dt'); inp%contour_interval_dt = to_real(val, key)
      case ('io.plot_64bit'); inp%plot_64bit = to_logical(val)
      case ('io.format'); inp%contour_format = val
      case ('io.plot_volume'); inp%plot_volume = to_logical(val)
      case ('general.units'); inp%unit_system = val
      case ('grid.n_ghost_layers'); inp%n_ghost_layers = int(to_real(val, key))
      end select
    end do
    close(u)
  end subroutine

  function cstr(s) result(r)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: r
    r = trim(s)//c_null_char
  end function

  !> 2D dataset of doubles, stored transposed: returns a Fortran (n1, n2) array.
  subroutine read_2d(f, name, a)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: name
    real(rk), allocatable, target, intent(out) :: a(:, :)
    integer(c_int64_t) :: d, s, dims(2)
    integer(c_int) :: r
    d = H5Dopen2(f, cstr(name), H5P_DEFAULT)
    if (d < 0) error stop 'initial conditions: missing dataset '//trim(name)
    s = H5Dget_space(d)
    if (H5Sget_simple_extent_ndims(s) /= 2) error stop 'initial conditions: dataset is not 2D: '//trim(name)
    r = H5Sget_simple_extent_dims(s, dims, c_null_ptr)
    allocate(a(dims(2), dims(1)))
    r = H5Dread(d, h5t_native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, c_loc(a))
    if (r < 0) error stop 'initial conditions: cannot read '//trim(name)
    r = H5Sclose(s)
    r = H5Dclose(d)
  end subroutine

  !> The 'units' attribute of a dataset ('' when absent; fixed-length strings as h5fortran writes them).
  function units_of(f, name) result(u)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: name
    character(len=32) :: u
    character(kind=c_char, len=256), target :: buf
    integer(c_int64_t) :: a, ft, mt
    integer(c_size_t) :: n
    integer(c_int) :: r
    u = ''
    if (H5Aexists_by_name(f, cstr(name), cstr('units'), H5P_DEFAULT) <= 0) return
    a = H5Aopen_by_name(f, cstr(name), cstr('units'), H5P_DEFAULT, H5P_DEFAULT)
    ft = H5Aget_type(a)
    if (H5Tis_variable_str(ft) > 0) error stop 'restart: variable-length units attributes are not supported'
    n = min(H5Tget_size(ft), int(len(buf), c_size_t))
    mt = H5Tcopy(h5t_c_s1)
    r = H5Tset_size(mt, n)
    buf = ' '
    r = H5Aread(a, mt, c_loc(buf))
    r = H5Tclose(mt)
    r = H5Tclose(ft)
    r = H5Aclose(a)
    u = buf(1:min(int(n), 32))
    if (index(u, c_null_char) > 0) u = u(1:index(u, c_null_char) - 1)
  end function

  !> One scalar (any numeric type, any shape with one element) read as a double.
  real(rk) function read_scalar(f, name) result(x)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: name
    real(rk), target :: buf(1)
    integer(c_int64_t) :: d
    integer(c_int) :: r
    d = H5Dopen2(f, cstr(name), H5P_DEFAULT)
    if (d < 0) error stop 'restart: missing dataset '//trim(name)
    r = H5Dread(d, h5t_native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, c_loc(buf))
    if (r < 0) error stop 'restart: cannot read '//trim(name)
    r = H5Dclose(d)
    x = buf(1)
  end function

  !> [restart] restart_file: master_puppeteer.f90:106-121 (/time in 's' or 'ns', /iteration) and
  !> fluid_t%initialize_from_hdf5 (mod_fluid.f90:273-345: density 'g/cc', velocity 'km/s' or 'cm/s',
  !> pressure 'barye' or 'Mbar'; halo'd fields).  A file written by write_state (it holds /rho_E) is a
  !> checkpoint of the conserved state instead: conserved = .true. and rho/u/v/p hold rho, rho u, rho v,
  !> rho E on the interior.  Datasets without a units attribute are taken in code units.
  subroutine read_restart(filename, time, iteration, conserved, rho, u, v, p)
    character(len=*), intent(in) :: filename
    real(rk), intent(out) :: time
    integer, intent(out) :: iteration
    logical, intent(out) :: conserved
    real(rk), allocatable, intent(out) :: rho(:, :), u(:, :), v(:, :), p(:, :)
    integer(c_int64_t) :: f
    integer(c_int) :: r
    character(len=32) :: un
    r = H5open()
    f = H5Fopen(cstr(filename), H5F_ACC_RDONLY, H5P_DEFAULT)
    if (f < 0) error stop 'restart: cannot open '//trim(filename)
    time = read_scalar(f, '/time')
    iteration = nint(read_scalar(f, '/iteration'))
    un = units_of(f, '/time')
    select case (trim(un))
    case ('ns'); time = time * 1.0e-9_rk
    case ('s', '')
    case default; error stop "Unknown time units in .h5 file. Acceptable units are 'ns' or 's'."
    end select
    conserved = H5Lexists(f, cstr('/rho_E'), H5P_DEFAULT) > 0
    if (conserved) then
      call read_2d(f, '/density', rho)
      call read_2d(f, '/rho_u', u)
      call read_2d(f, '/rho_v', v)
      call read_2d(f, '/rho_E', p)
    else
      call read_2d(f, '/density', rho)
      un = units_of(f, '/density')
      if (len_trim(un) > 0 .and. trim(un) /= 'g/cc') &
        error stop "Unknown density units in .h5 file. Acceptable units are 'g/cc'"
      call read_2d(f, '/x_velocity', u)
      call read_2d(f, '/y_velocity', v)
      un = units_of(f, '/x_velocity')
      select case (trim(un))
      case ('km/s')
        u = u * 1.0e5_rk
        v = v * 1.0e5_rk
      case ('cm/s', '')
      case default; error stop "Unknown velocity units in .h5 file. Acceptable units are 'km/s' and 'cm/s'"
      end select
      call read_2d(f, '/pressure', p)
      un = units_of(f, '/pressure')
      select case (trim(un))
      case ('Mbar'); p = p * 1.0e12_rk
      case ('barye', '')
      case default; error stop "Unknown pressure units in .h5 file. Acceptable units are 'barye', 'Mbar'"
      end select
    end if
    r = H5Fclose(f)
  end subroutine

  !> The deck: node coordinates and cell primitives incl. ghost cells (grid_block_2d.f90:194-265).
  subroutine read_initial_conditions(filename, n_ghost, x, y, rho, u, v, p)
    character(len=*), intent(in) :: filename
    integer, intent(out) :: n_ghost
    real(rk), allocatable, intent(out) :: x(:, :), y(:, :), rho(:, :), u(:, :), v(:, :), p(:, :)
    integer(c_int64_t) :: f, d
    integer(c_int), target :: nh(1)
    integer(c_int) :: r
    r = H5open()
    f = H5Fopen(cstr(filename), H5F_ACC_RDONLY, H5P_DEFAULT)
    if (f < 0) error stop 'initial conditions: cannot open '//trim(filename)
    d = H5Dopen2(f, cstr('/n_ghost_layers'), H5P_DEFAULT)
    if (d < 0) error stop 'initial conditions: missing /n_ghost_layers'
    r = H5Dread(d, h5t_native_int, H5S_ALL, H5S_ALL, H5P_DEFAULT, c_loc(nh))
    r = H5Dclose(d)
    n_ghost = nh(1)
    call read_2d(f, '/x', x)
    call read_2d(f, '/y', y)
    call read_2d(f, '/density', rho)
    call read_2d(f, '/x_velocity', u)
    call read_2d(f, '/y_velocity', v)
    call read_2d(f, '/pressure', p)
    r = H5Fclose(f)
  end subroutine

  subroutine write_2d(f, name, a)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: name
    real(rk), target, intent(in) :: a(:, :)
    integer(c_int64_t) :: s, d, dims(2)
    integer(c_int) :: r
    dims = [int(size(a, 2), c_int64_t), int(size(a, 1), c_int64_t)]
    s = H5Screate_simple(2_c_int, dims, c_null_ptr)
    d = H5Dcreate2(f, cstr(name), h5t_native_double, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
    if (d < 0) error stop 'output: cannot create dataset '//trim(name)
    r = H5Dwrite(d, h5t_native_double, H5S_ALL, H5S_ALL, H5P_DEFAULT, c_loc(a))
    if (r < 0) error stop 'output: cannot write '//trim(name)
    r = H5Dclose(d)
    r = H5Sclose(s)
  end subroutine

  subroutine write_scalar(f, name, val)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: name
    real(rk), intent(in) :: val
    real(rk), target :: a(1, 1)
    a(1, 1) = val
    call write_2d(f, name, a)
  end subroutine

  subroutine write_scalar_value(f, name, mtype, buf)
    integer(c_int64_t), intent(in) :: f, mtype
    character(len=*), intent(in) :: name
    type(c_ptr), intent(in) :: buf
    integer(c_int64_t) :: sp, d
    integer(c_int) :: r
    sp = H5Screate(H5S_SCALAR)
    d = H5Dcreate2(f, cstr(name), mtype, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
    if (d < 0) error stop 'contour: cannot create dataset '//trim(name)
    r = H5Dwrite(d, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, buf)
    r = H5Dclose(d)
    r = H5Sclose(sp)
  end subroutine

  subroutine write_string(f, name, text)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: name, text
    character(kind=c_char, len=max(len_trim(text), 1)), target :: buf
    integer(c_int64_t) :: t
    integer(c_int) :: r
    buf = trim(text)
    if (len_trim(text) == 0) buf = ' '
    t = H5Tcopy(h5t_c_s1)
    r = H5Tset_size(t, int(len(buf), c_size_t))
    call write_scalar_value(f, name, t, c_loc(buf))
    r = H5Tclose(t)
  end subroutine

  !> A fixed-length string attribute (as h5fortran's writeattr stores them).
  subroutine write_attr(f, obj, name, text)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: obj, name, text
    character(kind=c_char, len=max(len_trim(text), 1)), target :: buf
    integer(c_int64_t) :: t, sp, a
    integer(c_int) :: r
    buf = trim(text)
    if (len_trim(text) == 0) buf = ' '
    t = H5Tcopy(h5t_c_s1)
    r = H5Tset_size(t, int(len(buf), c_size_t))
    sp = H5Screate(H5S_SCALAR)
    a = H5Acreate_by_name(f, cstr(obj), cstr(name), t, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
    if (a < 0) error stop 'contour: cannot create attribute '//trim(name)//' of '//trim(obj)
    r = H5Awrite(a, t, c_loc(buf))
    r = H5Aclose(a)
    r = H5Sclose(sp)
    r = H5Tclose(t)
  end subroutine

  !> contour_writer_t%write_2d_real_data (contour_writer.f90:544-587): FP64, or FP32 with values beyond
  !> huge(real32) clipped to +-huge and magnitudes below tiny(real32) written as 0; description and units.
  subroutine write_field(f, name, a, plot_64bit, description, units)
    integer(c_int64_t), intent(in) :: f
    character(len=*), intent(in) :: name, description, units
    real(rk), intent(in) :: a(:, :)
    logical, intent(in) :: plot_64bit
    real(c_float), allocatable, target :: b(:, :)
    integer(c_int64_t) :: s, d, dims(2)
    integer(c_int) :: r
    integer :: i, j
    if (plot_64bit) then
      call write_2d(f, name, a)
    else
      allocate(b(size(a, 1), size(a, 2)))
      do j = 1, size(a, 2)
        do i = 1, size(a, 1)
          if (abs(a(i, j)) > huge(1.0_c_float)) then
            b(i, j) = sign(huge(1.0_c_float), real(a(i, j), c_float))
          else if (abs(a(i, j)) < tiny(1.0_c_float)) then
            b(i, j) = 0.0_c_float
          else
            b(i, j) = real(a(i, j), c_float)
          end if
        end do
      end do
      dims = [int(size(b, 2), c_int64_t), int(size(b, 1), c_int64_t)]
      s = H5Screate_simple(2_c_int, dims, c_null_ptr)
      d = H5Dcreate2(f, cstr(name), h5t_native_float, s, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
      if (d < 0) error stop 'output: cannot create dataset '//trim(name)
      r = H5Dwrite(d, h5t_native_float, H5S_ALL, H5S_ALL, H5P_DEFAULT, c_loc(b))
      r = H5Dclose(d)
      r = H5Sclose(s)
    end if
    call write_attr(f, name, 'description', description)
    call write_attr(f, name, 'units', units)
  end subroutine

  !> contour_writer_t%write_hdf5 (src/lib/io/contour_writer.f90:193-384): results/step_NNNNNNN.h5 with the
  !> scalars and the halo'd node and cell fields (ghost cells included, as plot_ghost_cells defaults to true)
  !> in the deck's output unit system (set_output_unit_system, units.f90:94-143: 'cgs' or 'icf'; the code
  !> units are cgs with the unit reference state this host takes), each with its description and units
  !> attributes; FP32 unless plot_64bit.  The xdmf companion is not written.
  subroutine write_contour(folder, title, iteration, time, delta_t, n_ghost, x, y, rho, u, v, p, cs, volume, &
                           unit_system, plot_64bit)
    character(len=*), intent(in) :: folder, title, unit_system
    integer, intent(in) :: iteration, n_ghost
    real(rk), intent(in) :: time, delta_t
    real(rk), intent(in) :: x(:, :), y(:, :), rho(:, :), u(:, :), v(:, :), p(:, :), cs(:, :)
    real(rk), intent(in), optional :: volume(:, :)
    logical, intent(in) :: plot_64bit
    character(len=512) :: path
    integer(c_int64_t) :: f
    integer(c_int), target :: ival
    real(rk), target :: rval
    integer(c_int) :: r
    real(rk) :: tu, lu, vu, pu
    character(len=10) :: tl, ll, vl, pl
    select case (trim(unit_system))
    case ('icf')
      tu = 1e9_rk; tl = 'ns'; lu = 1e4_rk; ll = 'um'; vu = 1e-5_rk; vl = 'km/s'; pu = 1e-12_rk; pl = 'Mbar'
    case default
      tu = 1.0_rk; tl = 's'; lu = 1.0_rk; ll = 'cm'; vu = 1.0_rk; vl = 'cm/s'; pu = 1.0_rk; pl = 'barye'
    end select
    write(path, '(a, a, i0.7, a)') trim(folder), '/step_', iteration, '.h5'
    r = H5open()
    f = H5Fcreate(cstr(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    if (f < 0) error stop 'contour: cannot create '//trim(path)
    call write_string(f, '/title', title)
    ival = iteration
    call write_scalar_value(f, '/iteration', h5t_native_int, c_loc(ival))
    call write_attr(f, '/iteration', 'description', 'Iteration Count')
    call write_attr(f, '/iteration', 'units', 'dimensionless')
    rval = time * tu
    call write_scalar_value(f, '/time', h5t_native_double, c_loc(rval))
    call write_attr(f, '/time', 'description', 'Simulation Time')
    call write_attr(f, '/time', 'units', tl)
    rval = delta_t
    call write_scalar_value(f, '/delta_t', h5t_native_double, c_loc(rval))
    call write_attr(f, '/delta_t', 'description', 'Simulation Timestep')
    call write_attr(f, '/delta_t', 'units', 'seconds')
    ival = n_ghost
    call write_scalar_value(f, '/n_ghost_layers', h5t_native_int, c_loc(ival))
    call write_field(f, '/x', x * lu, plot_64bit, 'X Coordinate', ll)
    call write_field(f, '/y', y * lu, plot_64bit, 'Y Coordinate', ll)
    call write_field(f, '/density', rho, plot_64bit, 'Cell Density', 'g/cc')
    call write_field(f, '/x_velocity', u * vu, plot_64bit, 'Cell X Velocity', vl)
    call write_field(f, '/y_velocity', v * vu, plot_64bit, 'Cell Y Velocity', vl)
    call write_field(f, '/pressure', p * pu, plot_64bit, 'Cell Pressure', pl)
    call write_field(f, '/sound_speed', cs * vu, plot_64bit, 'Cell Sound Speed', vl)
    if (present(volume)) call write_field(f, '/volume', volume, plot_64bit, 'Cell Volume', 'cc')
    r = H5Fclose(f)
  end subroutine

  !> contour_writer_t%write_xdmf (contour_writer.f90:386-528): the XDMF companion of step_NNNNNNN.h5 that
  !> points a viewer at its datasets (structured 2D mesh from /x, /y; density, sound speed, pressure, the
  !> velocity vector, the Mach-number function and the volume as cell attributes, with their unit labels).
  subroutine write_xdmf(folder, iteration, time, nx_cells, ny_cells, unit_system, plot_volume)
    character(len=*), intent(in) :: folder, unit_system
    integer, intent(in) :: iteration, nx_cells, ny_cells
    real(rk), intent(in) :: time
    logical, intent(in) :: plot_volume
    character(len=512) :: path
    character(len=32) :: h5
    character(len=40) :: cells, nodes
    character(len=10) :: tl, vl, pl
    real(rk) :: tu
    integer :: u
    select case (trim(unit_system))
    case ('icf')
      tu = 1e9_rk; tl = 'ns'; vl = 'km/s'; pl = 'Mbar'
    case default
      tu = 1.0_rk; tl = 's'; vl = 'cm/s'; pl = 'barye'
    end select
    write(h5, '(a, i0.7, a)') 'step_', iteration, '.h5'
    write(path, '(a, a, i0.7, a)') trim(folder), '/step_', iteration, '.xdmf'
    write(cells, '(i0, 1x, i0)') ny_cells, nx_cells
    write(nodes, '(i0, 1x, i0)') ny_cells + 1, nx_cells + 1
    open(newunit=u, file=trim(path), status='replace', action='write')
    write(u, '(a)') '<?xml version="1.0" ?>'
    write(u, '(a)') '<Xdmf version="2.2">'
    write(u, '(a)') '  <Domain>'
    write(u, '(a)') '    <Grid GridType="Uniform" Name="grid">'
    write(u, '(a, g0.3, a)') '      <Time Value="', time * tu, ' '//trim(tl)//'"/>'
    write(u, '(a)') '      <Topology NumberOfElements="'//trim(nodes)//'" TopologyType="2DSMesh"/>'
    write(u, '(a)') '      <Geometry GeometryType="X_Y">'
    call item('        ', nodes, '/x')
    call item('        ', nodes, '/y')
    write(u, '(a)') '      </Geometry>'
    call scalar('Density [g/cc]', '/density')
    call scalar('Sound Speed ['//trim(vl)//']', '/sound_speed')
    call scalar('Pressure ['//trim(pl)//']', '/pressure')
    write(u, '(a)') '      <Attribute AttributeType="Vector" Center="Cell" Name="Velocity ['//trim(vl)// &
      ']" Dimensions="'//trim(cells)//' 2">'
    write(u, '(a)') '        <DataItem Dimensions="'//trim(cells)//' 2" ItemType="Function" Function="JOIN($0, $1)">'
    call func_item('/x_velocity')
    call func_item('/y_velocity')
    write(u, '(a)') '        </DataItem>'
    write(u, '(a)') '      </Attribute>'
    write(u, '(a)') '      <Attribute AttributeType="Vector" Center="Cell" Name="Mach" Dimensions="'//trim(cells)//' 2">'
    write(u, '(a)') '        <DataItem Dimensions="'//trim(cells)// &
      ' 2" ItemType="Function" Function="JOIN(ABS($0/$2), ABS($1/$2))">'
    call func_item('/x_velocity')
    call func_item('/y_velocity')
    call func_item('/sound_speed')
    write(u, '(a)') '        </DataItem>'
    write(u, '(a)') '      </Attribute>'
    if (plot_volume) call scalar('Volume [cc]', '/volume')
    write(u, '(a)') '    </Grid>'
    write(u, '(a)') '  </Domain>'
    write(u, '(a)') '</Xdmf>'
    close(u)
  contains
    subroutine item(indent, dims, ds)
      character(len=*), intent(in) :: indent, dims, ds
      write(u, '(a)') indent//'<DataItem Dimensions="'//trim(dims)//'" Format="HDF" NumberType="Float" Precision="4">'// &
        trim(h5)//':'//ds//'</DataItem>'
    end subroutine
    subroutine scalar(label, ds)
      character(len=*), intent(in) :: label, ds
      write(u, '(a)') '      <Attribute AttributeType="Scalar" Center="Cell" Name="'//label//'">'
      call item('        ', cells, ds)
      write(u, '(a)') '      </Attribute>'
    end subroutine
    subroutine func_item(ds)
      character(len=*), intent(in) :: ds
      write(u, '(a)') '          <DataItem DataType="Float" Dimensions="'//trim(cells)//'" Format="HDF" Precision="4">'// &
        trim(h5)//':'//ds//'</DataItem>'
    end subroutine
  end subroutine

  !> Interior state of a solve: conserved and primitive fields, time and iteration count.
  subroutine write_state(filename, time, iteration, rho, rho_u, rho_v, rho_E, u, v, p)
    character(len=*), intent(in) :: filename
    real(rk), intent(in) :: time
    integer, intent(in) :: iteration
    real(rk), intent(in) :: rho(:, :), rho_u(:, :), rho_v(:, :), rho_E(:, :), u(:, :), v(:, :), p(:, :)
    integer(c_int64_t) :: f
    integer(c_int) :: r
    r = H5open()
    f = H5Fcreate(cstr(filename), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    if (f < 0) error stop 'output: cannot create '//trim(filename)
    call write_scalar(f, '/time', time)
    call write_scalar(f, '/iteration', real(iteration, rk))
    call write_2d(f, '/density', rho)
    call write_2d(f, '/rho_u', rho_u)
    call write_2d(f, '/rho_v', rho_v)
    call write_2d(f, '/rho_E', rho_E)
    call write_2d(f, '/x_velocity', u)
    call write_2d(f, '/y_velocity', v)
    call write_2d(f, '/pressure', p)
    r = H5Fclose(f)
  end subroutine

end module cato_deck
