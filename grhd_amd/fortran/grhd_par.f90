! =============================================================================
! grhd_par.f90 -- the .par operator surface of the reference driver.
!
! Same file format and semantics as the reference (parameters.f90:95-356):
! "name = value" lines, '#' comments, case-insensitive names and symbols,
! duplicate and unknown names are fatal, theta/phi bounds in units of pi
! (fractions a/b allowed), initial-condition names namespaced per problem.
! The problem preset is applied first (initialization.f90:483-821) and the
! file has the last word.  Fields the update reads go into grhd_config;
! run-control fields (perturbation, strides, output folder) into run_control.
! Validation of the physics fields is done by grhd_create (the C library
! applies parameters.f90:359-588).
! =============================================================================
module grhd_par
  use grhd_c
  implicit none
  private
  public :: run_control, read_par, read_restart_par, problem_label

  real(c_double), parameter :: PI = 3.141592653589793d0

  type :: run_control
    character(len=256) :: output_folder = 'output'
    character(len=64) :: output_prefix = 'out'
    integer :: perturbation = 0          ! 0 none, 1 pressure noise, 2 density noise
    integer :: perturbation_seed = 3435
    real(c_double) :: perturbation_time = 1000.0d0
    real(c_double) :: perturbation_amplitude = 0.01d0
    integer :: mass_monitor_stride = 1000
    real(c_double) :: save_interval = 1.0d0
    real(c_double) :: checkpoint_interval = 100.0d0   ! variables.f90:264
    ! observers (variables.f90:132-148): accretion rate every extraction_stride
    ! steps, azimuthal modes and global diagnostics every diagnostic_stride
    logical :: extract_mdot = .false., ppi_diagnostics = .false.
    integer :: extraction_stride = 10, diagnostic_stride = 100
    ! VTK frames every save_interval (output.f90:154-471); vtk_mapping
    logical :: vtk_untwisted = .false.
    logical :: restart = .false.
    character(len=1024) :: restart_file = ' '
  end type run_control

contains

  pure function lower(s) result(o)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: o
    integer :: i, c
    o = s
    do i = 1, len(s)
      c = iachar(s(i:i))
      if (c >= iachar('A') .and. c <= iachar('Z')) o(i:i) = achar(c + 32)
    end do
  end function lower

  function problem_label(id) result(label)
    integer, intent(in) :: id
    character(len=32) :: label
    select case(id)
    case(1); label = 'sod'
    case(2); label = 'strong_shock'
    case(3); label = 'shu_osher'
    case(4); label = 'kelvin_helmholtz'
    case(5); label = 'axisymmetric_jet'
    case(6); label = 'convergence'
    case(7); label = 'michel'
    case(8); label = 'dust'
    case(9); label = 'offaxis_blast'
    case(10); label = 'fishbone_equatorial'
    case(11); label = 'fishbone_sagittal'
    case(12); label = 'fishbone_3d'
    case default; label = 'unknown'
    end select
  end function problem_label

  subroutine fail(file_name, line, msg)
    character(len=*), intent(in) :: file_name, msg
    integer, intent(in) :: line
    if (line > 0) then
      write(*,'(A,1X,A,A,I0,A,1X,A)') 'ERROR in', trim(file_name), ':', line, ':', trim(msg)
    else
      write(*,'(A,1X,A,A,1X,A)') 'ERROR in', trim(file_name), ':', trim(msg)
    end if
    error stop 1
  end subroutine fail

  integer function problem_id(text)
    character(len=*), intent(in) :: text
    select case(trim(lower(adjustl(text))))
    case('sod'); problem_id = 1
    case('strong_shock', 'blast_wave'); problem_id = 2
    case('shu_osher'); problem_id = 3
    case('kelvin_helmholtz', 'khi'); problem_id = 4
    case('axisymmetric_jet', 'jet'); problem_id = 5
    case('convergence', 'convergence_test'); problem_id = 6
    case('michel', 'michel_accretion'); problem_id = 7
    case('dust', 'dust_accretion'); problem_id = 8
    case('offaxis_blast', 'off_axis_blast'); problem_id = 9
    case('fishbone_equatorial', 'fishbone_moncrief_equatorial'); problem_id = 10
    case('fishbone_sagittal', 'fishbone_sagital', 'fishbone_moncrief_sagittal'); problem_id = 11
    case('fishbone_3d', 'fishbone_moncrief_3d'); problem_id = 12
    case default; problem_id = -1
    end select
  end function problem_id

  ! Module defaults (variables.f90:111-243) and the problem preset
  ! (initialization.f90:483-821).
  subroutine apply_preset(c, id)
    type(grhd_config), intent(out) :: c
    integer, intent(in) :: id
    c%problem = id
    c%metric = 0; c%geometry = 0; c%log_r = 0
    c%adiabatic_index = 1.4d0; c%bh_mass = 0.0d0; c%spin = 0.0d0
    c%nx = 1; c%ny = 1; c%nz = 1
    c%r_min = 0.0d0; c%r_max = 1.0d0; c%y_min = 0.0d0; c%y_max = 1.0d0
    c%z_min = 0.0d0; c%z_max = 1.0d0
    c%reconstruction = 5; c%riemann_solver = 1; c%tvd_limiter = 3; c%shock_sensor = 1
    c%precision = 0; c%cfl = 0.4d0; c%final_time = 0.4d0
    c%sod_interface = 0.5d0; c%sod_density_left = 1.0d0; c%sod_density_right = 0.125d0
    c%sod_pressure_left = 1.0d0; c%sod_pressure_right = 0.1d0
    c%strong_interface = 0.5d0; c%strong_density_left = 1.0d0; c%strong_density_right = 1.0d0
    c%strong_pressure_left = 1000.0d0; c%strong_pressure_right = 0.01d0
    c%shu_interface = 0.5d0; c%shu_density_left = 5.0d0; c%shu_pressure_left = 50.0d0
    c%shu_density_right = 2.0d0; c%shu_density_amplitude = 0.3d0
    c%shu_wave_number = 50.0d0; c%shu_pressure_right = 5.0d0
    c%khi_half_width = 0.25d0; c%khi_density_inner = 2.0d0; c%khi_density_outer = 1.0d0
    c%khi_velocity_inner = 0.5d0; c%khi_velocity_outer = -0.5d0; c%khi_pressure = 2.5d0
    c%khi_perturbation_amplitude = 0.01d0; c%khi_perturbation_wave_number = 10.0d0
    c%khi_perturbation_width = 0.05d0
    c%jet_ambient_density = 10.0d0; c%jet_ambient_pressure = 0.01d0
    c%jet_ambient_velocity = 0.0d0; c%jet_nozzle_radius = 1.0d0; c%jet_nozzle_length = 1.0d0
    c%jet_density = 0.1d0; c%jet_pressure = 0.01d0; c%jet_velocity = 0.99d0
    c%convergence_density = 1.0d0; c%convergence_amplitude = 0.1d0
    c%convergence_velocity = 0.5d0; c%convergence_pressure = 1.0d0
    c%convergence_wave_number = 2.0d0
    c%michel_critical_radius = 100.0d0; c%michel_critical_density = 0.1d0
    c%dust_accretion_constant = -0.5d0
    c%offaxis_radial_center = 15.0d0; c%offaxis_phi_center = PI; c%offaxis_width = 1.5d0
    c%offaxis_background_density = 0.1d0; c%offaxis_background_pressure = 0.1d0
    c%offaxis_density_amplitude = 1.0d0; c%offaxis_pressure_amplitude = 1.0d0
    c%fm_inner_radius = 6.25d0; c%fm_pressure_max_radius = 9.25d0
    c%fm_polytropic_constant = 0.0015d0
    c%allow_open_theta = 0; c%reserved = 0
    select case(id)
    case(1)
      c%shock_sensor = 0; c%adiabatic_index = 1.4d0; c%nx = 200
    case(2, 3)
      c%shock_sensor = 0; c%adiabatic_index = 5.0d0/3.0d0; c%nx = 400
    case(4)
      c%shock_sensor = 0; c%adiabatic_index = 5.0d0/3.0d0
      c%nx = 400; c%r_min = -0.5d0; c%r_max = 0.5d0
      c%nz = 400; c%z_min = -0.5d0; c%z_max = 0.5d0
      c%final_time = 3.0d0; c%cfl = 0.25d0
    case(5)
      c%geometry = 1; c%shock_sensor = 1; c%adiabatic_index = 5.0d0/3.0d0
      c%nx = 320; c%r_min = 0.0d0; c%r_max = 16.0d0; c%y_max = PI
      c%nz = 800; c%z_min = 0.0d0; c%z_max = 40.0d0
      c%final_time = 120.0d0; c%cfl = 0.25d0
    case(6)
      c%shock_sensor = 0; c%adiabatic_index = 4.0d0/3.0d0; c%nx = 2048
      c%final_time = 0.7d0; c%cfl = 0.10d0
    case(7, 8, 9)
      c%geometry = 2; c%metric = 1; c%adiabatic_index = 4.0d0/3.0d0; c%bh_mass = 1.0d0
      c%y_max = PI; c%z_max = 2.0d0*PI
      if (id == 7) then
        c%shock_sensor = 0; c%nx = 400; c%r_min = 1.01d0; c%r_max = 51.0d0; c%nz = 200
        c%final_time = 500.0d0; c%cfl = 0.5d0
      else if (id == 8) then
        c%shock_sensor = 1; c%nx = 500; c%r_min = 1.0d0; c%r_max = 51.0d0; c%nz = 4
        c%final_time = 500.0d0; c%cfl = 0.5d0
      else
        c%shock_sensor = 1; c%nx = 400; c%r_min = 1.0d0; c%r_max = 121.0d0; c%nz = 200
        c%final_time = 2000.0d0; c%cfl = 0.4d0
      end if
    case(10)
      c%geometry = 3; c%metric = 2; c%shock_sensor = 0; c%log_r = 1
      c%adiabatic_index = 4.0d0/3.0d0; c%bh_mass = 1.0d0; c%spin = 0.9d0
      c%nx = 400; c%r_min = 1.2d0; c%r_max = 40.0d0; c%y_max = PI
      c%nz = 200; c%z_max = 2.0d0*PI; c%final_time = 1000.0d0; c%cfl = 0.4d0
    case(11)
      c%geometry = 2; c%metric = 1; c%shock_sensor = 1; c%log_r = 1
      c%adiabatic_index = 4.0d0/3.0d0; c%bh_mass = 1.0d0
      c%nx = 400; c%r_min = 2.0d0; c%r_max = 200.0d0; c%ny = 200; c%y_max = PI
      c%z_max = 2.0d0*PI; c%final_time = 0.0d0; c%cfl = 0.4d0
    case(12)
      c%geometry = 2; c%metric = 1; c%shock_sensor = 1
      c%adiabatic_index = 4.0d0/3.0d0; c%bh_mass = 1.0d0
      c%fm_inner_radius = 6.0d0; c%fm_pressure_max_radius = 12.0d0
      c%fm_polytropic_constant = 0.015d0
      c%nx = 100; c%r_min = 1.0d0; c%r_max = 121.0d0; c%ny = 100; c%y_max = PI
      c%nz = 100; c%z_max = 2.0d0*PI; c%final_time = 35.0d0; c%cfl = 0.4d0
    end select
    if (id == 10 .or. id == 11) then
      c%fm_inner_radius = 6.25d0; c%fm_pressure_max_radius = 9.25d0
      c%fm_polytropic_constant = 0.0015d0
    end if
  end subroutine apply_preset

  subroutine parse_real(text, v, f, ln)
    character(len=*), intent(in) :: text, f
    real(c_double), intent(out) :: v
    integer, intent(in) :: ln
    integer :: ios
    read(text, *, iostat=ios) v
    if (ios /= 0) call fail(f, ln, 'expected a real number: '//trim(text))
  end subroutine parse_real

  subroutine parse_int(text, v, f, ln)
    character(len=*), intent(in) :: text, f
    integer(c_int32_t), intent(out) :: v
    integer, intent(in) :: ln
    integer :: ios
    read(text, *, iostat=ios) v
    if (ios /= 0) call fail(f, ln, 'expected an integer: '//trim(text))
  end subroutine parse_int

  integer function parse_bool(text, f, ln)
    character(len=*), intent(in) :: text, f
    integer, intent(in) :: ln
    select case(trim(lower(adjustl(text))))
    case('1', 'true', '.true.', 'yes', 'si'); parse_bool = 1
    case('0', 'false', '.false.', 'no'); parse_bool = 0
    case default
      parse_bool = 0
      call fail(f, ln, 'expected true/false, yes/no or 1/0: '//trim(text))
    end select
  end function parse_bool

  ! parse_pi_coefficient (parameters.f90:808-835): a real or a fraction a/b.
  real(c_double) function pi_coeff(text, f, ln)
    character(len=*), intent(in) :: text, f
    integer, intent(in) :: ln
    integer :: s
    real(c_double) :: a, b
    s = index(trim(text), '/')
    if (s == 0) then
      call parse_real(text, pi_coeff, f, ln)
      return
    end if
    if (s == 1 .or. s == len_trim(text) .or. index(text(s+1:), '/') > 0) &
      call fail(f, ln, 'invalid fraction of pi: '//trim(text))
    call parse_real(text(:s-1), a, f, ln)
    call parse_real(text(s+1:), b, f, ln)
    if (b == 0.0d0) call fail(f, ln, 'zero denominator: '//trim(text))
    pi_coeff = a / b
  end function pi_coeff

  subroutine read_par(file_name, c, rc)
    character(len=*), intent(in) :: file_name
    type(grhd_config), intent(out) :: c
    type(run_control), intent(out) :: rc
    character(len=1024) :: line, key, val
    character(len=64) :: seen(128)
    integer :: u, ios, ln, eq, hash, nseen, i, pid, pass
    character(len=16) :: prefix

    pid = -1
    do pass = 1, 2
      open(newunit=u, file=trim(file_name), status='old', action='read', iostat=ios)
      if (ios /= 0) call fail(file_name, 0, 'cannot open the parameter file')
      if (pass == 2) then
        call apply_preset(c, pid)
        prefix = ic_prefix(pid)
        ! the presets' observers (initialization.f90:656-802): the accretion
        ! rate for Michel, dust and the sagittal and 3D tori, the azimuthal
        ! modes for the equatorial and 3D tori
        rc%extract_mdot = any(pid == [7, 8, 11, 12])
        rc%ppi_diagnostics = any(pid == [10, 12])
        call preset_output(rc, pid)
      end if
      ln = 0
      nseen = 0
      do
        read(u, '(A)', iostat=ios) line
        if (ios < 0) exit
        ln = ln + 1
        hash = index(line, '#')
        if (hash > 0) line(hash:) = ' '
        if (len_trim(line) == 0) cycle
        eq = index(line, '=')
        if (eq <= 1 .or. eq == len_trim(line)) call fail(file_name, ln, 'expected name = value')
        if (index(line(eq+1:), '=') > 0) call fail(file_name, ln, 'more than one =')
        key = lower(trim(adjustl(line(:eq-1))))
        val = trim(adjustl(line(eq+1:)))
        if (len_trim(key) > 64) call fail(file_name, ln, 'parameter name too long')
        do i = 1, nseen
          if (trim(seen(i)) == trim(key)) call fail(file_name, ln, 'duplicated parameter: '//trim(key))
        end do
        nseen = nseen + 1
        if (nseen > size(seen)) call fail(file_name, ln, 'too many parameters')
        seen(nseen) = trim(key)
        if (pass == 1) then
          if (trim(key) == 'problem') then
            pid = problem_id(val)
            if (pid < 0) call fail(file_name, ln, 'unknown problem: '//trim(val))
          else if (trim(key) == 'restart') then
            rc%restart = parse_bool(val, file_name, ln) == 1
          else if (trim(key) == 'restart_file') then
            rc%restart_file = unquote_long(val)
          end if
          cycle
        end if
        if (is_ic_key(key) .and. index(trim(key), trim(prefix)) /= 1) &
          call fail(file_name, ln, 'parameter does not belong to the selected problem: '//trim(key))
        call apply_key(c, rc, key, val, file_name, ln)
      end do
      close(u)
      ! a restart takes the problem from the checkpoint (parameters.f90:84-91);
      ! the caller reads it, then applies the restart-control keys
      if (pass == 1 .and. rc%restart) then
        if (len_trim(rc%restart_file) == 0) &
          call fail(file_name, 0, 'restart = true requires the parameter restart_file')
        return
      end if
      if (pass == 1 .and. pid < 0) &
        call fail(file_name, 0, 'missing mandatory parameter problem for a cold start')
    end do
    if (c%metric == 2 .and. c%geometry == 2) c%geometry = 3   ! parameters.f90:428-431
  end subroutine read_par

  ! The presets' frame interval, file prefix and folder (initialization.f90:
  ! 504-820)
  subroutine preset_output(rc, id)
    type(run_control), intent(inout) :: rc
    integer, intent(in) :: id
    select case(id)
    case(1); rc%save_interval = 0.1d0; rc%output_prefix = 'Sod'; rc%output_folder = 'ShockTubes'
    case(2); rc%save_interval = 0.1d0; rc%output_prefix = 'Strong'; rc%output_folder = 'ShockTubes'
    case(3); rc%save_interval = 0.1d0; rc%output_prefix = 'Shu'; rc%output_folder = 'ShockTubes'
    case(4); rc%save_interval = 0.1d0; rc%output_prefix = 'KHI'; rc%output_folder = 'KelvinHelmholtz'
    case(5)
      rc%save_interval = 1.0d0; rc%output_prefix = 'JET'; rc%output_folder = 'AxisymmetricJet_extreme'
    case(6)
      rc%save_interval = 0.1d0; rc%output_prefix = 'Conv'; rc%output_folder = 'ConvergenceTest_N2048'
    case(7); rc%save_interval = 1.0d0; rc%output_prefix = 'MA'; rc%output_folder = 'Michel_test_2D'
    case(8); rc%save_interval = 50.0d0; rc%output_prefix = 'D'; rc%output_folder = 'Dust2D'
    case(9); rc%save_interval = 25.0d0; rc%output_prefix = 'OR'; rc%output_folder = 'OffAxis'
    case(10)
      rc%save_interval = 10.0d0; rc%output_prefix = 'FM_KS_a0.9d0'; rc%output_folder = 'FM_KS_a0.9d0_data'
    case(11)
      rc%save_interval = 1.0d0; rc%output_prefix = 'FM_Sag'; rc%output_folder = 'FishboneMoncrief_Sagital'
    case(12)
      rc%save_interval = 1.0d0; rc%output_prefix = 'FM_3D'; rc%output_folder = 'FishboneMoncrief_3D'
    end select
  end subroutine preset_output

  ! Keys a restart may change (is_restart_control, parameters.f90:716-731), plus
  ! precision (the FP64 or mixed face arithmetic, which the reference does not
  ! have); every other key belongs to the checkpoint.
  logical function is_restart_control(key)
    character(len=*), intent(in) :: key
    select case(trim(key))
    case('restart', 'restart_file', 'final_time', 'cfl', 'save_interval', &
         'checkpoint_interval', 'output_prefix', 'output_folder', 'vtk_mapping', &
         'shock_sensor', 'extract_gw', 'extract_mdot', 'ppi_diagnostics', 'diagnostic_stride', &
         'extraction_stride', 'mass_monitor_stride', 'perturbation', 'perturbation_seed', &
         'perturbation_time', 'perturbation_amplitude', 'perturbation_mode', 'precision')
      is_restart_control = .true.
    case default
      is_restart_control = .false.
    end select
  end function is_restart_control

  ! The .par of a restart over the configuration read from the checkpoint
  ! (read_parameter_file with restart_mode, parameters.f90:95-160).
  subroutine read_restart_par(file_name, c, rc)
    character(len=*), intent(in) :: file_name
    type(grhd_config), intent(inout) :: c
    type(run_control), intent(inout) :: rc
    character(len=1024) :: line, key, val
    integer :: u, ios, ln, eq, hash
    open(newunit=u, file=trim(file_name), status='old', action='read', iostat=ios)
    if (ios /= 0) call fail(file_name, 0, 'cannot open the parameter file')
    ln = 0
    do
      read(u, '(A)', iostat=ios) line
      if (ios < 0) exit
      ln = ln + 1
      hash = index(line, '#')
      if (hash > 0) line(hash:) = ' '
      if (len_trim(line) == 0) cycle
      eq = index(line, '=')
      key = lower(trim(adjustl(line(:eq-1))))
      val = trim(adjustl(line(eq+1:)))
      if (.not. is_restart_control(key)) call fail(file_name, ln, &
        'this parameter belongs to the checkpoint and cannot change on restart: '//trim(key))
      call apply_key(c, rc, key, val, file_name, ln)
    end do
    close(u)
  end subroutine read_restart_par

  function ic_prefix(id) result(p)
    integer, intent(in) :: id
    character(len=16) :: p
    select case(id)
    case(1); p = 'sod_'
    case(2); p = 'strong_'
    case(3); p = 'shu_'
    case(4); p = 'khi_'
    case(5); p = 'jet_'
    case(6); p = 'convergence_'
    case(7); p = 'michel_'
    case(8); p = 'dust_'
    case(9); p = 'offaxis_'
    case default; p = 'fm_'
    end select
  end function ic_prefix

  logical function is_ic_key(key)
    character(len=*), intent(in) :: key
    is_ic_key = index(key, 'sod_') == 1 .or. index(key, 'strong_') == 1 .or. &
                index(key, 'shu_') == 1 .or. index(key, 'khi_') == 1 .or. &
                index(key, 'jet_') == 1 .or. index(key, 'convergence_') == 1 .or. &
                index(key, 'michel_') == 1 .or. index(key, 'dust_') == 1 .or. &
                index(key, 'offaxis_') == 1 .or. index(key, 'fm_') == 1
  end function is_ic_key

  subroutine apply_key(c, rc, key, val, f, ln)
    type(grhd_config), intent(inout) :: c
    type(run_control), intent(inout) :: rc
    character(len=*), intent(in) :: key, val, f
    integer, intent(in) :: ln
    real(c_double) :: x
    integer(c_int32_t) :: n
    select case(trim(key))
    case('problem')
    case('restart'); rc%restart = parse_bool(val, f, ln) == 1
    case('restart_file'); rc%restart_file = unquote_long(val)
    case('checkpoint_interval')
      call parse_real(val, x, f, ln)
      if (.not. (x > 0.0d0)) call fail(f, ln, 'checkpoint_interval must be > 0')
      rc%checkpoint_interval = x
    case('extract_mdot'); rc%extract_mdot = parse_bool(val, f, ln) == 1
    case('ppi_diagnostics'); rc%ppi_diagnostics = parse_bool(val, f, ln) == 1
    case('diagnostic_stride', 'extraction_stride')
      call parse_int(val, n, f, ln)
      if (n < 1) call fail(f, ln, trim(key)//' must be >= 1')
      if (trim(key) == 'diagnostic_stride') then
        rc%diagnostic_stride = n
      else
        rc%extraction_stride = n
      end if
    case('vtk_mapping')
      ! parse_vtk_mapping (parameters.f90:966-980)
      select case(trim(lower(adjustl(val))))
      case('physical', 'ks_cartesian', 'complete', 'completo'); rc%vtk_untwisted = .false.
      case('untwisted', 'unrolled', 'desenrollado'); rc%vtk_untwisted = .true.
      case default; call fail(f, ln, 'vtk_mapping must be physical or untwisted')
      end select
    case('extract_gw')
      ! the gravitational-wave strain is not built
    case('metric')
      select case(trim(lower(adjustl(val))))
      case('minkowski'); c%metric = 0
      case('eddington_finkelstein', 'ef'); c%metric = 1
      case('kerr_schild', 'ks'); c%metric = 2
      case default; call fail(f, ln, 'unknown metric: '//trim(val))
      end select
    case('geometry')
      select case(trim(lower(adjustl(val))))
      case('cartesian'); c%geometry = 0
      case('cylindrical'); c%geometry = 1
      case('spherical'); c%geometry = 2
      case('spheroidal'); c%geometry = 3
      case default; call fail(f, ln, 'unknown geometry: '//trim(val))
      end select
    case('reconstruction')
      select case(trim(lower(adjustl(val))))
      case('godunov'); c%reconstruction = 1
      case('tvd'); c%reconstruction = 2
      case('weno3'); c%reconstruction = 3
      case('mp5'); c%reconstruction = 4
      case('weno5'); c%reconstruction = 5
      case default; call fail(f, ln, 'reconstruction must be godunov, tvd, weno3, mp5 or weno5')
      end select
    case('riemann_solver')
      select case(trim(lower(adjustl(val))))
      case('hlle'); c%riemann_solver = 1
      case('hllc'); c%riemann_solver = 2
      case default; call fail(f, ln, 'riemann_solver must be hlle or hllc')
      end select
    case('tvd_limiter')
      select case(trim(lower(adjustl(val))))
      case('minmod'); c%tvd_limiter = 1
      case('superbee'); c%tvd_limiter = 2
      case('mc'); c%tvd_limiter = 3
      case default; call fail(f, ln, 'tvd_limiter must be minmod, superbee or mc')
      end select
    case('precision')
      select case(trim(lower(adjustl(val))))
      case('fp64'); c%precision = 0
      case('mixed'); c%precision = 1
      case default; call fail(f, ln, 'precision must be fp64 or mixed')
      end select
    case('perturbation')
      select case(trim(lower(adjustl(val))))
      case('none'); rc%perturbation = 0
      case('pressure_noise'); rc%perturbation = 1
      case('density_noise'); rc%perturbation = 2
      case default; call fail(f, ln, 'perturbation must be none, pressure_noise or density_noise')
      end select
    case('shock_sensor'); c%shock_sensor = parse_bool(val, f, ln)
    case('logarithmic_r'); c%log_r = parse_bool(val, f, ln)
    case('nx'); call parse_int(val, c%nx, f, ln)
    case('ny'); call parse_int(val, c%ny, f, ln)
    case('nz'); call parse_int(val, c%nz, f, ln)
    case('perturbation_seed'); call parse_int(val, n, f, ln); rc%perturbation_seed = n
    case('mass_monitor_stride'); call parse_int(val, n, f, ln); rc%mass_monitor_stride = n
    case('theta_min'); c%y_min = pi_coeff(val, f, ln) * PI
    case('theta_max'); c%y_max = pi_coeff(val, f, ln) * PI
    case('phi_min'); c%z_min = pi_coeff(val, f, ln) * PI
    case('phi_max'); c%z_max = pi_coeff(val, f, ln) * PI
    case('convergence_wave_number'); c%convergence_wave_number = pi_coeff(val, f, ln)
    case('offaxis_phi_center'); c%offaxis_phi_center = pi_coeff(val, f, ln) * PI
    case('output_folder'); rc%output_folder = unquote(val)
    case('output_prefix'); rc%output_prefix = unquote(val)
    case default
      call parse_real(val, x, f, ln)
      select case(trim(key))
      case('bh_mass'); c%bh_mass = x
      case('spin'); c%spin = x
      case('adiabatic_index'); c%adiabatic_index = x
      case('r_min'); c%r_min = x
      case('r_max'); c%r_max = x
      case('y_min'); c%y_min = x
      case('y_max'); c%y_max = x
      case('z_min'); c%z_min = x
      case('z_max'); c%z_max = x
      case('final_time'); c%final_time = x
      case('cfl'); c%cfl = x
      case('save_interval'); rc%save_interval = x
      case('perturbation_time'); rc%perturbation_time = x
      case('perturbation_amplitude'); rc%perturbation_amplitude = x
      case('perturbation_mode')
      case('sod_interface'); c%sod_interface = x
      case('sod_density_left'); c%sod_density_left = x
      case('sod_density_right'); c%sod_density_right = x
      case('sod_pressure_left'); c%sod_pressure_left = x
      case('sod_pressure_right'); c%sod_pressure_right = x
      case('strong_interface'); c%strong_interface = x
      case('strong_density_left'); c%strong_density_left = x
      case('strong_density_right'); c%strong_density_right = x
      case('strong_pressure_left'); c%strong_pressure_left = x
      case('strong_pressure_right'); c%strong_pressure_right = x
      case('shu_interface'); c%shu_interface = x
      case('shu_density_left'); c%shu_density_left = x
      case('shu_pressure_left'); c%shu_pressure_left = x
      case('shu_density_right'); c%shu_density_right = x
      case('shu_density_amplitude'); c%shu_density_amplitude = x
      case('shu_wave_number'); c%shu_wave_number = x
      case('shu_pressure_right'); c%shu_pressure_right = x
      case('khi_half_width'); c%khi_half_width = x
      case('khi_density_inner'); c%khi_density_inner = x
      case('khi_density_outer'); c%khi_density_outer = x
      case('khi_velocity_inner'); c%khi_velocity_inner = x
      case('khi_velocity_outer'); c%khi_velocity_outer = x
      case('khi_pressure'); c%khi_pressure = x
      case('khi_perturbation_amplitude'); c%khi_perturbation_amplitude = x
      case('khi_perturbation_wave_number'); c%khi_perturbation_wave_number = x
      case('khi_perturbation_width'); c%khi_perturbation_width = x
      case('jet_ambient_density'); c%jet_ambient_density = x
      case('jet_ambient_pressure'); c%jet_ambient_pressure = x
      case('jet_ambient_velocity'); c%jet_ambient_velocity = x
      case('jet_nozzle_radius'); c%jet_nozzle_radius = x
      case('jet_nozzle_length'); c%jet_nozzle_length = x
      case('jet_density'); c%jet_density = x
      case('jet_pressure'); c%jet_pressure = x
      case('jet_velocity'); c%jet_velocity = x
      case('convergence_density'); c%convergence_density = x
      case('convergence_amplitude'); c%convergence_amplitude = x
      case('convergence_velocity'); c%convergence_velocity = x
      case('convergence_pressure'); c%convergence_pressure = x
      case('michel_critical_radius'); c%michel_critical_radius = x
      case('michel_critical_density'); c%michel_critical_density = x
      case('dust_accretion_constant'); c%dust_accretion_constant = x
      case('offaxis_radial_center'); c%offaxis_radial_center = x
      case('offaxis_width'); c%offaxis_width = x
      case('offaxis_background_density'); c%offaxis_background_density = x
      case('offaxis_background_pressure'); c%offaxis_background_pressure = x
      case('offaxis_density_amplitude'); c%offaxis_density_amplitude = x
      case('offaxis_pressure_amplitude'); c%offaxis_pressure_amplitude = x
      case('fm_inner_radius'); c%fm_inner_radius = x
      case('fm_pressure_max_radius'); c%fm_pressure_max_radius = x
      case('fm_polytropic_constant'); c%fm_polytropic_constant = x
      case default
        call fail(f, ln, 'unknown parameter: '//trim(key))
      end select
    end select
  end subroutine apply_key

  function unquote_long(text) result(s)
    character(len=*), intent(in) :: text
    character(len=1024) :: s
    integer :: n
    s = adjustl(text)
    n = len_trim(s)
    if (n >= 2) then
      if ((s(1:1) == '"' .and. s(n:n) == '"') .or. (s(1:1) == "'" .and. s(n:n) == "'")) &
        s = s(2:n-1)
    end if
  end function unquote_long

  function unquote(text) result(s)
    character(len=*), intent(in) :: text
    character(len=256) :: s
    integer :: n
    s = adjustl(text)
    n = len_trim(s)
    if (n >= 2) then
      if ((s(1:1) == '"' .and. s(n:n) == '"') .or. (s(1:1) == "'" .and. s(n:n) == "'")) &
        s = s(2:n-1)
    end if
  end function unquote

end module grhd_par
