! =============================================================================
! grhd_ckpt -- GRHDCP_V4 checkpoints (output.f90:474-609), written and read
! record for record with the reference's types and lengths, so that a run
! restarts from a checkpoint of either code: the driver's own checkpoints
! restart the reference, and the reference's restart this driver.
!
! Record order: magic; nghost nx ny nz; r_min r_max y_min y_max z_min z_max;
! bh_mass adb_idx a_spin; five flags; metric_type geom_type; final_time CFL
! save_interval; output_prefix output_folder; case_id case_name scheme_name;
! rec_method_id tvd_limiter_id riemann_solver_id; the nine records of the
! problems' initial parameters (output.f90:607-629); the perturbation state
! (three records); the time; the step; up(5,nx,ny,nz).
! =============================================================================
module grhd_ckpt
  use, intrinsic :: iso_c_binding
  use grhd_c
  use grhd_par
  implicit none
  private
  public :: checkpoint_state, write_checkpoint, read_checkpoint, checkpoint_file_name, ghost_depth_of
  public :: scheme_name_of

  character(len=12), parameter :: MAGIC_V4 = 'GRHDCP_V4'

  ! The driver's state that a checkpoint carries beyond grhd_config
  type :: checkpoint_state
    real(c_double) :: t = 0.0d0
    integer :: n_steps = 0
    logical :: perturbation_applied = .false.
  end type checkpoint_state

contains

  ! nghost of the reference (initialization.f90:27-52)
  integer function ghost_depth_of(rec)
    integer(c_int32_t), intent(in) :: rec
    select case(rec)
    case(1); ghost_depth_of = 1
    case(2, 3); ghost_depth_of = 2
    case default; ghost_depth_of = 3
    end select
  end function ghost_depth_of

  ! case_name of the reference presets (initialization.f90:484-798)
  function case_name_of(problem) result(name)
    integer(c_int32_t), intent(in) :: problem
    character(len=50) :: name
    select case(problem)
    case(1); name = 'Sod'
    case(2); name = 'Strong'
    case(3); name = 'Shu'
    case(4); name = 'KHI'
    case(5); name = 'Jet'
    case(6); name = 'Conv'
    case(7); name = 'MichelA'
    case(8); name = 'Dust'
    case(9); name = 'OffAxis'
    case(10); name = 'FishMoncEqu'
    case(11); name = 'FishMoncSag'
    case default; name = 'FishMonc3D'
    end select
  end function case_name_of

  ! metric_type / geom_type strings (parameters.f90:924-963)
  function metric_name_of(metric) result(name)
    integer(c_int32_t), intent(in) :: metric
    character(len=30) :: name
    select case(metric)
    case(0); name = 'Minkowski'
    case(1); name = 'Eddington-Finkelstein'
    case default; name = 'Kerr-Schild'
    end select
  end function metric_name_of

  function geom_name_of(geometry) result(name)
    integer(c_int32_t), intent(in) :: geometry
    character(len=20) :: name
    select case(geometry)
    case(0); name = 'Cartesian'
    case(1); name = 'Cylindrical'
    case(2); name = 'Spherical'
    case default; name = 'Spheroidal'
    end select
  end function geom_name_of

  ! scheme_name (choose_numerical_architecture, initialization.f90:23-79)
  function scheme_name_of(c) result(name)
    type(grhd_config), intent(in) :: c
    character(len=20) :: name
    select case(c%reconstruction)
    case(1); name = 'godunov'
    case(2)
      select case(c%tvd_limiter)
      case(1); name = 'minmod'
      case(2); name = 'superbee'
      case default; name = 'mc'
      end select
    case(3); name = 'weno3'
    case(4); name = 'mp5'
    case default; name = 'weno5'
    end select
    if (c%riemann_solver == 2) then
      name = trim(name)//'_hllc'
    else
      name = trim(name)//'_hlle'
    end if
  end function scheme_name_of

  ! <output_folder>/checkpoint_<scheme_name>_step_<n, 9 digits>.rst (output.f90:481-482)
  function checkpoint_file_name(c, rc, n_steps) result(path)
    type(grhd_config), intent(in) :: c
    type(run_control), intent(in) :: rc
    integer, intent(in) :: n_steps
    character(len=1024) :: path
    write(path, '(A, "/checkpoint_", A, "_step_", I9.9, ".rst")') &
      trim(rc%output_folder), trim(scheme_name_of(c)), n_steps
  end function checkpoint_file_name

  ! save_checkpoint (output.f90:474-527).  The accretion-rate and azimuthal-mode
  ! observer flags and diagnostic_stride are the run's; gravitational-wave
  ! extraction is not part of this path and is written .false.;
  ! perturbation_mode keeps the reference's default.
  subroutine write_checkpoint(path, c, rc, st, up)
    character(len=*), intent(in) :: path
    type(grhd_config), intent(in) :: c
    type(run_control), intent(in) :: rc
    type(checkpoint_state), intent(in) :: st
    real(c_double), intent(in) :: up(:,:,:,:)
    integer :: u, ios
    integer :: nghost, nx, ny, nz, case_id, rec_id, lim_id, rs_id
    integer :: pert_type, pert_seed, diag_stride, step_num
    logical :: use_log_r, use_sensor, no_gw, apply_pert, applied
    character(len=30) :: metric_type
    character(len=20) :: geom_type, scheme_name
    character(len=50) :: output_prefix, case_name
    character(len=256) :: output_folder
    real(c_double) :: save_interval, pert_mode

    nghost = ghost_depth_of(c%reconstruction)
    nx = c%nx; ny = c%ny; nz = c%nz
    use_log_r = c%log_r /= 0
    use_sensor = c%shock_sensor /= 0
    no_gw = .false.
    metric_type = metric_name_of(c%metric)
    geom_type = geom_name_of(c%geometry)
    save_interval = rc%save_interval
    output_prefix = rc%output_prefix
    output_folder = rc%output_folder
    case_id = c%problem
    case_name = case_name_of(c%problem)
    scheme_name = scheme_name_of(c)
    rec_id = c%reconstruction; lim_id = c%tvd_limiter; rs_id = c%riemann_solver
    apply_pert = rc%perturbation /= 0
    applied = st%perturbation_applied
    pert_type = rc%perturbation
    pert_seed = rc%perturbation_seed
    diag_stride = rc%diagnostic_stride
    pert_mode = 1.0d0
    step_num = st%n_steps

    open(newunit=u, file=trim(path), status='replace', action='write', form='unformatted', iostat=ios)
    if (ios /= 0) then
      write(*,'(A,A)') 'ERROR: cannot create the checkpoint ', trim(path)
      error stop 1
    end if
    write(u) MAGIC_V4
    write(u) nghost, nx, ny, nz
    write(u) c%r_min, c%r_max, c%y_min, c%y_max, c%z_min, c%z_max
    write(u) c%bh_mass, c%adiabatic_index, c%spin
    write(u) use_log_r, use_sensor, no_gw, rc%extract_mdot, rc%ppi_diagnostics
    write(u) metric_type, geom_type
    write(u) c%final_time, c%cfl, save_interval
    write(u) output_prefix, output_folder
    write(u) case_id, case_name, scheme_name
    write(u) rec_id, lim_id, rs_id
    write(u) c%sod_interface, c%sod_density_left, c%sod_density_right, &
      c%sod_pressure_left, c%sod_pressure_right
    write(u) c%strong_interface, c%strong_density_left, c%strong_density_right, &
      c%strong_pressure_left, c%strong_pressure_right
    write(u) c%shu_interface, c%shu_density_left, c%shu_pressure_left, &
      c%shu_density_right, c%shu_density_amplitude, c%shu_wave_number, c%shu_pressure_right
    write(u) c%khi_half_width, c%khi_density_inner, c%khi_density_outer, &
      c%khi_velocity_inner, c%khi_velocity_outer, c%khi_pressure, c%khi_perturbation_amplitude, &
      c%khi_perturbation_wave_number, c%khi_perturbation_width
    write(u) c%jet_ambient_density, c%jet_ambient_pressure, &
      c%jet_ambient_velocity, c%jet_nozzle_radius, c%jet_nozzle_length, &
      c%jet_density, c%jet_pressure, c%jet_velocity
    write(u) c%convergence_density, c%convergence_amplitude, &
      c%convergence_velocity, c%convergence_pressure, c%convergence_wave_number
    write(u) c%michel_critical_radius, c%michel_critical_density, c%dust_accretion_constant
    write(u) c%offaxis_radial_center, c%offaxis_phi_center, c%offaxis_width, &
      c%offaxis_background_density, c%offaxis_background_pressure, &
      c%offaxis_density_amplitude, c%offaxis_pressure_amplitude
    write(u) c%fm_inner_radius, c%fm_pressure_max_radius, c%fm_polytropic_constant
    write(u) apply_pert, applied
    write(u) pert_type, pert_seed, diag_stride
    write(u) rc%perturbation_time, rc%perturbation_amplitude, pert_mode
    write(u) st%t
    write(u) step_num
    write(u) up
    close(u)
  end subroutine write_checkpoint

  ! read_checkpoint_metadata + read_checkpoint_state (output.f90:529-604) of a
  ! V4 checkpoint into a configuration, the run control and the state; up is
  ! allocated (5, nx, ny, nz).  Earlier versions and perturbation kinds this
  ! build does not implement are rejected.
  subroutine read_checkpoint(path, c, rc, st, up)
    character(len=*), intent(in) :: path
    type(grhd_config), intent(inout) :: c
    type(run_control), intent(inout) :: rc
    type(checkpoint_state), intent(out) :: st
    real(c_double), allocatable, intent(out) :: up(:,:,:,:)
    integer :: u, ios
    integer :: nghost, nx, ny, nz, case_id, rec_id, lim_id, rs_id
    integer :: pert_type, pert_seed, diag_stride, step_num
    logical :: use_log_r, use_sensor, obs(3), apply_pert, applied
    character(len=12) :: magic
    character(len=30) :: metric_type
    character(len=20) :: geom_type, scheme_name
    character(len=50) :: output_prefix, case_name
    character(len=256) :: output_folder
    real(c_double) :: save_interval, pert_time, pert_amp, pert_mode

    c%precision = 0
    c%allow_open_theta = 0
    c%reserved = 0
    open(newunit=u, file=trim(path), status='old', action='read', form='unformatted', iostat=ios)
    if (ios /= 0) call ckpt_fail(path, 'cannot open the checkpoint')
    read(u, iostat=ios) magic
    if (ios /= 0 .or. magic /= MAGIC_V4) call ckpt_fail(path, 'not a GRHDCP_V4 checkpoint')
    read(u) nghost, nx, ny, nz
    c%nx = nx; c%ny = ny; c%nz = nz
    read(u) c%r_min, c%r_max, c%y_min, c%y_max, c%z_min, c%z_max
    read(u) c%bh_mass, c%adiabatic_index, c%spin
    read(u) use_log_r, use_sensor, obs(1), obs(2), obs(3)
    c%log_r = merge(1, 0, use_log_r)
    c%shock_sensor = merge(1, 0, use_sensor)
    if (obs(1)) write(*,'(A)') 'note: the checkpoint asks for gravitational-wave extraction, ' // &
      'which this driver does not run'
    rc%extract_mdot = obs(2)
    rc%ppi_diagnostics = obs(3)
    read(u) metric_type, geom_type
    select case(trim(metric_type))
    case('Minkowski'); c%metric = 0
    case('Eddington-Finkelstein'); c%metric = 1
    case('Kerr-Schild'); c%metric = 2
    case default; call ckpt_fail(path, 'unknown metric_type '//trim(metric_type))
    end select
    select case(trim(geom_type))
    case('Cartesian'); c%geometry = 0
    case('Cylindrical'); c%geometry = 1
    case('Spherical'); c%geometry = 2
    case('Spheroidal'); c%geometry = 3
    case default; call ckpt_fail(path, 'unknown geom_type '//trim(geom_type))
    end select
    read(u) c%final_time, c%cfl, save_interval
    rc%save_interval = save_interval
    read(u) output_prefix, output_folder
    rc%output_prefix = output_prefix
    rc%output_folder = output_folder
    read(u) case_id, case_name, scheme_name
    if (case_id < 1 .or. case_id > 12) call ckpt_fail(path, 'unknown case_id')
    c%problem = case_id
    read(u) rec_id, lim_id, rs_id
    c%reconstruction = rec_id; c%tvd_limiter = lim_id; c%riemann_solver = rs_id
    if (nghost /= ghost_depth_of(c%reconstruction)) &
      call ckpt_fail(path, 'nghost does not match the reconstruction')
    read(u) c%sod_interface, c%sod_density_left, c%sod_density_right, &
      c%sod_pressure_left, c%sod_pressure_right
    read(u) c%strong_interface, c%strong_density_left, c%strong_density_right, &
      c%strong_pressure_left, c%strong_pressure_right
    read(u) c%shu_interface, c%shu_density_left, c%shu_pressure_left, &
      c%shu_density_right, c%shu_density_amplitude, c%shu_wave_number, c%shu_pressure_right
    read(u) c%khi_half_width, c%khi_density_inner, c%khi_density_outer, &
      c%khi_velocity_inner, c%khi_velocity_outer, c%khi_pressure, c%khi_perturbation_amplitude, &
      c%khi_perturbation_wave_number, c%khi_perturbation_width
    read(u) c%jet_ambient_density, c%jet_ambient_pressure, &
      c%jet_ambient_velocity, c%jet_nozzle_radius, c%jet_nozzle_length, &
      c%jet_density, c%jet_pressure, c%jet_velocity
    read(u) c%convergence_density, c%convergence_amplitude, &
      c%convergence_velocity, c%convergence_pressure, c%convergence_wave_number
    read(u) c%michel_critical_radius, c%michel_critical_density, c%dust_accretion_constant
    read(u) c%offaxis_radial_center, c%offaxis_phi_center, c%offaxis_width, &
      c%offaxis_background_density, c%offaxis_background_pressure, &
      c%offaxis_density_amplitude, c%offaxis_pressure_amplitude
    read(u) c%fm_inner_radius, c%fm_pressure_max_radius, c%fm_polytropic_constant
    read(u) apply_pert, applied
    read(u) pert_type, pert_seed, diag_stride
    read(u) pert_time, pert_amp, pert_mode
    rc%perturbation = merge(pert_type, 0, apply_pert)
    if (rc%perturbation /= 0 .and. .not. applied .and. (pert_type < 1 .or. pert_type > 2)) &
      call ckpt_fail(path, 'pending perturbation kind not supported (density_mode)')
    rc%perturbation_seed = pert_seed
    if (diag_stride > 0) rc%diagnostic_stride = diag_stride
    rc%perturbation_time = pert_time
    rc%perturbation_amplitude = pert_amp
    st%perturbation_applied = applied
    read(u) st%t
    read(u) step_num
    st%n_steps = step_num
    allocate(up(5, nx, ny, nz))
    read(u, iostat=ios) up
    if (ios /= 0) call ckpt_fail(path, 'truncated conserved state')
    close(u)
  end subroutine read_checkpoint

  subroutine ckpt_fail(path, msg)
    character(len=*), intent(in) :: path, msg
    write(*,'(A,A,A,A)') 'ERROR in checkpoint ', trim(path), ': ', trim(msg)
    error stop 1
  end subroutine ckpt_fail

end module grhd_ckpt
