! =============================================================================
! grhd_c.f90 -- iso_c_binding interface to libgrhd_hip.so (include/grhd_c.h).
!
! type(grhd_config) mirrors struct grhd_config field by field (bind(C) gives it
! the C layout); every entry point returns an integer status, 0 on success.
! Host arrays are passed in the reference's own layout:
!   p(5, -ng:nx+ng, -ng:ny+ng, -ng:nz+ng),  u(5, nx, ny, nz).
! =============================================================================
module grhd_c
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: grhd_config, c_ptr, c_null_ptr, c_int, c_double, c_int32_t, c_int64_t
  public :: grhd_create, grhd_validate, grhd_destroy, grhd_last_error, grhd_dims, grhd_initial_conditions
  public :: grhd_set_primitives, grhd_get_primitives, grhd_set_conserved, grhd_get_conserved
  public :: grhd_compute_dt, grhd_step, grhd_advance, grhd_perturb, grhd_counters
  public :: grhd_total_mass, grhd_convergence_norms, grhd_global_diagnostics, grhd_ppi_modes
  public :: grhd_accretion_rate, grhd_lapse_shift
  public :: grhd_last_timing, grhd_synchronize, grhd_error_message

  integer, parameter, public :: GRHD_OK = 0

  type, bind(C) :: grhd_config
    integer(c_int32_t) :: problem, metric, geometry, log_r
    real(c_double) :: adiabatic_index, bh_mass, spin
    integer(c_int32_t) :: nx, ny, nz, reconstruction
    real(c_double) :: r_min, r_max, y_min, y_max, z_min, z_max
    integer(c_int32_t) :: riemann_solver, tvd_limiter, shock_sensor, precision
    real(c_double) :: cfl, final_time
    real(c_double) :: sod_interface, sod_density_left, sod_density_right
    real(c_double) :: sod_pressure_left, sod_pressure_right
    real(c_double) :: strong_interface, strong_density_left, strong_density_right
    real(c_double) :: strong_pressure_left, strong_pressure_right
    real(c_double) :: shu_interface, shu_density_left, shu_pressure_left, shu_density_right
    real(c_double) :: shu_density_amplitude, shu_wave_number, shu_pressure_right
    real(c_double) :: khi_half_width, khi_density_inner, khi_density_outer
    real(c_double) :: khi_velocity_inner, khi_velocity_outer, khi_pressure
    real(c_double) :: khi_perturbation_amplitude, khi_perturbation_wave_number
    real(c_double) :: khi_perturbation_width
    real(c_double) :: jet_ambient_density, jet_ambient_pressure, jet_ambient_velocity
    real(c_double) :: jet_nozzle_radius, jet_nozzle_length, jet_density, jet_pressure
    real(c_double) :: jet_velocity
    real(c_double) :: convergence_density, convergence_amplitude, convergence_velocity
    real(c_double) :: convergence_pressure, convergence_wave_number
    real(c_double) :: michel_critical_radius, michel_critical_density
    real(c_double) :: dust_accretion_constant
    real(c_double) :: offaxis_radial_center, offaxis_phi_center, offaxis_width
    real(c_double) :: offaxis_background_density, offaxis_background_pressure
    real(c_double) :: offaxis_density_amplitude, offaxis_pressure_amplitude
    real(c_double) :: fm_inner_radius, fm_pressure_max_radius, fm_polytropic_constant
    integer(c_int32_t) :: allow_open_theta
    integer(c_int32_t) :: reserved(7)
  end type grhd_config

  interface
    integer(c_int) function grhd_create(cfg, ctx) bind(C, name='grhd_create')
      import :: c_int, c_ptr, grhd_config
      type(grhd_config), intent(in) :: cfg
      type(c_ptr), intent(out) :: ctx
    end function grhd_create
    integer(c_int) function grhd_validate(cfg) bind(C, name='grhd_validate')
      import :: c_int, grhd_config
      type(grhd_config), intent(in) :: cfg
    end function grhd_validate
    integer(c_int) function grhd_destroy(ctx) bind(C, name='grhd_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function grhd_destroy
    type(c_ptr) function grhd_last_error(ctx) bind(C, name='grhd_last_error')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function grhd_last_error
    integer(c_int) function grhd_dims(ctx, dims) bind(C, name='grhd_dims')
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: dims(4)
    end function grhd_dims
    integer(c_int) function grhd_initial_conditions(ctx) bind(C, name='grhd_initial_conditions')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function grhd_initial_conditions
    integer(c_int) function grhd_set_primitives(ctx, p) bind(C, name='grhd_set_primitives')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: p(*)
    end function grhd_set_primitives
    integer(c_int) function grhd_lapse_shift(cfg, alpha, beta) bind(C, name='grhd_lapse_shift')
      import :: c_int, c_double, grhd_config
      type(grhd_config), intent(in) :: cfg
      real(c_double), intent(out) :: alpha(*), beta(*)
    end function grhd_lapse_shift
    integer(c_int) function grhd_get_primitives(ctx, p) bind(C, name='grhd_get_primitives')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: p(*)
    end function grhd_get_primitives
    integer(c_int) function grhd_set_conserved(ctx, u) bind(C, name='grhd_set_conserved')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
    end function grhd_set_conserved
    integer(c_int) function grhd_get_conserved(ctx, u) bind(C, name='grhd_get_conserved')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: u(*)
    end function grhd_get_conserved
    integer(c_int) function grhd_compute_dt(ctx, dt) bind(C, name='grhd_compute_dt')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
    end function grhd_compute_dt
    integer(c_int) function grhd_step(ctx, dt, nsteps) bind(C, name='grhd_step')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int32_t), value :: nsteps
    end function grhd_step
    integer(c_int) function grhd_advance(ctx, t_final, max_steps, t, steps_done, last_dt) &
        bind(C, name='grhd_advance')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: t_final
      integer(c_int32_t), value :: max_steps
      real(c_double), intent(inout) :: t
      integer(c_int32_t), intent(out) :: steps_done
      real(c_double), intent(out) :: last_dt
    end function grhd_advance
    integer(c_int) function grhd_perturb(ctx, kind, amplitude, seed) bind(C, name='grhd_perturb')
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kind
      real(c_double), value :: amplitude
      integer(c_int64_t), value :: seed
    end function grhd_perturb
    integer(c_int) function grhd_counters(ctx, counters) bind(C, name='grhd_counters')
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: counters(4)
    end function grhd_counters
    integer(c_int) function grhd_total_mass(ctx, mass) bind(C, name='grhd_total_mass')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mass
    end function grhd_total_mass
    integer(c_int) function grhd_convergence_norms(ctx, t, norms, imax) bind(C, name='grhd_convergence_norms')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), value :: t
      real(c_double), intent(out) :: norms(3)
      integer(c_int32_t), intent(out) :: imax
    end function grhd_convergence_norms
    integer(c_int) function grhd_global_diagnostics(ctx, out, counts) bind(C, name='grhd_global_diagnostics')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(8)
      integer(c_int64_t), intent(out) :: counts(2)
    end function grhd_global_diagnostics
    integer(c_int) function grhd_ppi_modes(ctx, out) bind(C, name='grhd_ppi_modes')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(17)
    end function grhd_ppi_modes
    integer(c_int) function grhd_accretion_rate(ctx, mdot, probe) bind(C, name='grhd_accretion_rate')
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: mdot
      integer(c_int32_t), intent(out) :: probe
    end function grhd_accretion_rate
    integer(c_int) function grhd_last_timing(ctx, ms, launches) bind(C, name='grhd_last_timing')
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms
      integer(c_int64_t), intent(out) :: launches
    end function grhd_last_timing
    integer(c_int) function grhd_synchronize(ctx) bind(C, name='grhd_synchronize')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function grhd_synchronize
  end interface

contains

  ! grhd_last_error as a Fortran string.
  function grhd_error_message(ctx) result(msg)
    type(c_ptr), value :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: cp
    integer :: i
    msg = ' '
    cp = grhd_last_error(ctx)
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, s, [256])
    do i = 1, 256
      if (s(i) == c_null_char) exit
      msg(i:i) = s(i)
    end do
  end function grhd_error_message

end module grhd_c
