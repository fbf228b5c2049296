! ============================================================================
! ref_harness.f90 -- C-callable harness around the UNMODIFIED reference modules.
!
! Test infrastructure only.  This file is compiled together with the reference's
! own Fortran sources (variables, parameters, metrics, conditions, equations,
! fluxes, initialization, reconstruction, evolution), read in place from the
! reference checkout by oracle/Makefile, into oracle/_ref/libgrhd_ref.so.  The
! harness owns no physics: every number it returns is computed by a reference
! routine.  It only (a) drives the reference's cold-start path from a .par file,
! (b) replays the body of the reference time loop (grhd2.f90:179-265, without
! perturbation, diagnostics or output), and (c) exposes single reference
! routines for point-wise known-answer checks.
!
! Array convention at the C boundary: the reference's own Fortran order, i.e.
! p(neq, -ng:nx+ng, -ng:ny+ng, -ng:nz+ng) and u(neq, nx, ny, nz), passed as flat
! double buffers.
! ============================================================================
module grhd_ref_harness
  use, intrinsic :: iso_c_binding
  use variables
  use parameters
  use metrics
  use conditions
  use output, only: save_vtk_at_time, save_checkpoint, read_checkpoint_metadata, read_checkpoint_state, &
                    calc_global_diagnostics, calc_ppi_modes, calc_m_dot
  use equations
  use fluxes
  use initialization
  use reconstruction
  use evolution
  implicit none
  private
  logical, save :: problem_ready = .false.

contains

  subroutine c_to_f_string(cstr, n, fstr)
    character(kind=c_char), intent(in) :: cstr(*)
    integer(c_int), intent(in) :: n
    character(len=*), intent(out) :: fstr
    integer :: i
    fstr = ' '
    do i = 1, min(int(n), len(fstr))
      fstr(i:i) = cstr(i)
    end do
  end subroutine c_to_f_string

  ! Restores the module defaults that presets do not reset (variables.f90
  ! initialisers), so several problems can be set up in one process.
  subroutine reset_globals()
    n_steps = 0
    integration_time = 0.0d0
    use_log_r = .false.
    bh_mass = 0.0d0
    a_spin = 0.0d0
    use_shock_sensor = .true.
    rec_method_id = REC_WENO5
    riemann_solver_id = RS_HLLE
    tvd_limiter_id = LIM_MC
    is_mhd = .false.
    do_mdot_extraction = .false.
    do_gw_extraction = .false.
    do_ppi_diagnostics = .false.
    apply_perturbation = .false.
    perturbation_applied = .false.
    do_restart = .false.
    save_interval = 1.0d0
    checkpoint_interval = 100.0d0
  end subroutine reset_globals

  subroutine release_problem()
    if (allocated(x)) deallocate(x)
    if (allocated(y)) deallocate(y)
    if (allocated(z)) deallocate(z)
    if (allocated(x_face)) deallocate(x_face)
    if (allocated(y_face)) deallocate(y_face)
    if (allocated(z_face)) deallocate(z_face)
    if (allocated(u)) deallocate(u)
    if (allocated(up)) deallocate(up)
    if (allocated(s)) deallocate(s)
    if (allocated(rhs)) deallocate(rhs)
    if (allocated(p)) deallocate(p)
    if (allocated(alpha_c)) deallocate(alpha_c)
    if (allocated(beta_c)) deallocate(beta_c)
    if (allocated(gamma_c)) deallocate(gamma_c)
    if (allocated(gamma_inv_c)) deallocate(gamma_inv_c)
    if (allocated(gmunu_c)) deallocate(gmunu_c)
    if (allocated(sqrt_gamma_c)) deallocate(sqrt_gamma_c)
    if (allocated(chris_c)) deallocate(chris_c)
    if (allocated(dg_c)) deallocate(dg_c)
    if (allocated(dlna_c)) deallocate(dlna_c)
    if (allocated(alpha_f_x)) deallocate(alpha_f_x)
    if (allocated(beta_f_x)) deallocate(beta_f_x)
    if (allocated(gamma_f_x)) deallocate(gamma_f_x)
    if (allocated(sqrt_gamma_f_x)) deallocate(sqrt_gamma_f_x)
    if (allocated(alpha_f_y)) deallocate(alpha_f_y)
    if (allocated(beta_f_y)) deallocate(beta_f_y)
    if (allocated(gamma_f_y)) deallocate(gamma_f_y)
    if (allocated(sqrt_gamma_f_y)) deallocate(sqrt_gamma_f_y)
    if (allocated(alpha_f_z)) deallocate(alpha_f_z)
    if (allocated(beta_f_z)) deallocate(beta_f_z)
    if (allocated(gamma_f_z)) deallocate(gamma_f_z)
    if (allocated(sqrt_gamma_f_z)) deallocate(sqrt_gamma_f_z)
    if (allocated(michel_injector)) deallocate(michel_injector)
    problem_ready = .false.
  end subroutine release_problem

  ! Cold start exactly as grhd2.f90:57,118-134,151: header, preset, .par
  ! overrides, validation, metric binding, floors, ghost depth, grid, metric
  ! cache, initial conditions, wave-speed binding, then up = P2C(p) and dt.
  integer(c_int) function ref_init(path, n) bind(C, name='ref_init')
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: n
    character(len=1024) :: fname
    integer :: i, j, k
    call c_to_f_string(path, n, fname)
    call release_problem()
    call reset_globals()
    call read_parameter_header(trim(fname))
    call initialize_problem(trim(fname), .false.)
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          call prim_to_cons(p(:,i,j,k), up(:,i,j,k), i, j, k)
        end do
      end do
    end do
    u = up
    n_steps = 0
    integration_time = 0.0d0
    call update_adaptive_dt()
    problem_ready = .true.
    ref_init = 0
  end function ref_init

  ! The restart branch of the reference main program (grhd2.f90:57,63-117):
  ! the .par header (restart and restart_file), the checkpoint's metadata, the
  ! .par's restart-control keys, validation, floors, metric binding, grid,
  ! metric cache, wave speeds, the checkpoint's state, the case's boundary
  ! data, u = up, cons_to_prim and the boundaries.  Returns 1 when the .par
  ! does not ask for a restart and 2 when the checkpoint cannot be opened.
  integer(c_int) function ref_restart(path, n) bind(C, name='ref_restart')
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: n
    character(len=1024) :: fname
    integer :: i, j, k, io_status
    call c_to_f_string(path, n, fname)
    call release_problem()
    call reset_globals()
    call read_parameter_header(trim(fname))
    if (.not. do_restart) then
      ref_restart = 1
      return
    end if
    open(20, file=trim(restart_file), status='old', form='unformatted', iostat=io_status)
    if (io_status /= 0) then
      ref_restart = 2
      return
    end if
    call read_checkpoint_metadata(20)
    call read_parameter_file(trim(fname), .true.)
    call validate_parameters(trim(fname))
    g1 = adb_idx / (adb_idx - 1.0d0)
    rho_floor = 1.0d-10
    p_floor   = rho_floor * 1.0d-3
    D_floor   = rho_floor
    tau_floor = p_floor * (g1 - 1.0d0)
    call set_metric_type()
    call allocate_and_grid()
    call precalculate_metric_cache()
    call init_wavespeed_solver()
    call read_checkpoint_state(20, n_steps, integration_time)
    close(20)
    call prepare_case_boundary_data()
    u = up
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          call cons_to_prim(u(:,i,j,k), p(:,i,j,k), i, j, k)
        end do
      end do
    end do
    call set_boundary_conditions(p)
    problem_ready = .true.
    ref_restart = 0
  end function ref_restart

  ! save_checkpoint (output.f90:474-527) of the current state and time into
  ! folder dir: <dir>/checkpoint_<scheme_name>_step_<n_steps, 9 digits>.rst
  subroutine ref_save_checkpoint(dir, n) bind(C, name='ref_save_checkpoint')
    character(kind=c_char), intent(in) :: dir(*)
    integer(c_int), value :: n
    character(len=256) :: d
    call c_to_f_string(dir, n, d)
    output_folder = d
    call save_checkpoint(n_steps, integration_time)
  end subroutine ref_save_checkpoint

  ! The reference's observers of the current state (output.f90:951-1125),
  ! appended to <dir>/<scheme_name>/{global_diagnostics,ppi_modes,m_dot}.dat;
  ! the azimuthal modes and global diagnostics are switched on for the call.
  subroutine ref_observers(dir, n) bind(C, name='ref_observers')
    character(kind=c_char), intent(in) :: dir(*)
    integer(c_int), value :: n
    character(len=256) :: d
    call c_to_f_string(dir, n, d)
    output_folder = d
    do_ppi_diagnostics = .true.
    call calc_global_diagnostics(integration_time)
    call calc_ppi_modes(integration_time)
    call calc_m_dot(integration_time)
    do_ppi_diagnostics = .false.
  end subroutine ref_observers

  ! save_vtk_at_time (output.f90:458-471) of the current state: the frame of
  ! time t under <dir>/<scheme_name>/
  subroutine ref_save_vtk(dir, n, t) bind(C, name='ref_save_vtk')
    character(kind=c_char), intent(in) :: dir(*)
    integer(c_int), value :: n
    real(c_double), value :: t
    character(len=256) :: d
    call c_to_f_string(dir, n, d)
    output_folder = d
    call save_vtk_at_time(t)
  end subroutine ref_save_vtk

  ! the centre metric cache: alpha_c(nx, ny, nz), beta_c(3, nx, ny, nz)
  ! (initialization.f90:229-230, 283-299)
  subroutine ref_lapse_shift(alpha, beta) bind(C, name='ref_lapse_shift')
    real(c_double), intent(out) :: alpha(nx, ny, nz), beta(3, nx, ny, nz)
    alpha = alpha_c(1:nx, 1:ny, 1:nz)
    beta = beta_c(:, 1:nx, 1:ny, 1:nz)
  end subroutine ref_lapse_shift

  subroutine ref_release() bind(C, name='ref_release')
    call release_problem()
  end subroutine ref_release

  ! dims(1:8) = nx, ny, nz, nghost, case_id, rec_method_id, riemann_solver_id, use_shock_sensor
  subroutine ref_dims(dims) bind(C, name='ref_dims')
    integer(c_int), intent(out) :: dims(8)
    dims(1) = nx; dims(2) = ny; dims(3) = nz; dims(4) = nghost
    dims(5) = case_id; dims(6) = rec_method_id; dims(7) = riemann_solver_id
    dims(8) = merge(1, 0, use_shock_sensor)
  end subroutine ref_dims

  ! vals(1:16) = dx, dy, dz, x_min, x_max, y_min, y_max, z_min, z_max, CFL,
  !              final_time, adb_idx, bh_mass, a_spin, rho_floor, p_floor
  subroutine ref_scalars(vals) bind(C, name='ref_scalars')
    real(c_double), intent(out) :: vals(16)
    vals = (/ dx, dy, dz, x_min, x_max, y_min, y_max, z_min, z_max, CFL, &
              final_time, adb_idx, bh_mass, a_spin, rho_floor, p_floor /)
  end subroutine ref_scalars

  subroutine ref_get_prim(buf) bind(C, name='ref_get_prim')
    real(c_double), intent(out) :: buf(*)
    integer :: nt
    nt = size(p)
    buf(1:nt) = reshape(p, (/ nt /))
  end subroutine ref_get_prim

  ! Loads a full ghosted primitive array and rebuilds the conserved state from
  ! it, as the reference does after a perturbation (grhd2.f90:206-218).
  subroutine ref_set_prim(buf) bind(C, name='ref_set_prim')
    real(c_double), intent(in) :: buf(*)
    integer :: i, j, k
    p = reshape(buf(1:size(p)), shape(p))
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          call prim_to_cons(p(:,i,j,k), up(:,i,j,k), i, j, k)
        end do
      end do
    end do
    u = up
  end subroutine ref_set_prim

  ! Overwrites p without touching the conserved arrays.
  subroutine ref_put_prim(buf) bind(C, name='ref_put_prim')
    real(c_double), intent(in) :: buf(*)
    p = reshape(buf(1:size(p)), shape(p))
  end subroutine ref_put_prim

  subroutine ref_get_cons(buf) bind(C, name='ref_get_cons')
    real(c_double), intent(out) :: buf(*)
    integer :: nt
    nt = size(up)
    buf(1:nt) = reshape(up, (/ nt /))
  end subroutine ref_get_cons

  subroutine ref_put_cons(buf) bind(C, name='ref_put_cons')
    real(c_double), intent(in) :: buf(*)
    up = reshape(buf(1:size(up)), shape(up))
    u = up
  end subroutine ref_put_cons

  subroutine ref_get_rhs(buf) bind(C, name='ref_get_rhs')
    real(c_double), intent(out) :: buf(*)
    integer :: nt
    nt = size(rhs)
    n_steps = max(n_steps, 1)
    call calc_rhs(p, rhs)
    buf(1:nt) = reshape(rhs, (/ nt /))
  end subroutine ref_get_rhs

  subroutine ref_boundaries() bind(C, name='ref_boundaries')
    call set_boundary_conditions(p)
  end subroutine ref_boundaries

  ! cons_to_prim over all interior cells: up -> p, up repaired in place.
  subroutine ref_con2prim_all() bind(C, name='ref_con2prim_all')
    integer :: i, j, k
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          call cons_to_prim(up(:,i,j,k), p(:,i,j,k), i, j, k)
        end do
      end do
    end do
  end subroutine ref_con2prim_all

  real(c_double) function ref_compute_dt() bind(C, name='ref_compute_dt')
    call update_adaptive_dt()
    ref_compute_dt = dt
  end function ref_compute_dt

  ! The body of the reference time loop (grhd2.f90:179-265) for at most nsteps
  ! steps, stopping at final_time.  Perturbations, diagnostics and output are
  ! not part of the update and are skipped.  state(1:3) = t, dt, n_steps.
  subroutine ref_step(nsteps, state) bind(C, name='ref_step')
    integer(c_int), value :: nsteps
    real(c_double), intent(out) :: state(3)
    integer :: istep, rk, i, j, k
    do istep = 1, nsteps
      if (.not. (integration_time < final_time)) exit
      call update_adaptive_dt()
      if (integration_time + dt > final_time) dt = final_time - integration_time
      n_steps = n_steps + 1
      integration_time = integration_time + dt
      u = up
      call set_boundary_conditions(p)
      do rk = 1, 3
        call calc_rhs(p, rhs)
        if (rk == 1) then
          up = u + rhs * dt
        else if (rk == 2) then
          up = 0.75d0 * u + 0.25d0 * (up + rhs * dt)
        else
          up = (u + 2.0d0*(up + rhs * dt)) / 3.0d0
        end if
        !$OMP PARALLEL DO PRIVATE(i, j, k)
        do k = 1, nz
          do j = 1, ny
            do i = 1, nx
              call cons_to_prim(up(:,i,j,k), p(:,i,j,k), i, j, k)
            end do
          end do
        end do
        !$OMP END PARALLEL DO
        call set_boundary_conditions(p)
      end do
    end do
    state(1) = integration_time
    state(2) = dt
    state(3) = dble(n_steps)
  end subroutine ref_step

  subroutine ref_set_time(t_now, t_final) bind(C, name='ref_set_time')
    real(c_double), value :: t_now, t_final
    integration_time = t_now
    final_time = t_final
  end subroutine ref_set_time

  ! ---------------------------------------------------------------------------
  ! Point probes.  ref_configure sets the globals the physics routines read
  ! (the same ones test_hllc_fluxes.f90:11-15 sets) without allocating a grid.
  ! metric: 0 Minkowski, 1 Eddington-Finkelstein, 2 Kerr-Schild
  ! geometry: 0 Cartesian, 1 Cylindrical, 2 Spherical, 3 Spheroidal
  ! ---------------------------------------------------------------------------
  subroutine ref_configure(metric_id, geom_id, log_r, gam, mass, spin, rec_id, rs_id, &
                           lim_id, sensor, n_y, n_ghost) bind(C, name='ref_configure')
    integer(c_int), value :: metric_id, geom_id, log_r, rec_id, rs_id, lim_id, sensor, n_y, n_ghost
    real(c_double), value :: gam, mass, spin
    call release_problem()
    call reset_globals()
    select case(metric_id)
    case(0); metric_type = 'Minkowski'
    case(1); metric_type = 'Eddington-Finkelstein'
    case default; metric_type = 'Kerr-Schild'
    end select
    select case(geom_id)
    case(0); geom_type = 'Cartesian'
    case(1); geom_type = 'Cylindrical'
    case(2); geom_type = 'Spherical'
    case default; geom_type = 'Spheroidal'
    end select
    use_log_r = (log_r /= 0)
    adb_idx = gam
    g1 = adb_idx / (adb_idx - 1.0d0)
    bh_mass = mass
    a_spin = spin
    rho_floor = 1.0d-10
    p_floor = rho_floor * 1.0d-3
    D_floor = rho_floor
    tau_floor = p_floor * (g1 - 1.0d0)
    rec_method_id = rec_id
    riemann_solver_id = rs_id
    tvd_limiter_id = lim_id
    use_shock_sensor = (sensor /= 0)
    ny = n_y
    nghost = n_ghost
    n_steps = 1
    call set_metric_type()
    call init_wavespeed_solver()
  end subroutine ref_configure

  ! out(1:64): alpha, beta(3), gamma(3,3), det, gmunu(0:3,0:3), dlna(0:3), ... packed:
  ! [1] alpha [2:4] beta [5:13] gamma (col-major) [14] det [15:30] gmunu (col-major)
  ! [31:34] dlnalpha.  chris(0:3,0:3,0:3) -> chris_out(64); dg(0:3,0:3,1:3) -> dg_out(48)
  subroutine ref_metric_point(xp, yp, out, chris_out, dg_out) bind(C, name='ref_metric_point')
    real(c_double), value :: xp, yp
    real(c_double), intent(out) :: out(34), chris_out(64), dg_out(48)
    real*8 :: a, b(3), g(3,3), det, gm(0:3,0:3), dl(0:3), ch(0:3,0:3,0:3), dgl(0:3,0:3,1:3)
    call calculate_metric(xp, yp, alpha=a, beta=b, gamma=g, gmunu=gm, det=det, dlnalpha=dl)
    call calculate_christoffel_symbols(xp, yp, ch)
    call calculate_metric_derivatives(xp, yp, dgl)
    out(1) = a
    out(2:4) = b
    out(5:13) = reshape(g, (/ 9 /))
    out(14) = det
    out(15:30) = reshape(gm, (/ 16 /))
    out(31:34) = dl
    chris_out = reshape(ch, (/ 64 /))
    dg_out = reshape(dgl, (/ 48 /))
  end subroutine ref_metric_point

  subroutine ref_riemann(qL, qR, dir, alpha, beta, g, sqg, flux) bind(C, name='ref_riemann')
    real(c_double), intent(in) :: qL(5), qR(5), beta(3), g(9)
    integer(c_int), value :: dir
    real(c_double), value :: alpha, sqg
    real(c_double), intent(out) :: flux(5)
    real*8 :: gm(3,3)
    gm = reshape(g, (/ 3, 3 /))
    call resolve_riemann_problem(qL, qR, int(dir), alpha, beta, gm, sqg, flux)
  end subroutine ref_riemann

  ! prim_1d(5, -ng:nodes+ng) -> qL, qR (5, 0:nodes)
  subroutine ref_reconstruct(prim, nodes, dir, qL, qR) bind(C, name='ref_reconstruct')
    integer(c_int), value :: nodes, dir
    real(c_double), intent(in) :: prim(5, -nghost:nodes+nghost)
    real(c_double), intent(out) :: qL(5, 0:nodes), qR(5, 0:nodes)
    call reconstruct_1d_core(prim, int(nodes), qL, qR, int(dir))
  end subroutine ref_reconstruct

  subroutine ref_physical_flux(q, dir, alpha, beta, g, sqg, flux) bind(C, name='ref_physical_flux')
    real(c_double), intent(in) :: q(5), beta(3), g(9)
    integer(c_int), value :: dir
    real(c_double), value :: alpha, sqg
    real(c_double), intent(out) :: flux(5)
    real*8 :: gm(3,3)
    gm = reshape(g, (/ 3, 3 /))
    call calc_fluxes(q, flux, int(dir), alpha, beta, gm, sqg)
  end subroutine ref_physical_flux

  ! Point con2prim and sources need the centre cache of a live problem: they
  ! act on interior cell (i,j,k) of the problem set up by ref_init.
  subroutine ref_cons_to_prim_cell(cons, prim, i, j, k) bind(C, name='ref_cons_to_prim_cell')
    real(c_double), intent(inout) :: cons(5)
    real(c_double), intent(out) :: prim(5)
    integer(c_int), value :: i, j, k
    call cons_to_prim(cons, prim, int(i), int(j), int(k))
  end subroutine ref_cons_to_prim_cell

  subroutine ref_prim_to_cons_cell(prim, cons, i, j, k) bind(C, name='ref_prim_to_cons_cell')
    real(c_double), intent(in) :: prim(5)
    real(c_double), intent(out) :: cons(5)
    integer(c_int), value :: i, j, k
    call prim_to_cons(prim, cons, int(i), int(j), int(k))
  end subroutine ref_prim_to_cons_cell

  subroutine ref_sources_cell(prim, src, i, j, k) bind(C, name='ref_sources_cell')
    real(c_double), intent(in) :: prim(5)
    real(c_double), intent(out) :: src(5)
    integer(c_int), value :: i, j, k
    call calc_sources(prim, src, int(i), int(j), int(k))
  end subroutine ref_sources_cell

end module grhd_ref_harness
