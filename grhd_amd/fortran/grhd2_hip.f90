! =============================================================================
! grhd2_hip -- driver of the GRHD2 update on one MI355X.
!
!   grhd2_hip [--check] file.par
!
! The reference main program (grhd2.f90:37-393) with its update replaced by
! the HIP library: the .par file is read with the reference's semantics
! (grhd_par), the library builds the problem (grhd_create validates it,
! grhd_initial_conditions sets the initial data and U = P2C(p)), and the time
! loop -- CFL dt, final-time clamp, SSP-RK3 stages with con2prim and boundaries
! -- runs on the device through grhd_advance.  The seeded perturbation is
! injected when t reaches perturbation_time (grhd2.f90:186-227) and the mass
! monitor, reduced on the device, is printed every mass_monitor_stride steps
! (grhd2.f90:282-301).  At the end the convergence norms of the advected wave,
! Michel and dust problems are printed (grhd2.f90:358) and the conserved state
! up(5,nx,ny,nz) is written in the reference's order to
! <output_folder>/<output_prefix>_final_u.bin (raw float64).
! =============================================================================
program grhd2_hip
  use grhd_c
  use grhd_par
  use grhd_ckpt
  use grhd_vtk
  implicit none

  type(grhd_config) :: cfg
  type(run_control) :: rc
  type(c_ptr) :: ctx
  integer(c_int) :: st
  integer(c_int32_t) :: dims(4), done, chunk, n_adv
  integer(c_int64_t) :: counters(4)
  character(len=1024) :: arg1, parfile
  logical :: check_only, perturbed
  real(c_double) :: t, last_dt, dt, mass, norms(3)
  integer(c_int32_t) :: imax
  real(c_double), allocatable :: u(:,:,:,:), u_restart(:,:,:,:)
  type(checkpoint_state) :: cks
  real(c_double) :: next_checkpoint, next_save
  integer :: nargs, nx, ny, nz, unit_out
  integer(c_int64_t) :: nsteps
  real(c_double) :: wall1
  integer(8) :: c0, c1, crate

  nargs = command_argument_count()
  check_only = .false.
  if (nargs == 1) then
    call get_command_argument(1, parfile)
  else if (nargs == 2) then
    call get_command_argument(1, arg1)
    call get_command_argument(2, parfile)
    if (trim(arg1) /= '--check') then
      write(*,*) 'usage: grhd2_hip [--check] file.par'
      error stop 1
    end if
    check_only = .true.
  else
    write(*,*) 'usage: grhd2_hip [--check] file.par'
    error stop 1
  end if

  call read_par(trim(parfile), cfg, rc)
  if (rc%restart) then
    ! restart branch of grhd2.f90:63-117: the problem, its parameters and
    ! state come from the checkpoint, the .par may change the run control only
    write(*,'(A,A)') '>>> restarting from checkpoint ', trim(rc%restart_file)
    call read_checkpoint(rc%restart_file, cfg, rc, cks, u_restart)
    call read_restart_par(trim(parfile), cfg, rc)
  end if

  st = grhd_validate(cfg)
  if (st /= GRHD_OK) then
    write(*,'(A,1X,A)') 'ERROR in '//trim(parfile)//':', trim(grhd_error_message(c_null_ptr))
    error stop 1
  end if
  write(*,'(A,A)') 'problem: ', trim(problem_label(cfg%problem))
  write(*,'(A,3(I0,1X))') 'grid nx ny nz: ', cfg%nx, cfg%ny, cfg%nz
  write(*,'(A,ES12.5,A,ES12.5)') 'final_time: ', cfg%final_time, '  cfl: ', cfg%cfl
  if (check_only) then
    if (rc%restart) then
      write(*,'(A,ES22.14,A,I0)') 'checkpoint at t = ', cks%t, '  step ', cks%n_steps
      write(*,'(A,L1,A,L1,A,I0)') 'observers: extract_mdot ', rc%extract_mdot, '  ppi_diagnostics ', &
        rc%ppi_diagnostics, '  diagnostic_stride ', rc%diagnostic_stride
      write(*,'(A)') 'parameter file and checkpoint valid; simulation not started.'
    else
      write(*,'(A)') 'parameter file valid; simulation not started.'
    end if
    stop
  end if

  st = grhd_create(cfg, ctx)
  if (st /= GRHD_OK) then
    write(*,'(A,1X,A)') 'ERROR in '//trim(parfile)//':', trim(grhd_error_message(c_null_ptr))
    error stop 1
  end if
  call require(grhd_dims(ctx, dims), 'grhd_dims')
  nx = dims(1); ny = dims(2); nz = dims(3)

  ! the initial conditions also build the case's boundary data (the Michel
  ! injector, prepare_case_boundary_data); a restart then replaces the state:
  ! U from the checkpoint, cons_to_prim and the boundaries (grhd2.f90:95-115)
  call require(grhd_initial_conditions(ctx), 'grhd_initial_conditions')
  if (rc%restart) call require(grhd_set_conserved(ctx, u_restart), 'grhd_set_conserved')
  allocate(u(5, nx, ny, nz))
  call require(grhd_total_mass(ctx, mass), 'grhd_total_mass')
  write(*,'(A,ES22.14)') 'Initial mass in domain: ', mass
  call require(grhd_compute_dt(ctx, dt), 'grhd_compute_dt')
  write(*,'(A,ES22.14)') 'dt (CFL, metric light cone): ', dt

  t = cks%t
  nsteps = cks%n_steps
  ! observers at the start (grhd2.f90:139-145)
  call execute_command_line('mkdir -p '//trim(rc%output_folder)//'/'//trim(scheme_name_of(cfg)))
  ! the cold start's frame of the initial state (grhd2.f90:136)
  if (.not. rc%restart) call write_vtk_frame(ctx, cfg, rc, t)
  call observers(rc%extract_mdot, rc%ppi_diagnostics)
  ! frames on the absolute calendar of save_interval, also after a restart
  ! (grhd2.f90:150)
  next_save = dble(int(t / rc%save_interval) + 1) * rc%save_interval
  perturbed = (rc%perturbation == 0) .or. cks%perturbation_applied
  ! checkpoints every checkpoint_interval of time (grhd2.f90:117,123,338-341)
  if (rc%restart) then
    next_checkpoint = dble(int(t / rc%checkpoint_interval) + 1) * rc%checkpoint_interval
  else
    next_checkpoint = rc%checkpoint_interval
  end if
  call execute_command_line('mkdir -p '//trim(rc%output_folder))
  chunk = int(max(1, rc%mass_monitor_stride), c_int32_t)
  if (rc%mass_monitor_stride <= 0) chunk = huge(chunk)
  call system_clock(c0, crate)
  do while (t < cfg%final_time)
    if (.not. perturbed .and. t >= rc%perturbation_time) then
      call require(grhd_perturb(ctx, int(rc%perturbation, c_int32_t), rc%perturbation_amplitude, &
                                int(rc%perturbation_seed, c_int64_t)), 'grhd_perturb')
      perturbed = .true.
      write(*,'(A,ES14.6)') '>>> perturbation injected at t = ', t
      call observers(.false., rc%ppi_diagnostics)   ! grhd2.f90:222-225
    end if
    ! up to the next multiple of the monitor stride (grhd2.f90:282-301)
    n_adv = chunk
    if (rc%mass_monitor_stride > 0) &
      n_adv = int(rc%mass_monitor_stride - mod(nsteps, int(rc%mass_monitor_stride, 8)), c_int32_t)
    if (rc%extract_mdot) n_adv = min(n_adv, int(rc%extraction_stride - &
                                     mod(nsteps, int(rc%extraction_stride, 8)), c_int32_t))
    if (rc%ppi_diagnostics) n_adv = min(n_adv, int(rc%diagnostic_stride - &
                                        mod(nsteps, int(rc%diagnostic_stride, 8)), c_int32_t))
    if (.not. perturbed) n_adv = min(n_adv, steps_until(rc%perturbation_time))
    n_adv = min(n_adv, steps_until(next_save))
    n_adv = max(1_c_int32_t, min(n_adv, steps_until(next_checkpoint)))
    call require(grhd_advance(ctx, cfg%final_time, n_adv, t, done, last_dt), 'grhd_advance')
    nsteps = nsteps + done
    if (done == 0) exit
    ! observers every extraction_stride / diagnostic_stride steps (grhd2.f90:266-280)
    call observers(rc%extract_mdot .and. mod(nsteps, int(rc%extraction_stride, 8)) == 0, &
                   rc%ppi_diagnostics .and. mod(nsteps, int(rc%diagnostic_stride, 8)) == 0)
    if (rc%mass_monitor_stride > 0 .and. mod(nsteps, int(rc%mass_monitor_stride, 8)) == 0) then
      call require(grhd_total_mass(ctx, mass), 'grhd_total_mass')
      write(*,'(A,I0,A,ES14.6,A,ES22.14)') 'step ', nsteps, '  t = ', t, '  total mass: ', mass
    end if
    ! then the frame and the checkpoint (grhd2.f90:333-341)
    if (t >= next_save) then
      call write_vtk_frame(ctx, cfg, rc, next_save)
      next_save = next_save + rc%save_interval
    end if
    if (t >= next_checkpoint) then
      call require(grhd_get_conserved(ctx, u), 'grhd_get_conserved')
      cks%t = t
      cks%n_steps = int(nsteps)
      cks%perturbation_applied = perturbed .and. rc%perturbation /= 0
      call write_checkpoint(checkpoint_file_name(cfg, rc, int(nsteps)), cfg, rc, cks, u)
      write(*,'(A,A)') '>>> checkpoint ', trim(checkpoint_file_name(cfg, rc, int(nsteps)))
      next_checkpoint = next_checkpoint + rc%checkpoint_interval
    end if
  end do
  ! a last observer sample at the final time when the strides missed it
  ! (grhd2.f90:344-354)
  call observers(rc%extract_mdot .and. mod(nsteps, int(rc%extraction_stride, 8)) /= 0, &
                 rc%ppi_diagnostics .and. mod(nsteps, int(rc%diagnostic_stride, 8)) /= 0)
  call require(grhd_synchronize(ctx), 'grhd_synchronize')
  call system_clock(c1)
  wall1 = dble(c1 - c0) / dble(crate)

  call require(grhd_get_conserved(ctx, u), 'grhd_get_conserved')
  call require(grhd_counters(ctx, counters), 'grhd_counters')
  write(*,'(A)') '========================================================='
  write(*,'(A,F14.6)') ' finished at t = ', t
  write(*,'(A,I0)') ' time steps: ', nsteps
  write(*,'(A,F12.3,A)') ' wall time: ', wall1, ' s'
  if (wall1 > 0.0d0) write(*,'(A,ES12.4)') ' cell-updates/s: ', &
      dble(nx) * dble(ny) * dble(nz) * dble(nsteps) / wall1
  write(*,'(A,I0)') ' cells reset to atmosphere by con2prim: ', counters(1)
  write(*,'(A)') '========================================================='
  ! calc_convergence_norms (grhd2.f90:358): advected wave, Michel and dust only
  if (cfg%problem == 6 .or. cfg%problem == 7 .or. cfg%problem == 8) then
    call require(grhd_convergence_norms(ctx, t, norms, imax), 'grhd_convergence_norms')
    write(*,'(A,I0,A,ES20.12,A,ES20.12,A,ES20.12,A,I0)') ' N = ', nx, ' | L1 = ', norms(1), &
      ' | L2 = ', norms(2), ' | Linf = ', norms(3), ' at cell ', imax
  end if

  open(newunit=unit_out, file=trim(rc%output_folder)//'/'//trim(rc%output_prefix)//'_final_u.bin', &
       access='stream', form='unformatted', status='replace')
  write(unit_out) u
  close(unit_out)
  call require(grhd_destroy(ctx), 'grhd_destroy')

contains

  ! Steps of the reference loop (dt, final clamp, t += dt) from t until
  ! t >= t_event, counted with the same accumulation grhd_advance performs.
  integer(c_int32_t) function steps_until(t_event)
    real(c_double), intent(in) :: t_event
    real(c_double) :: tt, d
    tt = t
    steps_until = 0
    do while (tt < t_event .and. tt < cfg%final_time .and. steps_until < chunk)
      d = dt
      if (tt + d > cfg%final_time) d = cfg%final_time - tt
      tt = tt + d
      steps_until = steps_until + 1
    end do
  end function steps_until

  ! calc_m_dot, calc_ppi_modes and calc_global_diagnostics (output.f90:951-1125)
  ! from the device reductions, appended to <output_folder>/<scheme>/ with the
  ! reference's record formats
  subroutine observers(do_mdot, do_ppi)
    logical, intent(in) :: do_mdot, do_ppi
    character(len=1024) :: dir
    real(c_double) :: mdot, modes(17), glob(8)
    integer(c_int64_t) :: cnt(2)
    integer(c_int32_t) :: probe
    integer :: uo
    dir = trim(rc%output_folder)//'/'//trim(scheme_name_of(cfg))
    ! calc_m_dot returns for flat metrics and M < 0.1 (output.f90:959)
    if (do_mdot .and. cfg%metric /= 0 .and. cfg%bh_mass >= 0.1d0) then
      call require(grhd_accretion_rate(ctx, mdot, probe), 'grhd_accretion_rate')
      open(newunit=uo, file=trim(dir)//'/m_dot.dat', position='append', status='unknown')
      write(uo, '(2(E15.7, 2X))') t, mdot
      close(uo)
    end if
    if (do_ppi) then
      if (nz >= 2) then
        call require(grhd_ppi_modes(ctx, modes), 'grhd_ppi_modes')
        open(newunit=uo, file=trim(dir)//'/ppi_modes.dat', position='append', status='unknown', action='write')
        write(uo, '(*(ES24.16E3,1X))') t, modes
        close(uo)
      end if
      call require(grhd_global_diagnostics(ctx, glob, cnt), 'grhd_global_diagnostics')
      open(newunit=uo, file=trim(dir)//'/global_diagnostics.dat', position='append', status='unknown', &
           action='write')
      write(uo, '(9(ES24.16E3,1X),2(I0,1X))') t, glob, cnt
      close(uo)
    end if
  end subroutine observers

  subroutine require(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what
    if (status /= GRHD_OK) then
      write(*,'(A,A,A,I0,A,A)') 'ERROR: ', what, ' failed (', status, '): ', &
        trim(grhd_error_message(ctx))
      error stop 1
    end if
  end subroutine require

end program grhd2_hip
