! =============================================================================
! grhd_vtk -- VTK frames of the state, in the reference's file format
! (output.f90:154-471): legacy VTK, BINARY, big-endian.  Cell data are the
! five primitives (the radial velocity times r on logarithmic grids), the
! lapse and the radial shift (likewise scaled).  The grid is built from the
! true cell faces: a STRUCTURED_GRID, or, when the domain covers the whole
! sphere (theta in [0, pi], phi over 2 pi, ny > 1, nz > 1), an
! UNSTRUCTURED_GRID with one point per pole and radial face, wedges at the
! poles and hexahedra elsewhere.  Polar positions follow vtk_mapping:
! "physical" places the Kerr-Schild spheroidal coordinates in Cartesian space,
! "untwisted" drops the azimuthal twist (output.f90:743-795).
!
! Frame <output_prefix>_<scheme>_<nint(t / save_interval), 4 digits>.vtk under
! <output_folder>/<scheme>/ (save_vtk_at_time, output.f90:458-471).
! =============================================================================
module grhd_vtk
  use, intrinsic :: iso_c_binding
  use grhd_c
  use grhd_par
  use grhd_ckpt, only: scheme_name_of, ghost_depth_of
  implicit none
  private
  public :: write_vtk_frame

  real(c_double), parameter :: PI = 3.141592653589793d0
  character(len=*), parameter :: NL = char(10)
  character(len=10), parameter :: VAR_NAMES(5) = &
    [character(len=10) :: 'Density', 'Pressure', 'Velocity_X', 'Velocity_Y', 'Velocity_Z']

  ! the grid of the configuration last written (faces, centres, lapse and
  ! shift), kept between frames: the driver runs one configuration per process
  type :: vtk_grid
    integer :: nx = 0, ny = 0, nz = 0
    real(c_double), allocatable :: xf(:), yf(:), zf(:), xc(:), yc(:), zc(:)
    real(c_double), allocatable :: alpha(:,:,:), beta(:,:,:,:)
  end type vtk_grid
  type(vtk_grid), save :: g

contains

  ! faces and centres (allocate_and_grid, initialization.f90:88-206) and the
  ! centre lapse and shift, once per configuration
  subroutine prepare_grid(c)
    type(grhd_config), intent(in) :: c
    real(c_double) :: x0, x1, dx, dy, dz
    integer :: i
    if (g%nx == c%nx .and. g%ny == c%ny .and. g%nz == c%nz .and. allocated(g%xf)) return
    if (allocated(g%xf)) deallocate(g%xf, g%yf, g%zf, g%xc, g%yc, g%zc, g%alpha, g%beta)
    g%nx = c%nx; g%ny = c%ny; g%nz = c%nz
    if (c%log_r /= 0) then
      x0 = log(c%r_min); x1 = log(c%r_max)
    else
      x0 = c%r_min; x1 = c%r_max
    end if
    dx = (x1 - x0) / dble(max(1, c%nx))
    dy = (c%y_max - c%y_min) / dble(max(1, c%ny))
    dz = (c%z_max - c%z_min) / dble(max(1, c%nz))
    allocate(g%xf(0:c%nx), g%yf(0:c%ny), g%zf(0:c%nz), g%xc(c%nx), g%yc(c%ny), g%zc(c%nz))
    do i = 0, c%nx
      g%xf(i) = x0 + dble(i) * dx
    end do
    do i = 1, c%nx
      g%xc(i) = x0 + (dble(i) - 0.5d0) * dx
    end do
    do i = 0, c%ny
      g%yf(i) = c%y_min + dble(i) * dy
    end do
    do i = 1, c%ny
      g%yc(i) = c%y_min + (dble(i) - 0.5d0) * dy
    end do
    do i = 0, c%nz
      g%zf(i) = c%z_min + dble(i) * dz
    end do
    do i = 1, c%nz
      g%zc(i) = c%z_min + (dble(i) - 0.5d0) * dz
    end do
    allocate(g%alpha(c%nx, c%ny, c%nz), g%beta(3, c%nx, c%ny, c%nz))
    if (grhd_lapse_shift(c, g%alpha, g%beta) /= GRHD_OK) then
      write(*,'(A,1X,A)') 'ERROR in grhd_lapse_shift:', trim(grhd_error_message(c_null_ptr))
      error stop 1
    end if
  end subroutine prepare_grid

  ! kerr_schild_spheroidal_position and vtk_polar_position (output.f90:743-795)
  subroutine polar_position(r, theta, phi, spin, untwisted, pos)
    real(c_double), intent(in) :: r, theta, phi, spin
    logical, intent(in) :: untwisted
    real(c_double), intent(out) :: pos(3)
    real(c_double) :: st, ct, sp, cp, rho
    st = sin(theta)
    ct = cos(theta)
    ! sin(pi) is not zero in floating point: both poles collapse exactly
    if (abs(st) <= 64.0d0 * epsilon(1.0d0)) then
      pos = [0.0d0, 0.0d0, sign(r, ct)]
      return
    end if
    if (untwisted) then
      rho = sqrt(r*r + spin*spin) * st
      pos(1) = rho * cos(phi)
      pos(2) = rho * sin(phi)
      pos(3) = r * ct
    else
      sp = sin(phi)
      cp = cos(phi)
      pos(1) = (r*cp - spin*sp) * st
      pos(2) = (r*sp + spin*cp) * st
      pos(3) = r * ct
    end if
  end subroutine polar_position

  pure real(c_double) function radius(c, x)
    type(grhd_config), intent(in) :: c
    real(c_double), intent(in) :: x
    if (c%log_r /= 0) then
      radius = exp(x)
    else
      radius = x
    end if
  end function radius

  ! save_vtk_at_time (output.f90:458-471) of the device state at time t
  subroutine write_vtk_frame(ctx, c, rc, t)
    type(c_ptr), intent(in) :: ctx
    type(grhd_config), intent(in) :: c
    type(run_control), intent(in) :: rc
    real(c_double), intent(in) :: t
    character(len=1024) :: path
    character(len=20) :: scheme
    real(c_double), allocatable :: p(:,:,:,:)
    real(c_double) :: spin, tol
    integer :: frame, ng, u, ios
    logical :: polar, full_sphere

    call prepare_grid(c)
    frame = 0
    if (rc%save_interval > 0.0d0) frame = nint(t / rc%save_interval)
    scheme = scheme_name_of(c)
    write(path, '(A, "/", A, "/", A, "_", A, "_", I4.4, ".vtk")') trim(rc%output_folder), trim(scheme), &
      trim(rc%output_prefix), trim(scheme), frame
    write(*,'(A,A)') '>>> VTK frame ', trim(path)

    ng = ghost_depth_of(c%reconstruction)
    allocate(p(5, -ng:c%nx+ng, -ng:c%ny+ng, -ng:c%nz+ng))
    if (grhd_get_primitives(ctx, p) /= GRHD_OK) then
      write(*,'(A,1X,A)') 'ERROR in grhd_get_primitives:', trim(grhd_error_message(ctx))
      error stop 1
    end if

    spin = 0.0d0
    if (c%metric == 2) spin = c%spin
    tol = 128.0d0 * epsilon(1.0d0)
    polar = c%geometry == 2 .or. c%geometry == 3    ! is_polar_geometry (variables.f90:271-276)
    full_sphere = polar .and. c%ny > 1 .and. c%nz > 1 .and. abs(g%yf(0)) <= tol .and. &
      abs(g%yf(c%ny) - PI) <= tol * max(1.0d0, abs(PI)) .and. &
      abs((g%zf(c%nz) - g%zf(0)) - 2.0d0*PI) <= tol * max(1.0d0, abs(2.0d0*PI))

    open(newunit=u, file=trim(path), status='replace', access='stream', form='unformatted', &
         convert='big_endian', iostat=ios)
    if (ios /= 0) then
      write(*,'(A,A)') 'ERROR: cannot create the VTK frame ', trim(path)
      return
    end if
    write(u) '# vtk DataFile Version 3.0'//NL
    write(u) 'GRHD Simulation Data'//NL
    write(u) 'BINARY'//NL
    if (full_sphere) then
      call unstructured_grid(u, c, spin, rc%vtk_untwisted)
    else
      call structured_grid(u, c, spin, polar, rc%vtk_untwisted)
    end if
    call cell_data(u, c, p)
    close(u)
  end subroutine write_vtk_frame

  ! save_vtk_data (output.f90:153-244): points at the faces of every active
  ! direction, the centre of an inactive one
  subroutine structured_grid(u, c, spin, polar, untwisted)
    integer, intent(in) :: u
    type(grhd_config), intent(in) :: c
    real(c_double), intent(in) :: spin
    logical, intent(in) :: polar, untwisted
    real(c_double), allocatable :: pts(:,:,:,:)
    real(c_double) :: r, th, ph
    integer :: npx, npy, npz, i, j, k
    character(len=100) :: line
    npx = c%nx + 1
    npy = 1
    npz = 1
    if (c%ny > 1) npy = c%ny + 1
    if (c%nz > 1) npz = c%nz + 1
    write(u) 'DATASET STRUCTURED_GRID'//NL
    write(line, '(A, I0, 1X, I0, 1X, I0)') 'DIMENSIONS ', npx, npy, npz
    write(u) trim(line)//NL
    write(line, '(A, I0, A)') 'POINTS ', npx * npy * npz, ' double'
    write(u) trim(line)//NL
    allocate(pts(3, npx, npy, npz))
    do k = 1, npz
      if (c%nz > 1) then
        ph = g%zf(k - 1)
      else
        ph = g%zc(1)
      end if
      do j = 1, npy
        if (c%ny > 1) then
          th = g%yf(j - 1)
        else
          th = g%yc(1)
        end if
        do i = 1, npx
          r = radius(c, g%xf(i - 1))
          if (polar) then
            call polar_position(r, th, ph, spin, untwisted, pts(:, i, j, k))
          else if (c%geometry == 1) then
            pts(1, i, j, k) = r * cos(th)
            pts(2, i, j, k) = r * sin(th)
            pts(3, i, j, k) = ph
          else
            pts(:, i, j, k) = [r, th, ph]
          end if
        end do
      end do
    end do
    write(u) pts
    write(u) NL
  end subroutine structured_grid

  ! save_vtk_polar_unstructured (output.f90:247-398): per radial face a north
  ! point, the rings j = 1..ny-1 for every k, a south point; cells k-major,
  ! then j, then i, with the south wedges' azimuthal order reversed to keep
  ! their orientation
  subroutine unstructured_grid(u, c, spin, untwisted)
    integer, intent(in) :: u
    type(grhd_config), intent(in) :: c
    real(c_double), intent(in) :: spin
    logical, intent(in) :: untwisted
    integer :: nx, ny, nz, npts, ncells, listn, i, j, k, kn, id, cid, o
    integer(c_int32_t), allocatable :: ring(:,:,:), north(:), south(:), cells(:), types(:)
    real(c_double), allocatable :: pts(:,:)
    real(c_double) :: r
    character(len=100) :: line
    nx = c%nx; ny = c%ny; nz = c%nz
    npts = (nx + 1) * ((ny - 1) * nz + 2)
    ncells = nx * ny * nz
    listn = nx * nz * (9 * ny - 4)
    allocate(pts(3, npts), ring(0:nx, 1:ny-1, 0:nz-1), north(0:nx), south(0:nx))
    id = 0
    do i = 0, nx
      r = radius(c, g%xf(i))
      north(i) = id
      call polar_position(r, 0.0d0, g%zf(0), spin, untwisted, pts(:, id + 1))
      id = id + 1
      do k = 0, nz - 1
        do j = 1, ny - 1
          ring(i, j, k) = id
          call polar_position(r, g%yf(j), g%zf(k), spin, untwisted, pts(:, id + 1))
          id = id + 1
        end do
      end do
      south(i) = id
      call polar_position(r, PI, g%zf(0), spin, untwisted, pts(:, id + 1))
      id = id + 1
    end do
    allocate(cells(listn), types(ncells))
    o = 1
    cid = 0
    do k = 0, nz - 1
      kn = mod(k + 1, nz)
      do j = 1, ny
        do i = 0, nx - 1
          cid = cid + 1
          if (j == 1) then
            cells(o:o+6) = [6_c_int32_t, north(i), ring(i, 1, k), ring(i, 1, kn), north(i+1), &
                            ring(i+1, 1, k), ring(i+1, 1, kn)]
            types(cid) = 13            ! VTK_WEDGE
            o = o + 7
          else if (j == ny) then
            cells(o:o+6) = [6_c_int32_t, south(i), ring(i, ny-1, kn), ring(i, ny-1, k), south(i+1), &
                            ring(i+1, ny-1, kn), ring(i+1, ny-1, k)]
            types(cid) = 13
            o = o + 7
          else
            cells(o:o+8) = [8_c_int32_t, ring(i, j-1, k), ring(i+1, j-1, k), ring(i+1, j, k), &
                            ring(i, j, k), ring(i, j-1, kn), ring(i+1, j-1, kn), ring(i+1, j, kn), &
                            ring(i, j, kn)]
            types(cid) = 12            ! VTK_HEXAHEDRON
            o = o + 9
          end if
        end do
      end do
    end do
    write(u) 'DATASET UNSTRUCTURED_GRID'//NL
    write(line, '(A, I0, A)') 'POINTS ', npts, ' double'
    write(u) trim(line)//NL
    write(u) pts
    write(u) NL
    write(line, '(A, I0, 1X, I0)') 'CELLS ', ncells, listn
    write(u) trim(line)//NL
    write(u) cells
    write(u) NL
    write(line, '(A, I0)') 'CELL_TYPES ', ncells
    write(u) trim(line)//NL
    write(u) types
    write(u) NL
  end subroutine unstructured_grid

  ! write_vtk_cell_data (output.f90:401-455)
  subroutine cell_data(u, c, p)
    integer, intent(in) :: u
    type(grhd_config), intent(in) :: c
    real(c_double), intent(in) :: p(:,:,:,:)
    real(c_double), allocatable :: buf(:,:,:)
    integer :: v, i, j, k, o
    character(len=100) :: line
    ! p comes in with lower bounds 1: the interior starts at ng + 1
    o = ghost_depth_of(c%reconstruction)
    write(line, '(A, I0)') 'CELL_DATA ', c%nx * c%ny * c%nz
    write(u) trim(line)//NL
    allocate(buf(c%nx, c%ny, c%nz))
    do v = 1, 5
      write(u) 'SCALARS '//trim(VAR_NAMES(v))//' double 1'//NL
      write(u) 'LOOKUP_TABLE default'//NL
      do k = 1, c%nz
        do j = 1, c%ny
          do i = 1, c%nx
            buf(i, j, k) = p(v, o + 1 + i, o + 1 + j, o + 1 + k)
            if (c%log_r /= 0 .and. v == 3) buf(i, j, k) = buf(i, j, k) * exp(g%xc(i))
          end do
        end do
      end do
      write(u) buf
      write(u) NL
    end do
    write(u) 'SCALARS Lapse_alpha double 1'//NL
    write(u) 'LOOKUP_TABLE default'//NL
    write(u) g%alpha
    write(u) NL
    write(u) 'SCALARS Shift_beta_r double 1'//NL
    write(u) 'LOOKUP_TABLE default'//NL
    do k = 1, c%nz
      do j = 1, c%ny
        do i = 1, c%nx
          buf(i, j, k) = g%beta(1, i, j, k)
          if (c%log_r /= 0) buf(i, j, k) = buf(i, j, k) * exp(g%xc(i))
        end do
      end do
    end do
    write(u) buf
    write(u) NL
  end subroutine cell_data

end module grhd_vtk
