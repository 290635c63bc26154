! default_callbacks.F90 -- the default problem callbacks of taxila_gpu.
!
! initialize_walls: no walls (default_initialize_walls.F90:19-33).
! initialize_state: a bubble of -rho_inner densities in -rho_outer densities,
! a square (2-D) / cube (3-D) of half-width -bubble_halfwidth around
! ((N+1)/2, ...) in 1-based coordinates, zero velocity -- the initial state
! of the reference's bubble regression tests
! (tests/bubble_2D/initialize_state.F90:104-203, half-widths 26 in 2-D and
! 10 in 3-D).  With -bubble_radius r > 0 the bubble is a disc / sphere.
subroutine initialize_walls(walls, nx, ny, nz, opts)
  use txg_options
  implicit none
  integer, intent(in) :: nx, ny, nz
  double precision, intent(inout) :: walls(nx * ny * nz)
  type(options_type), intent(in) :: opts
  walls = 0.d0
end subroutine initialize_walls

subroutine initialize_state(rho, u, walls, nx, ny, nz, nc, nd, opts)
  use txg_options
  implicit none
  integer, intent(in) :: nx, ny, nz, nc, nd
  double precision, intent(inout) :: rho(nc * nx * ny * nz), u(nc * nd * nx * ny * nz)
  double precision, intent(in) :: walls(nx * ny * nz)
  type(options_type), intent(in) :: opts
  double precision :: rin(4), rout(4), radius
  integer :: half, nmax, i, j, k, node, lx, rx, ly, ry, lz, rz
  logical :: flag, inside

  rin = 0.d0
  rout = 0.d0
  nmax = nc
  call OptionsGetRealArray(opts, '-rho_inner', rin, nmax, flag)
  nmax = nc
  call OptionsGetRealArray(opts, '-rho_outer', rout, nmax, flag)
  half = merge(26, 10, nd == 2)
  call OptionsGetInt(opts, '-bubble_halfwidth', half, flag)
  radius = 0.d0
  call OptionsGetReal(opts, '-bubble_radius', radius, flag)
  lx = (nx + 1) / 2 - half
  rx = (nx + 1) / 2 + half
  ly = (ny + 1) / 2 - half
  ry = (ny + 1) / 2 + half
  lz = (nz + 1) / 2 - half
  rz = (nz + 1) / 2 + half
  u = 0.d0
  do k = 1, nz
     do j = 1, ny
        do i = 1, nx
           node = (i - 1) + nx * ((j - 1) + ny * (k - 1))
           if (walls(node + 1) /= 0.d0) cycle
           if (radius > 0.d0) then
              inside = dble(i - (nx + 1) / 2)**2 + dble(j - (ny + 1) / 2)**2 + &
                   merge(dble(k - (nz + 1) / 2)**2, 0.d0, nd == 3) <= radius * radius
           else
              inside = i >= lx .and. i <= rx .and. j >= ly .and. j <= ry
              if (nd == 3) inside = inside .and. k >= lz .and. k <= rz
           end if
           if (inside) then
              rho(node * nc + 1:node * nc + nc) = rin(1:nc)
           else
              rho(node * nc + 1:node * nc + nc) = rout(1:nc)
           end if
        end do
     end do
  end do
end subroutine initialize_state
