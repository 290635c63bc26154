! default_callbacks.F90 -- the default problem callbacks of taxila_gpu.
!
! initialize_state_transport: species densities -psi_inner / -psi_outer
! (per species, default 0) on the same bubble, zero fluxes; with neither
! option it leaves the zero state, as the reference's default
! (default_initialize_state_transport.F90:19-21, initialize_state_zeroconc).
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
  double precision :: rin(4), rout(4)
  integer :: nmax
  logical :: flag

  rin = 0.d0
  rout = 0.d0
  nmax = nc
  call OptionsGetRealArray(opts, '-rho_inner', rin, nmax, flag)
  nmax = nc
  call OptionsGetRealArray(opts, '-rho_outer', rout, nmax, flag)
  u = 0.d0
  call bubble_fill(rho, walls, nx, ny, nz, nc, nd, rin, rout, opts)
end subroutine initialize_state

! a(C,node) = vin inside the bubble, vout outside it, at fluid nodes
subroutine bubble_fill(a, walls, nx, ny, nz, nc, nd, vin, vout, opts)
  use txg_options
  implicit none
  integer, intent(in) :: nx, ny, nz, nc, nd
  double precision, intent(inout) :: a(nc * nx * ny * nz)
  double precision, intent(in) :: walls(nx * ny * nz), vin(4), vout(4)
  type(options_type), intent(in) :: opts
  double precision :: radius
  integer :: half, i, j, k, node, lx, rx, ly, ry, lz, rz
  logical :: flag, inside

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
              a(node * nc + 1:node * nc + nc) = vin(1:nc)
           else
              a(node * nc + 1:node * nc + nc) = vout(1:nc)
           end if
        end do
     end do
  end do
end subroutine bubble_fill

subroutine initialize_state_transport(psi, flux, walls, nx, ny, nz, ns, nd, opts)
  use txg_options
  implicit none
  integer, intent(in) :: nx, ny, nz, ns, nd
  double precision, intent(inout) :: psi(ns * nx * ny * nz), flux(ns * nd * nx * ny * nz)
  double precision, intent(in) :: walls(nx * ny * nz)
  type(options_type), intent(in) :: opts
  double precision :: pin(4), pout(4)
  integer :: nmax
  logical :: f1, f2

  pin = 0.d0
  pout = 0.d0
  nmax = ns
  call OptionsGetRealArray(opts, '-psi_inner', pin, nmax, f1)
  nmax = ns
  call OptionsGetRealArray(opts, '-psi_outer', pout, nmax, f2)
  flux = 0.d0
  if (.not. (f1 .or. f2)) return
  call bubble_fill(psi, walls, nx, ny, nz, ns, nd, pin, pout, opts)
end subroutine initialize_state_transport
