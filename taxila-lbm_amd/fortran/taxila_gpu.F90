! taxila_gpu.F90 -- the Fortran-90 driver: main.F90 + LBMRun2 of the
! reference (src/lbm/main.F90:19-108, lbm.F90:183-422) for the flow, with
! the step on the MI355X.
!
!   taxila_gpu <options file> [-option value ...]
!
! Sequence: options file (argv 1) and command line -> flow from options ->
! device set-up -> walls (-walls_type 1: -walls_file PETSc Vec; 2: the
! initialize_walls callback) -> initial state (initialize_state callback, or
! -restart -istep k reading <prefix>fi<k/kwrite>.dat) -> init -> npasses
! steps with output every -kwrite steps and the optional steady-state exit.
! With -transport_discretization the passive-scalar species distributions
! step beside the flow (txg_transport.F90); their initial state comes from
! the initialize_state_transport callback (zero on restart, as the
! reference's restart reloads the flow only, lbm.F90:519-544).
!
! Callbacks are linked at build time, as in the reference's test Makefiles:
!   subroutine initialize_walls(walls, nx, ny, nz, opts)
!   subroutine initialize_state(rho, u, walls, nx, ny, nz, nc, nd, opts)
!   subroutine initialize_state_transport(psi, flux, walls, nx, ny, nz, ns, nd, opts)
! with plain owned arrays in the C ABI layouts (walls(node), rho(C*node),
! u(C*D*node), node = x + nx*(y + ny*z)).
program taxila_gpu
  use iso_fortran_env, only: real64, int64
  use txg_options
  use txg_flow
  use txg_transport
  implicit none

  interface
     subroutine initialize_walls(walls, nx, ny, nz, opts)
       use txg_options
       integer, intent(in) :: nx, ny, nz
       double precision, intent(inout) :: walls(nx * ny * nz)
       type(options_type), intent(in) :: opts
     end subroutine initialize_walls
     subroutine initialize_state(rho, u, walls, nx, ny, nz, nc, nd, opts)
       use txg_options
       integer, intent(in) :: nx, ny, nz, nc, nd
       double precision, intent(inout) :: rho(nc * nx * ny * nz), u(nc * nd * nx * ny * nz)
       double precision, intent(in) :: walls(nx * ny * nz)
       type(options_type), intent(in) :: opts
     end subroutine initialize_state
     subroutine initialize_state_transport(psi, flux, walls, nx, ny, nz, ns, nd, opts)
       use txg_options
       integer, intent(in) :: nx, ny, nz, ns, nd
       double precision, intent(inout) :: psi(ns * nx * ny * nz), flux(ns * nd * nx * ny * nz)
       double precision, intent(in) :: walls(nx * ny * nz)
       type(options_type), intent(in) :: opts
     end subroutine initialize_state_transport
  end interface

  type(options_type) :: opts
  type(flow_type) :: flow
  type(transport_type) :: tr
  character(len=512) :: infile, prefix, wfile
  integer :: ierr, npasses, kwrite, istep, walls_type, counter, lcv, n, restart_counter
  logical :: flag, restart, run_ss, supress_ic, output_last
  double precision :: tol, norm
  double precision, allocatable :: walls(:), rho(:), u(:), fi(:)
  integer(int64) :: t0, t1, rate
  character(len=3) :: cnt

  if (command_argument_count() < 1) then
     write(*, *) 'usage: taxila_gpu <options file> [-option value ...]'
     error stop 2
  end if
  call get_command_argument(1, infile)
  call OptionsLoadFile(opts, infile, ierr)
  if (ierr /= 0) then
     write(*, *) 'ERROR: cannot read options file ', trim(infile)
     error stop 2
  end if
  call OptionsLoadArgs(opts, 2)

  ! timestepping and i/o options (lbm_options.F90:180-262)
  npasses = 1
  call OptionsGetInt(opts, '-npasses', npasses, flag)
  kwrite = -1
  call OptionsGetInt(opts, '-kwrite', kwrite, flag)
  istep = 0
  call OptionsGetInt(opts, '-istep', istep, flag)
  restart = .false.
  call OptionsGetBool(opts, '-restart', restart, flag)
  restart_counter = -1
  call OptionsGetInt(opts, '-restart_counter', restart_counter, flag)
  run_ss = .false.
  call OptionsGetBool(opts, '-run_to_steadystate', run_ss, flag)
  tol = 1.d-6
  call OptionsGetReal(opts, '-steadystate_tolerance', tol, flag)
  supress_ic = .false.
  call OptionsGetBool(opts, '-supress_ic_output', supress_ic, flag)
  output_last = .false.
  call OptionsGetBool(opts, '-output_last', output_last, flag)
  prefix = 'test_solution/'
  call OptionsGetString(opts, '-output_file_prefix', prefix, flag)

  call FlowSetFromOptions(flow, opts)
  call TransportSetFromOptions(tr, opts, flow%desc)
  call FlowSetUp(flow)
  if (tr%enabled) call TransportSetUp(tr)

  ! walls (lbm_walls.F90:151-177)
  allocate(walls(flow%nnodes))
  walls = 0.d0
  walls_type = 2
  call OptionsGetInt(opts, '-walls_type', walls_type, flag)
  if (walls_type == 1) then
     wfile = ''
     call OptionsGetString(opts, '-walls_file', wfile, flag)
     block
       double precision, allocatable :: w(:)
       call read_petsc_vec(wfile, w, n)
       if (n /= flow%nnodes) then
          write(*, *) 'ERROR: walls file has ', n, ' values, the grid ', flow%nnodes
          error stop 1
       end if
       walls = w
     end block
  else
     call initialize_walls(walls, flow%nx, flow%ny, flow%nz, opts)
  end if
  call WallsSetValues(flow, walls)
  if (tr%enabled) call TransportSetWalls(tr, walls)

  ! initial state (lbm.F90:444-453, restart lbm.F90:519-544)
  counter = 0
  if (restart) then
     if (restart_counter > -1) then
        counter = restart_counter
     else if (kwrite > 0) then
        counter = istep / kwrite
     else
        write(*, *) 'ERROR: must set -restart_counter to define which file to restart from'
        error stop 1
     end if
     write(cnt, '(I3.3)') counter
     call read_petsc_vec(trim(prefix) // 'fi' // cnt // '.dat', fi, n)
     call FlowLoadFi(flow, fi)
  else
     allocate(rho(flow%nnodes * flow%ncomponents), u(flow%nnodes * flow%ncomponents * flow%ndims))
     rho = 0.d0
     u = 0.d0
     call initialize_state(rho, u, walls, flow%nx, flow%ny, flow%nz, flow%ncomponents, flow%ndims, opts)
     call FlowInitializeState(flow, rho, u)
  end if
  if (tr%enabled) then
     block
       double precision, allocatable :: psi(:), flux(:)
       allocate(psi(tr%nnodes * tr%nspecies), flux(tr%nnodes * tr%nspecies * tr%ndims))
       psi = 0.d0
       flux = 0.d0
       if (.not. restart) &
            call initialize_state_transport(psi, flux, walls, tr%nx, tr%ny, tr%nz, tr%nspecies, tr%ndims, opts)
       call TransportInitializeState(tr, psi, flux)
     end block
  end if
  call FlowInit(flow)
  if (tr%enabled) call TransportFiInit(tr)
  if (.not. restart .and. .not. supress_ic) then
     write(*, '(A,I0,A,I0)') ' outputing step ', istep, ' to file ', counter
     call FlowOutput(flow, prefix, counter)
     if (tr%enabled) call TransportOutputDiagnostics(tr, prefix, counter)
  end if
  counter = counter + 1

  call system_clock(t0, rate)
  lcv = istep
  do lcv = istep + 1, npasses
     call FlowCollision(flow)
     if (tr%enabled) call TransportCollision(tr)
     call FlowStream(flow)
     if (tr%enabled) call TransportStream(tr)
     call FlowBounceback(flow)
     if (tr%enabled) call TransportReactWithWalls(tr)
     call FlowApplyBCs(flow)
     call FlowUpdateFlux(flow)
     if (tr%enabled) call TransportUpdateMoments(tr)
     if (kwrite > 0) then
        if (mod(lcv, kwrite) == 0) then
           write(*, '(A,I0,A,I0)') ' outputing step ', lcv, ' to file ', counter
           call FlowOutput(flow, prefix, counter)
           if (tr%enabled) call TransportOutputDiagnostics(tr, prefix, counter)
           counter = counter + 1
        end if
     end if
     if (run_ss) then
        call FlowDeltaNorm(flow, norm)
        if (norm < tol .and. lcv > istep + 2) then
           write(*, *) 'Exiting steady state solve, norm (tol):', norm, '(', tol, ')'
           exit
        end if
     end if
  end do
  if (output_last) then
     call FlowOutput(flow, prefix, counter)
     if (tr%enabled) call TransportOutputDiagnostics(tr, prefix, counter)
  end if
  call system_clock(t1)
  if (npasses > istep) write(*, '(A,ES12.4,A)') ' Timing: Simulation took ', &
       dble(t1 - t0) / dble(rate) / dble(max(1, npasses - istep)), ' s (per unit timestep)'
  call TransportDestroy(tr)
  call FlowDestroy(flow)
end program taxila_gpu
