! fesomc_run: Fortran host driver of the GPU internal step.
!
! Follows the reference program's order (src/fv_main.F90): READ_DATA_RUN
! reads fv_run.dat, read_mesh_ser reads the MESHPATH files, the mesh
! preprocessing and array set-up build the context, initial_state sets T/S,
! then the main loop (src/fv_main.F90:389-746) advances the clock and calls
! oce_timestep's call sites one by one, writing energy.dat every IREP steps
! in the reference's record format (src/fv_main.F90:614-619).
!
! usage: fesomc_run RUN_DIR [NSTEPS] [--mesh-only]
!   RUN_DIR holds fv_run.dat; MESHPATH inside it is relative to RUN_DIR.
!   NSTEPS overrides fv_run.dat's nsteps.  --mesh-only stops after the mesh
!   preprocessing and prints its sizes (needs no GPU).
program fesomc_run
  use iso_c_binding
  use fesomc_gpu
  implicit none

  integer, parameter :: maxkeys = 400
  character(len=64) :: keys(maxkeys)
  character(len=256) :: vals(maxkeys)
  integer :: nkeys = 0

  character(len=512) :: rundir, arg, meshdir
  integer :: nsteps, irep, irestart, ini_time, n_dt, i, k, n, nargs, L, iverbosity
  logical :: mesh_only
  type(fcg_mesh_cfg) :: mc
  type(fcg_params) :: par
  type(fcg_mesh) :: view
  type(fcg_diag) :: dg
  type(c_ptr) :: mobj, ctx
  integer(c_int) :: rc
  integer :: nod2D, elem2D, nobn, n_harm
  real(c_double), allocatable, target :: xy(:, :), depth(:), ob(:, :), T(:, :), S(:, :)
  integer(c_int), allocatable :: ind(:), elem(:, :)
  real(c_double) :: time, time_jd
  ! river table (src/fv_rivers.F90): records, mouths, interpolation state
  integer :: ntime_riv = 0, nriv = 0, t_riv = 0, t_riv_p1 = 0
  logical :: one_field_riv = .false.
  real(c_double), allocatable :: time_riv(:), data_riv(:, :), coef_a_riv(:), coef_b_riv(:), runoff_riv(:), &
       runoff_node(:)
  integer, allocatable :: nodes_riv(:)
  character(len=3), parameter :: harm(12) = [character(len=3) :: 'M2', 'S2', 'N2', 'K2', 'K1', 'O1', &
       'P1', 'Q1', 'Mf', 'Mm', 'Ssa', 'M4']

  nargs = command_argument_count()
  if (nargs < 1) then
     write(*,*) 'usage: fesomc_run RUN_DIR [NSTEPS] [--mesh-only]'
     stop 2
  end if
  call get_command_argument(1, rundir)
  nsteps = -1
  mesh_only = .false.
  do i = 2, nargs
     call get_command_argument(i, arg)
     if (trim(arg) == '--mesh-only') then
        mesh_only = .true.
     else
        read(arg, *) nsteps
     end if
  end do

  ! ---- READ_DATA_RUN (src/fv_read_run.F90) -------------------------------
  call read_run(trim(rundir) // '/fv_run.dat')
  if (nsteps < 0) nsteps = geti('nsteps', 0)
  irep = geti('IREP', 100)
  irestart = geti('IRESTART', huge(0))
  iverbosity = geti('VERBOSITY', 1)   ! src/fv_read_run.F90:149-155
  call refuse_unsupported()

  mc%cartesian = merge(1, 0, getl('cartesian'))
  mc%coriolis_tf = merge(1, 0, getl('Coriolis_TF'))
  mc%nsigma = geti('nsigma', 10)
  mc%ind_sigma = geti('ind_sigma', 1)
  mc%is_LE = merge(1, 0, trim(gets('TITLE', '')) == 'LE')
  mc%T_potential = merge(1, 0, getl('T_potential'))
  mc%tf_presence = merge(1, 0, getl('TF_presence'))
  mc%lat_cartesian_deg = getf('lat_cartesian', 0.0d0)
  mc%cyclic_length_deg = getf('cyclic_length', 360.0d0)
  mc%pow_sig = getf('pow_sig', 1.0d0)
  mc%lev_bot = getf('lev_bot', 1.0d0)
  mc%lev_sur = getf('lev_sur', 1.0d0)
  mc%T_const = getf('T_const', 10.0d0)
  mc%S_const = getf('S_const', 35.0d0)
  mc%SL_obn = merge(1, 0, getl('SL_obn') .and. .not. getl('SL_obn_user_def'))
  mc%SL_radius = getf('SL_radius', 0.0d0)
  mc%harmonics = 0
  if (getl('TF_presence') .or. getl('T_potential')) call parse_harmonics(gets_line('Harmonics_tf'), mc%harmonics)
  n_harm = sum(mc%harmonics)

  par%dt = getf('dt', 0.0d0)
  par%dt_2D = getf('dt_2D', 0.0d0)
  par%time_jd0 = getf('JD0', 0.0d0)
  par%Dmin = getf('Dmin', 0.25d0)
  par%C_d = getf('Cd', 0.0d0)
  par%z0b_min = getf('z0b_min', 0.01d0)
  par%z0s_min = getf('z0s_min', 0.02d0)
  par%snul = getf('snul', 7.5d-4)
  par%Cd_bz = getf('Cd_bz', 0.0d0)
  par%Bp_bz = getf('Bp_bz', 0.0d0)
  par%clim_relax = 1.0d0 / (getf('clim_relax', 15.0d0) * 3600.0d0 * 24.0d0)       ! fv_read_run.F90:884
  par%relax2clim_ac = 1.0d0 / (getf('relax2clim_ac', 60.0d0) * 3600.0d0 * 24.0d0)
  par%nsigma = mc%nsigma
  if (has('Mt')) then
     par%Mt = geti('Mt', 1)
  else
     par%Mt = int(par%dt / par%dt_2D)                                                ! fv_read_run.F90:1060
  end if
  par%tracer_adv = geti('tracer_adv', 1)
  par%muscl_call = merge(1, 0, getl('MUSCL_CALL'))
  par%mom_adv_2D = geti('mom_adv_2D', 2)
  par%mom_adv_3D = geti('mom_adv_3D', 1)
  par%vert_adv = geti('vert_adv', 1)
  par%im_vert_visc = merge(1, 0, getl('im_vert_visc'))
  par%im_vert_visc_fic = merge(1, 0, getl('im_vert_visc_fic', .true.))
  par%fic_point = geti('fic_point', 1)
  par%wet_dry_on = merge(1, 0, getl('WET_DRY_ON'))
  par%tf_presence = mc%tf_presence
  par%restart = merge(1, 0, trim(gets('RESTART', 'cold_start')) == 'hot_start')
  par%cl_relax = merge(1, 0, getl('cl_relax'))
  par%precision = FCG_FP64
  par%closure = FCG_CLOSURE_D3END
  par%key_atm = merge(1, 0, getl('key_atm'))
  par%type_swr_body = geti('type_swr_body', 2)
  par%Rjc = getf('R', 0.58d0)
  par%ajc = getf('a', 0.35d0)
  par%bjc = getf('b', 23.0d0)
  par%key_rivers = merge(1, 0, getl('key_rivers'))
  par%filt_2D = merge(1, 0, getl('filt_2D'))
  par%filt_3D = merge(1, 0, getl('filt_3D'))
  par%tau_inv_filt = getf('tau_inv_filt', 0.0d0)
  par%filt_bi_2D = merge(1, 0, getl('filt_bi_2D'))
  par%filt_bi_3D = merge(1, 0, getl('filt_bi_3D'))
  par%tau_c = getf('tau_c', 0.4d0)
  par%bih_2D = merge(1, 0, getl('bih_2D'))
  par%bih_3D = merge(1, 0, getl('bih_3D'))
  par%Abh0 = getf('Abh0', 0.0d0)

  ! ---- read_mesh_ser (src/fv_main.F90:787-945) -----------------------------
  meshdir = trim(rundir) // '/' // trim(gets('MESHPATH', './MESH/'))
  call read_nod2d(trim(meshdir) // '/nod2d.out')
  call read_elem2d(trim(meshdir) // '/elem2d.out')
  allocate(depth(nod2D))
  open(21, file=trim(meshdir) // '/depth.out', status='old', action='read')
  read(21, *) depth
  close(21)
  nobn = 0
  if (mc%tf_presence == 1 .and. n_harm > 0) call read_m2(trim(meshdir) // '/m2.out')
  if (nobn == 0) allocate(ob(1, 1))

  rc = fcg_mesh_build(mc, nod2D, xy, ind, elem2D, elem, depth, nobn, &
       merge(c_loc(ob), c_null_ptr, nobn > 0), mobj)
  call fcg_check(rc, c_null_ptr, 'fcg_mesh_build')
  rc = fcg_mesh_view(mobj, view)
  call fcg_check(rc, c_null_ptr, 'fcg_mesh_view')
  write(*,'(a,5i10)') ' mesh: nod2D elem2D edge2D edge2D_in nobn', view%nod2D, view%elem2D, view%edge2D, &
       view%edge2D_in, view%nobn
  if (mesh_only) then
     call fcg_mesh_free(mobj)
     stop
  end if

  rc = fcg_create(view, par, ctx)
  call fcg_check(rc, ctx, 'fcg_create')
  write(*,*) fcg_version_string()

  ! ---- initial_state T/S (src/fv_ini.F90:476-546) --------------------------
  L = mc%nsigma - 1
  allocate(T(L, nod2D), S(L, nod2D))
  T = 0.0d0
  S = 0.0d0
  if (getl('TEMP_ON')) then
     T = mc%T_const
     if (mc%is_LE == 1) then          ! lock exchange: two water masses split at x = 0
        do n = 1, nod2D
           if (xy(1, n) < 0.0d0) then
              T(:, n) = 5.0d0 / 0.28d0
           else
              T(:, n) = 10.0d0 / 0.28d0
           end if
        end do
     end if
  end if
  if (getl('SALINITY_ON')) S = mc%S_const
  call put('TF', T)
  call put('T_old', T)
  call put('Tclim', T)
  call put('SF', S)
  call put('S_old', S)
  call put('Sclim', S)
  time_jd = par%time_jd0                                            ! fv_main.F90:326
  if (par%restart == 1) then          ! read_restart_separate (src/fv_restart.F90:507-613), fv_main.F90:336
     call read_restart('restarts/restart_0.out')
     nsteps = nsteps - ini_time
  end if
  if (par%key_rivers == 1) call rivers_ini(trim(meshdir) // '/rivers.dat')   ! fv_main.F90:352

  ! ---- main loop (src/fv_main.F90:389-746) ---------------------------------
  open(53, file='energy.dat')
  do n_dt = 1, nsteps
     call fcg_check(fcg_advance_clock(ctx), ctx, 'clock')            ! fv_main.F90:400-401
     time_jd = time_jd + par%dt / 86400.0d0                            ! the same clock, kept by the host
     if (par%key_rivers == 1) call rivers_do()                         ! fv_main.F90:412
     call fcg_check(fcg_pressure(ctx), ctx, 'pressure')              ! fv_main.F90:1996
     call fcg_check(fcg_vertical_mixing(ctx), ctx, 'd3_end')         ! fv_main.F90:2006, 2025
     call fcg_check(fcg_compute_vel_rhs(ctx), ctx, 'compute_vel_rhs') ! fv_main.F90:2034
     call fcg_check(fcg_barotropic_timestep(ctx), ctx, 'oce_barotropic_timestep') ! fv_main.F90:2041
     call fcg_check(fcg_jacobian(ctx), ctx, 'jacobian')              ! fv_main.F90:2050
     call fcg_check(fcg_update_3d_vel(ctx), ctx, 'update_3D_vel')    ! fv_main.F90:2057
     call fcg_check(fcg_vert_vel_sigma(ctx), ctx, 'vert_vel_sigma')  ! fv_main.F90:2064
     call fcg_check(fcg_solve_tracers(ctx), ctx, 'solve_tracers')    ! fv_main.F90:2076
     if (mod(n_dt, irep) == 0 .or. n_dt == nsteps) then
        call fcg_check(fcg_diagnostics(ctx, dg), ctx, 'fcg_diagnostics')
        call fcg_check(dg%status, ctx, 'solve_tracers')
        if (mod(n_dt, irep) == 0) then
           time = dg%time_jd * 86400.0d0
           call report(time)
        end if
     end if
     if (mod(n_dt, irestart) == 0) call write_restart(n_dt)           ! fv_main.F90:697
  end do
  close(53, status='KEEP')
  write(*,*) 'volume ', dg%volume, ' heat ', dg%heat, ' salt ', dg%salt
  call fcg_destroy(ctx)
  call fcg_mesh_free(mobj)

contains

  ! The control output of the main loop every IREP steps (src/fv_main.F90:
  ! 607-653, serial branch, type_task 3): the energy.dat record and the stdout
  ! block.  Extrema of eta_n, TF and SF come from the device reduction of
  ! fcg_diagnostics, those of U_n_2D from a host copy.  The reference never
  ! computes the mass-imbalance triple (ib, im_index, ib_mean stay zero) and
  ! prints icedyn_tmask, an ice variable that stays zero without ice.
  subroutine report(time)
    real(c_double), intent(in) :: time
    real(c_double), allocatable :: u2(:, :)
    write(*,*) 'Energy ENERGY ENERGY : ', dg%energy
    write(*,*) 'icedyn_tmask = ', 0.0d0
    write(53, '(4e16.7)') time / 86400.0d0 - par%time_jd0, dg%energy
    if (iverbosity >= 1) then
       write(*,*) 'time= ', time / 86400.0d0 - par%time_jd0, 'time_all= ', time / 86400.0d0 - par%time_jd0
       write(*,*) ' energy= ', dg%energy
       write(*,*) 'Mass Conserv. imbalance: max, index_max, mean', 0.0d0, 0, 0.0d0
       allocate(u2(2, view%elem2D))
       call get2('U_n_2D', u2)
       write(*,*) 'max_min_ssh:', dg%eta_max, dg%eta_min
       write(*,*) 'max_min_vel 2D:', maxval(u2), minval(u2)
       write(*,*) 'T_maxmin= ', dg%T_max, dg%T_min
       write(*,*) 'S_maxmin= ', dg%S_max, dg%S_min
       write(*,*) 'dt', par%dt
       write(*,*) 'dt_2D', par%dt_2D
       deallocate(u2)
    end if
  end subroutine report

  ! ---- restart records (src/fv_restart.F90:474-505): Fortran unformatted
  ! sequential, as write_restart_separate writes them for ver_mix = 3 and
  ! type_task = 3, one record per write statement
  subroutine write_restart(nstep)
    integer, intent(in) :: nstep
    character(len=10) :: dirname
    real(c_double), allocatable :: e0(:), e1(:), e2(:), u2(:, :), u21(:, :), u22(:, :)
    real(c_double), allocatable :: un(:, :), vn(:, :), ua(:, :), va(:, :), w(:, :), bt(:, :), sn(:, :), kv(:, :)
    real(c_double), allocatable :: tf(:, :), sf(:, :), to(:, :), so(:, :), tc(:, :), sc(:, :)
    integer :: ne, nk
    ne = view%elem2D
    nk = mc%nsigma
    allocate(e0(nod2D), e1(nod2D), e2(nod2D), u2(2, ne), u21(2, ne), u22(2, ne))
    allocate(un(nk - 1, ne), vn(nk - 1, ne), ua(nk - 1, ne), va(nk - 1, ne))
    allocate(w(nk, nod2D), bt(nk, nod2D), sn(nk, nod2D), kv(nk, nod2D))
    allocate(tf(nk - 1, nod2D), sf(nk - 1, nod2D), to(nk - 1, nod2D), so(nk - 1, nod2D), tc(nk - 1, nod2D), &
             sc(nk - 1, nod2D))
    call get1('eta_n', e0); call get1('eta_n_1', e1); call get1('eta_n_2', e2)
    call get2('U_n_2D', u2); call get2('U_n_1', u21); call get2('U_n_2', u22)
    call get2('U_n', un); call get2('V_n', vn); call get2('U_rhsAB', ua); call get2('V_rhsAB', va)
    call get2('Wvel', w); call get2('bt', bt); call get2('snu', sn); call get2('Kv', kv)
    call get2('TF', tf); call get2('SF', sf); call get2('T_old', to); call get2('S_old', so)
    call get2('Tclim', tc); call get2('Sclim', sc)
    call fcg_check(fcg_diagnostics(ctx, dg), ctx, 'fcg_diagnostics')
    write(*,*) ' Write Restart at step: ', nstep
    write(dirname, '(I10.10)') nstep
    open(25, file='restarts/' // dirname // '_restart_0.out', form='unformatted')
    write(25) nstep
    write(25) dg%time_jd, par%time_jd0, par%dt, par%dt_2D
    write(25) e0, e1, e2, u2, u21, u22
    write(25) un, vn, ua, va, w
    write(25) bt, sn, kv
    write(25) tf, sf, to, so
    write(25) tc, sc
    close(25)
  end subroutine write_restart

  subroutine read_restart(path)
    character(len=*), intent(in) :: path
    real(c_double) :: tjd, tjd0, rdt, rdt2
    real(c_double), allocatable :: e0(:), e1(:), e2(:), u2(:, :), u21(:, :), u22(:, :)
    real(c_double), allocatable :: un(:, :), vn(:, :), ua(:, :), va(:, :), w(:, :), bt(:, :), sn(:, :), kv(:, :)
    real(c_double), allocatable :: tf(:, :), sf(:, :), to(:, :), so(:, :), tc(:, :), sc(:, :)
    integer :: ne, nk
    ne = view%elem2D
    nk = mc%nsigma
    allocate(e0(nod2D), e1(nod2D), e2(nod2D), u2(2, ne), u21(2, ne), u22(2, ne))
    allocate(un(nk - 1, ne), vn(nk - 1, ne), ua(nk - 1, ne), va(nk - 1, ne))
    allocate(w(nk, nod2D), bt(nk, nod2D), sn(nk, nod2D), kv(nk, nod2D))
    allocate(tf(nk - 1, nod2D), sf(nk - 1, nod2D), to(nk - 1, nod2D), so(nk - 1, nod2D), tc(nk - 1, nod2D), &
             sc(nk - 1, nod2D))
    write(*,*) 'Reading restart'
    open(25, file=path, form='unformatted', status='old', action='read')
    read(25) ini_time
    read(25) tjd, tjd0, rdt, rdt2                                   ! time_jd0 of the file is not used
    read(25) e0, e1, e2, u2, u21, u22
    read(25) un, vn, ua, va, w
    read(25) bt, sn, kv
    read(25) tf, sf, to, so
    read(25) tc, sc
    close(25)
    call put1('eta_n', e0); call put1('eta_n_1', e1); call put1('eta_n_2', e2)
    call put('U_n_2D', u2); call put('U_n_1', u21); call put('U_n_2', u22)
    call put('U_n', un); call put('V_n', vn); call put('U_rhsAB', ua); call put('V_rhsAB', va)
    call put('Wvel', w); call put('bt', bt); call put('snu', sn); call put('Kv', kv)
    call put('TF', tf); call put('SF', sf); call put('T_old', to); call put('S_old', so)
    call put('Tclim', tc); call put('Sclim', sc)
    call fcg_check(fcg_hot_start(ctx, tjd), ctx, 'fcg_hot_start')
    time_jd = tjd
    write(*,*) 'Restart reading finished, time_jd= ', tjd, ',time_jd0=', par%time_jd0
  end subroutine read_restart

  ! rivers_ini (src/fv_rivers.F90:75-214), serial: the table read
  ! list-directed -- Ntime, Nnods, the record times (julian days), the mouth
  ! nodes, then one record of Nnods discharges [m3/s] per time -- and no
  ! discharge until the first rivers_do.
  subroutine rivers_ini(path)
    character(len=*), intent(in) :: path
    integer :: i, iost
    open(27, file=path, status='old', action='read', iostat=iost)
    if (iost /= 0) then
       write(*,*) 'ERROR: --> bad opening file   : ', trim(path), ' ; iostat=', iost
       stop 1
    end if
    read(27, *) ntime_riv
    read(27, *) nriv
    allocate(time_riv(ntime_riv), data_riv(ntime_riv, nriv), nodes_riv(nriv), coef_a_riv(nriv), &
             coef_b_riv(nriv), runoff_riv(nriv), runoff_node(nod2D))
    read(27, *) time_riv
    read(27, *) nodes_riv
    do i = 1, ntime_riv
       read(27, *) data_riv(i, :)
    end do
    close(27)
    if (nriv > 0) then
       if (minval(nodes_riv) < 1 .or. maxval(nodes_riv) > nod2D) then
          write(*,*) 'ERROR: rivers.dat: mouth node outside 1..nod2D'
          stop 1
       end if
    end if
    runoff_riv = 0.0d0
    if (nriv > 0) call rivers_coeff(time_jd)
    write(*,*) 'rivers: ', nriv, ' mouths, ', ntime_riv, ' records'
  end subroutine rivers_ini

  ! getcoeffld (src/fv_rivers.F90:216-261): the straight line through the
  ! records that bracket rdate; past the last record its value, after which
  ! the discharge is no longer updated (one_field).
  subroutine rivers_coeff(rdate)
    real(c_double), intent(in) :: rdate
    real(c_double) :: delta
    t_riv = rivers_search(rdate)
    if (t_riv == 0) then
       write(*,*) 'ERROR: model time ', rdate, ' before the first river record ', time_riv(1)
       stop 1
    end if
    if (t_riv < ntime_riv) then
       t_riv_p1 = t_riv + 1
       delta = time_riv(t_riv_p1) - time_riv(t_riv)
    else
       t_riv_p1 = t_riv
       delta = 1.0d0
       one_field_riv = .true.
    end if
    if (rdate < time_riv(1)) then
       t_riv_p1 = t_riv
       delta = 1.0d0
    end if
    coef_a_riv = (data_riv(t_riv_p1, :) - data_riv(t_riv, :)) / delta
    coef_b_riv = data_riv(t_riv, :) - coef_a_riv * time_riv(t_riv)
  end subroutine rivers_coeff

  ! binarysearch (src/fv_rivers.F90:305-350): the record within 1e-9 (a
  ! REAL(4) constant) of value, else the last one before it, 0 if none
  integer function rivers_search(value) result(ind)
    real(c_double), intent(in) :: value
    real(c_double) :: tol
    integer :: lo, hi, mid
    tol = real(1.0e-9, c_double)
    lo = 1
    hi = ntime_riv
    do while (lo <= hi)
       mid = (lo + hi + 1) / 2
       if (abs(time_riv(mid) - value) <= tol) then
          ind = mid
          return
       else if (time_riv(mid) > value) then
          hi = mid - 1
       else
          lo = mid + 1
       end if
    end do
    ind = hi
  end function rivers_search

  ! rivers_do (src/fv_rivers.F90:263-300) at the advanced clock, ramped over
  ! the first ten days; the per-node sum goes to the step as "runoff"
  subroutine rivers_do()
    integer :: j
    if (nriv == 0) return
    if (.not. one_field_riv) then
       if (time_jd > time_riv(t_riv_p1)) call rivers_coeff(time_jd)
       runoff_riv = time_jd * coef_a_riv + coef_b_riv
       if (time_jd - par%time_jd0 < 10.0d0) runoff_riv = runoff_riv * (time_jd - par%time_jd0) / 10.0d0
    end if
    runoff_node = 0.0d0
    do j = 1, nriv
       runoff_node(nodes_riv(j)) = runoff_node(nodes_riv(j)) + runoff_riv(j)
    end do
    call put1('runoff', runoff_node)
  end subroutine rivers_do

  subroutine get1(name, a)
    character(len=*), intent(in) :: name
    real(c_double), intent(inout) :: a(:)
    call fcg_check(fcg_get_state(ctx, cstr(name), a, int(size(a), c_long)), ctx, 'fcg_get_state ' // name)
  end subroutine get1

  subroutine get2(name, a)
    character(len=*), intent(in) :: name
    real(c_double), intent(inout) :: a(:, :)
    call fcg_check(fcg_get_state(ctx, cstr(name), a, int(size(a), c_long)), ctx, 'fcg_get_state ' // name)
  end subroutine get2

  subroutine put1(name, a)
    character(len=*), intent(in) :: name
    real(c_double), intent(in) :: a(:)
    call fcg_check(fcg_set_state(ctx, cstr(name), a, int(size(a), c_long)), ctx, 'fcg_set_state ' // name)
  end subroutine put1

  subroutine put(name, a)
    character(len=*), intent(in) :: name
    real(c_double), intent(in) :: a(:, :)
    call fcg_check(fcg_set_state(ctx, cstr(name), a, int(size(a), c_long)), ctx, 'fcg_set_state ' // name)
  end subroutine put

  ! fv_run.dat scanner: KEY = value lines, '!' comments, first occurrence wins
  ! (SCAN_FILE semantics, src/fv_utilit.F90:223-570)
  subroutine read_run(path)
    character(len=*), intent(in) :: path
    character(len=512) :: line
    integer :: ios, p, q
    open(20, file=path, status='old', action='read')
    do
       read(20, '(a)', iostat=ios) line
       if (ios /= 0) exit
       p = index(line, '!')
       if (p > 0) line = line(1:p - 1)
       q = index(line, '=')
       if (q == 0) cycle
       if (len_trim(adjustl(line(1:q - 1))) == 0) cycle
       if (has(trim(adjustl(line(1:q - 1))))) cycle
       if (nkeys == maxkeys) exit
       nkeys = nkeys + 1
       keys(nkeys) = trim(adjustl(line(1:q - 1)))
       vals(nkeys) = trim(adjustl(line(q + 1:)))
    end do
    close(20)
  end subroutine read_run

  integer function find(key)
    character(len=*), intent(in) :: key
    integer :: j
    find = 0
    do j = 1, nkeys
       if (trim(keys(j)) == key) then
          find = j
          return
       end if
    end do
  end function find

  logical function has(key)
    character(len=*), intent(in) :: key
    has = find(key) > 0
  end function has

  function gets_line(key) result(v)
    character(len=*), intent(in) :: key
    character(len=256) :: v
    integer :: j
    v = ''
    j = find(key)
    if (j > 0) v = vals(j)
  end function gets_line

  function gets(key, def) result(v)
    character(len=*), intent(in) :: key, def
    character(len=256) :: v
    integer :: j, sp
    v = def
    j = find(key)
    if (j > 0) then
       sp = index(trim(vals(j)), ' ')
       if (sp > 0) then
          v = vals(j)(1:sp - 1)
       else
          v = vals(j)
       end if
    end if
  end function gets

  real(c_double) function getf(key, def)
    character(len=*), intent(in) :: key
    real(c_double), intent(in) :: def
    getf = def
    if (has(key)) read(gets(key, ''), *) getf
  end function getf

  integer function geti(key, def)
    character(len=*), intent(in) :: key
    integer, intent(in) :: def
    geti = def
    if (has(key)) read(gets(key, ''), *) geti
  end function geti

  ! Options whose code this build does not hold stop the run before any set-up
  ! (the same list as the Python host's RunConfig.unsupported).
  subroutine refuse_unsupported()
    character(len=20), parameter :: opts(3) = [character(len=20) :: 'mixed_BP_grad_OB', 'T_potential', &
         'comp_sediment']
    integer :: j
    logical :: bad
    bad = .false.
    do j = 1, size(opts)
       if (getl(trim(opts(j)))) then
          write(*,*) 'ERROR: option not available in this build: ', trim(opts(j))
          bad = .true.
       end if
    end do
    if (geti('Vert_turb', 3) /= 3) then
       write(*,*) 'ERROR: option not available in this build: Vert_turb /= 3'
       bad = .true.
    end if
    if (getl('key_atm') .and. geti('type_swr_body', 2) == 1 .and. .not. getl('jer_const', .true.)) then
       write(*,*) 'ERROR: option not available in this build: jer_const = F (per-node Jerlov coefficients)'
       bad = .true.
    end if
    if (geti('type_swr_body', 2) /= 1 .and. geti('type_swr_body', 2) /= 2) then  ! fv_read_run.F90:799-802
       write(*,*) 'type_swr_body should be 1 or 2'
       bad = .true.
    end if
    if (.not. (getl('SALINITY_ON') .or. getl('TEMP_ON'))) then          ! type_task 3 (fv_read_run.F90)
       write(*,*) 'ERROR: option not available in this build: type_task /= 3 (baroclinic with T or S)'
       bad = .true.
    end if
    if (bad) stop 2
  end subroutine refuse_unsupported

  logical function getl(key, def)
    character(len=*), intent(in) :: key
    logical, intent(in), optional :: def
    character(len=256) :: v
    getl = .false.
    if (present(def)) getl = def
    if (.not. has(key)) return
    v = adjustl(gets(key, ''))
    getl = v(1:1) == 'T' .or. v(1:1) == 't' .or. v(1:2) == '.T' .or. v(1:2) == '.t'
  end function getl

  subroutine parse_harmonics(line, on)
    character(len=*), intent(in) :: line
    integer(c_int), intent(inout) :: on(12)
    character(len=256) :: rest
    character(len=16) :: tok
    integer :: sp, j
    rest = adjustl(line)
    do while (len_trim(rest) > 0)
       sp = index(trim(rest), ' ')
       if (sp == 0) then
          tok = trim(rest)
          rest = ''
       else
          tok = rest(1:sp - 1)
          rest = adjustl(rest(sp + 1:))
       end if
       do j = 1, 12
          if (trim(tok) == trim(harm(j))) on(j) = 1
       end do
    end do
  end subroutine parse_harmonics

  ! nod2d.out: nod2D, then "n x y index" (nodes stored by their id)
  subroutine read_nod2d(path)
    character(len=*), intent(in) :: path
    integer :: j, id, ix
    real(c_double) :: x, y
    open(21, file=path, status='old', action='read')
    read(21, *) nod2D
    allocate(xy(2, nod2D), ind(nod2D))
    do j = 1, nod2D
       read(21, *) id, x, y, ix
       xy(1, id) = x
       xy(2, id) = y
       ind(id) = ix
    end do
    close(21)
  end subroutine read_nod2d

  ! elem2d.out: elem2D, then 4 node ids per element (3 for triangle-only files)
  subroutine read_elem2d(path)
    character(len=*), intent(in) :: path
    character(len=256) :: line
    integer :: j, ncol, p
    logical :: intok
    open(21, file=path, status='old', action='read')
    read(21, *) elem2D
    read(21, '(a)') line
    ncol = 0
    intok = .false.
    do p = 1, len_trim(line)
       if (line(p:p) /= ' ' .and. .not. intok) ncol = ncol + 1
       intok = line(p:p) /= ' '
    end do
    rewind(21)
    read(21, *) elem2D
    allocate(elem(4, elem2D))
    if (ncol >= 4) then
       read(21, *) ((elem(p, j), p = 1, 4), j = 1, elem2D)
    else
       read(21, *) ((elem(p, j), p = 1, 3), j = 1, elem2D)
       elem(4, :) = elem(1, :)
    end if
    close(21)
  end subroutine read_elem2d

  ! m2.out: nobn, then "node amp phase ..." per open-boundary node
  subroutine read_m2(path)
    character(len=*), intent(in) :: path
    integer :: j, id
    open(21, file=path, status='old', action='read')
    read(21, *) nobn
    allocate(ob(2 * n_harm, nobn))
    do j = 1, nobn
       read(21, *) id, ob(:, j)
    end do
    close(21)
  end subroutine read_m2

end program fesomc_run
