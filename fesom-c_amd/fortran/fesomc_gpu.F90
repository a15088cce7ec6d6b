! ISO_C_BINDING interface of libfesomc_gpu for Fortran hosts.
!
! The reference hot path is a set of argument-less subroutines acting on the
! module globals of o_MESH / o_ARRAYS / o_PARAM (src/fv_var.F90:303-438),
! called from oce_timestep (src/fv_main.F90:1877-2106).  A Fortran host
! replaces each call site with the matching fcg_* function below; arrays
! keep their Fortran shapes (level index fastest) and 1-based ids, the
! library transposes them into its device layout.  Every function returns
! 0 or a negative FCG_E_* status; fcg_check prints the library's message and
! stops, as the reference's own STOP branches do.
module fesomc_gpu
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int), parameter :: FCG_OK = 0, FCG_E_ARG = -1, FCG_E_MESH = -2, FCG_E_HIP = -3, &
       FCG_E_NAN = -4, FCG_E_SALINITY = -5, FCG_E_TEMP = -6, FCG_E_NOGPU = -7
  integer(c_int), parameter :: FCG_FP64 = 0, FCG_MIXED = 1
  integer(c_int), parameter :: FCG_CLOSURE_D3END = 0, FCG_CLOSURE_MLP = 1

  ! fv_run.dat values after READ_DATA_RUN (src/fv_read_run.F90)
  type, bind(C) :: fcg_params
     real(c_double) :: dt, dt_2D, time_jd0, Dmin, C_d, z0b_min, z0s_min, snul, Cd_bz, Bp_bz, &
          clim_relax, relax2clim_ac
     integer(c_int) :: nsigma, Mt, tracer_adv, muscl_call, mom_adv_2D, mom_adv_3D, vert_adv, &
          im_vert_visc, im_vert_visc_fic, fic_point, wet_dry_on, tf_presence, restart, cl_relax, precision, closure, &
          key_atm, type_swr_body
     real(c_double) :: Rjc, ajc, bjc
     integer(c_int) :: key_rivers, filt_2D, filt_3D
     real(c_double) :: tau_inv_filt
     integer(c_int) :: filt_bi_2D, filt_bi_3D
     real(c_double) :: tau_c
     integer(c_int) :: bih_2D, bih_3D
     real(c_double) :: Abh0
  end type fcg_params

  ! the o_MESH arrays after the reference's preprocessing (pointers into host memory)
  type, bind(C) :: fcg_mesh
     integer(c_int) :: nod2D, elem2D, edge2D, edge2D_in, nsigma, nobn
     type(c_ptr) :: elem2D_nodes, edge_nodes, edge_tri, elem_edges, elem_neighbors, index_nod2D, in_obn
     type(c_ptr) :: depth, sigma, area, elem_area, elem_cos, coriolis, w_cv, gradient_sca, gradient_vec, &
          edge_dxdy, edge_cross_dxdy, ac, a_tp, ampt, fazt, amp_factor, phy_factor, coord_nod2D
     real(c_double) :: density_0, scale_area
  end type fcg_mesh

  type, bind(C) :: fcg_diag
     real(c_double) :: energy, volume, heat, salt, eta_min, eta_max, T_min, T_max, S_min, S_max, time_jd
     integer(c_int) :: n_dt, n_wet_elem
     integer(c_long_long) :: mlp_fallback
     integer(c_int) :: status
  end type fcg_diag

  type, bind(C) :: fcg_mesh_cfg
     integer(c_int) :: cartesian, coriolis_tf, nsigma, ind_sigma, is_LE, T_potential, tf_presence
     real(c_double) :: lat_cartesian_deg, cyclic_length_deg, pow_sig, lev_bot, lev_sur, T_const, S_const
     integer(c_int) :: harmonics(12)
     integer(c_int) :: SL_obn
     real(c_double) :: SL_radius
  end type fcg_mesh_cfg

  interface
     integer(c_int) function fcg_mesh_build(cfg, nod2D, xy, ind, elem2D, elem, depth, nobn, ob, out) &
          bind(C, name="fcg_mesh_build")
       import :: c_int, c_double, c_ptr, fcg_mesh_cfg
       type(fcg_mesh_cfg), intent(in) :: cfg
       integer(c_int), value :: nod2D, elem2D, nobn
       real(c_double), intent(in) :: xy(*), depth(*)
       integer(c_int), intent(in) :: ind(*), elem(*)
       type(c_ptr), value :: ob
       type(c_ptr), intent(out) :: out
     end function fcg_mesh_build

     integer(c_int) function fcg_mesh_view(m, view) bind(C, name="fcg_mesh_view")
       import :: c_int, c_ptr, fcg_mesh
       type(c_ptr), value :: m
       type(fcg_mesh), intent(out) :: view
     end function fcg_mesh_view

     subroutine fcg_mesh_free(m) bind(C, name="fcg_mesh_free")
       import :: c_ptr
       type(c_ptr), value :: m
     end subroutine fcg_mesh_free

     ! array_setup + jacobian + initial_state, device part (src/fv_main.F90:157-168, 320-362)
     integer(c_int) function fcg_create(mesh, par, ctx) bind(C, name="fcg_create")
       import :: c_int, c_ptr, fcg_mesh, fcg_params
       type(fcg_mesh), intent(in) :: mesh
       type(fcg_params), intent(in) :: par
       type(c_ptr), intent(out) :: ctx
     end function fcg_create

     subroutine fcg_destroy(ctx) bind(C, name="fcg_destroy")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine fcg_destroy

     integer(c_int) function fcg_sbc_set_coef(ctx, coef_a, coef_b) bind(C, name="fcg_sbc_set_coef")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: coef_a(*), coef_b(*)
     end function fcg_sbc_set_coef

     integer(c_int) function fcg_sbc_do(ctx, nm_prec_coef, nm_net_flux) bind(C, name="fcg_sbc_do")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: nm_prec_coef
       integer(c_int), value :: nm_net_flux
     end function fcg_sbc_do

     integer(c_int) function fcg_obc_set_coef(ctx, ndepth, depth, coef_a, coef_b) bind(C, name="fcg_obc_set_coef")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ndepth
       real(c_double), intent(in) :: depth(*), coef_a(*), coef_b(*)
     end function fcg_obc_set_coef

     integer(c_int) function fcg_obc_do(ctx, time_jd) bind(C, name="fcg_obc_do")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time_jd
     end function fcg_obc_do

     integer(c_int) function fcg_set_state(ctx, name, host, count) bind(C, name="fcg_set_state")
       import :: c_int, c_ptr, c_char, c_double, c_long
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), intent(in) :: host(*)
       integer(c_long), value :: count
     end function fcg_set_state

     integer(c_int) function fcg_get_state(ctx, name, host, count) bind(C, name="fcg_get_state")
       import :: c_int, c_ptr, c_char, c_double, c_long
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), intent(inout) :: host(*)
       integer(c_long), value :: count
     end function fcg_get_state

     integer(c_int) function fcg_set_clock(ctx, time_jd, lfirst, n_dt) bind(C, name="fcg_set_clock")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time_jd
       integer(c_int), value :: lfirst, n_dt
     end function fcg_set_clock

     ! one entry per call site of oce_timestep (src/fv_main.F90:400-401, 1996-2076)
     integer(c_int) function fcg_advance_clock(ctx) bind(C, name="fcg_advance_clock")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_advance_clock
     integer(c_int) function fcg_pressure(ctx) bind(C, name="fcg_pressure")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_pressure
     integer(c_int) function fcg_vertical_mixing(ctx) bind(C, name="fcg_vertical_mixing")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_vertical_mixing
     integer(c_int) function fcg_compute_vel_rhs(ctx) bind(C, name="fcg_compute_vel_rhs")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_compute_vel_rhs
     integer(c_int) function fcg_barotropic_timestep(ctx) bind(C, name="fcg_barotropic_timestep")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_barotropic_timestep
     integer(c_int) function fcg_jacobian(ctx) bind(C, name="fcg_jacobian")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_jacobian
     integer(c_int) function fcg_update_3d_vel(ctx) bind(C, name="fcg_update_3d_vel")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_update_3d_vel
     integer(c_int) function fcg_vert_vel_sigma(ctx) bind(C, name="fcg_vert_vel_sigma")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_vert_vel_sigma
     integer(c_int) function fcg_solve_tracers(ctx) bind(C, name="fcg_solve_tracers")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_solve_tracers
     integer(c_int) function fcg_oce_timestep(ctx) bind(C, name="fcg_oce_timestep")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_oce_timestep
     integer(c_int) function fcg_run(ctx, nsteps) bind(C, name="fcg_run")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function fcg_run

     integer(c_int) function fcg_diagnostics(ctx, d) bind(C, name="fcg_diagnostics")
       import :: c_int, c_ptr, fcg_diag
       type(c_ptr), value :: ctx
       type(fcg_diag), intent(out) :: d
     end function fcg_diagnostics
     integer(c_int) function fcg_sync(ctx) bind(C, name="fcg_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fcg_sync
     integer(c_int) function fcg_last_error(ctx, buf, len) bind(C, name="fcg_last_error")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(out) :: buf(*)
       integer(c_int), value :: len
     end function fcg_last_error
     type(c_ptr) function fcg_version() bind(C, name="fcg_version")
       import :: c_ptr
     end function fcg_version

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen

     ! ---- ensembles, learned closure, launch shapes (no reference counterpart)
     integer(c_int) function fcg_create_ensemble(members, nmemb, par, ctx) bind(C, name="fcg_create_ensemble")
       import :: c_int, c_ptr, fcg_mesh, fcg_params
       type(fcg_mesh), intent(in) :: members(*)
       integer(c_int), value :: nmemb
       type(fcg_params), intent(in) :: par
       type(c_ptr), intent(out) :: ctx
     end function fcg_create_ensemble
     integer(c_int) function fcg_diagnostics_member(ctx, member, d) bind(C, name="fcg_diagnostics_member")
       import :: c_int, c_ptr, fcg_diag
       type(c_ptr), value :: ctx
       integer(c_int), value :: member
       type(fcg_diag), intent(out) :: d
     end function fcg_diagnostics_member
     integer(c_int) function fcg_mlp_create(seed, net) bind(C, name="fcg_mlp_create")
       import :: c_int, c_ptr
       integer(c_int), value :: seed
       type(c_ptr), intent(out) :: net
     end function fcg_mlp_create
     subroutine fcg_mlp_destroy(net) bind(C, name="fcg_mlp_destroy")
       import :: c_ptr
       type(c_ptr), value :: net
     end subroutine fcg_mlp_destroy
     integer(c_int) function fcg_mlp_set_params(net, host, count) bind(C, name="fcg_mlp_set_params")
       import :: c_int, c_ptr, c_float, c_long
       type(c_ptr), value :: net
       real(c_float), intent(in) :: host(*)
       integer(c_long), value :: count
     end function fcg_mlp_set_params
     integer(c_int) function fcg_mlp_set_norm(net, x_mean, x_std, y_mean, y_std, x_lo, x_hi) &
         bind(C, name="fcg_mlp_set_norm")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: net
       real(c_float), intent(in) :: x_mean(5), x_std(5), y_mean(2), y_std(2), x_lo(5), x_hi(5)
     end function fcg_mlp_set_norm
     integer(c_int) function fcg_set_mlp(ctx, net) bind(C, name="fcg_set_mlp")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, net
     end function fcg_set_mlp
     integer(c_int) function fcg_hot_start(ctx, time_jd) bind(C, name="fcg_hot_start")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: time_jd
     end function fcg_hot_start
     integer(c_int) function fcg_set_launch(ctx, kernel, block) bind(C, name="fcg_set_launch")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: kernel(*)
       integer(c_int), value :: block
     end function fcg_set_launch
  end interface

contains

  ! NUL-terminated copy of a Fortran string for the name arguments
  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function cstr

  function fcg_version_string() result(s)
    character(len=:), allocatable :: s
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = fcg_version()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: s)
    do i = 1, n
       s(i:i) = chars(i)
    end do
  end function fcg_version_string

  ! Stop with the library's message on a non-zero status; the two tracer
  ! STOPs repeat the reference's text (src/fv_tracer.F90:1470, 1480).
  subroutine fcg_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, ignore
    if (rc == FCG_OK) return
    msg = ''
    if (c_associated(ctx)) then
       ignore = fcg_last_error(ctx, buf, 512)
       do i = 1, 512
          if (buf(i) == c_null_char) exit
          msg(i:i) = buf(i)
       end do
    end if
    select case (rc)
    case (FCG_E_SALINITY)
       write(*,*) 'Negative salinity, time resolution is insufficient, stop'
    case (FCG_E_TEMP)
       write(*,*) 'Negative temperature, time resolution is insufficient, stop '
    case default
       write(*,*) 'fesomc_gpu: ', what, ' failed with status ', rc, ' ', trim(msg)
    end select
    stop 1
  end subroutine fcg_check

end module fesomc_gpu
