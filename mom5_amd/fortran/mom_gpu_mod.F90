!-----------------------------------------------------------------------
! mom_gpu_mod -- ISO_C_BINDING interface to libmom5_gpu (include/mom5_gpu.h)
!
! This is the Fortran side of the drop-in boundary: MOM's own routines keep
! their argument lists and call these bindings with the addresses of
! device-resident copies of their arrays.  Derived-type fields are passed
! as type(c_ptr) device addresses; scalars by value.  Every function
! returns 0 on success; mom_gpu_check turns a non-zero code into a stop
! with the library's message, the counterpart of mpp_error(FATAL, ...).
!
! Paths in the comments are relative to src/mom5/ of mom-ocean/MOM5.
!-----------------------------------------------------------------------
module mom_gpu_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MOM_GPU_OK = 0
  ! ocean_core/ocean_parameters.F90:148-163
  integer(c_int), parameter, public :: MOM_ADVECT_QUICKER    = 5
  integer(c_int), parameter, public :: MOM_ADVECT_MDFL_SWEBY = 9
  integer(c_int), parameter, public :: MOM_ADVECT_DST_LINEAR = 10

  ! ocean_core/ocean_types.F90:922-935 bounds, ocean_domains.F90 cyclic flags
  type, bind(C), public :: mom_gpu_domain_t
     integer(c_int) :: isd, ied, jsd, jed
     integer(c_int) :: isc, iec, jsc, jec
     integer(c_int) :: nk
     integer(c_int) :: cyclic_x, cyclic_y
  end type mom_gpu_domain_t

  ! ocean_core/ocean_types.F90:747-920 (device addresses)
  type, bind(C), public :: mom_gpu_grid_t
     type(c_ptr) :: dxt, dyt, dat, datr
     type(c_ptr) :: dxte, dyte, dxtn, dytn
     type(c_ptr) :: dxu, dyu, dxur, dyur
     type(c_ptr) :: f
     type(c_ptr) :: kmt, kmu
     type(c_ptr) :: tmask, umask
     type(c_ptr) :: dzt
  end type mom_gpu_grid_t

  ! ocean_barotropic_nml (ocean_core/ocean_barotropic.F90:600-630)
  type, bind(C), public :: mom_gpu_bt_params_t
     integer(c_int) :: nts
     real(c_double) :: dtbt
     real(c_double) :: pred_corr_gamma
     real(c_double) :: rho0
     integer(c_int) :: stepping_a
     integer(c_int) :: initsum_with_bar
  end type mom_gpu_bt_params_t

  ! the scheme and fields of mom_gpu_update_ocean_barotropic_variant
  type, bind(C), public :: mom_gpu_bt_variant_t
     integer(c_int) :: cgrid = 0, pressure_based = 0
     real(c_double) :: grav = 9.80d0
     type(c_ptr) :: x_tau = c_null_ptr, x_bar_tau = c_null_ptr, udrho_tau = c_null_ptr
     type(c_ptr) :: forcing_bt = c_null_ptr, thick = c_null_ptr, tmasken = c_null_ptr
     type(c_ptr) :: patm = c_null_ptr, rho_g = c_null_ptr, dpatm_dt = c_null_ptr
     type(c_ptr) :: pme = c_null_ptr, river = c_null_ptr, source = c_null_ptr
     type(c_ptr) :: smooth = c_null_ptr
     type(c_ptr) :: x_bar_taup1 = c_null_ptr, udrho_taup1 = c_null_ptr
     type(c_ptr) :: conv_rho_ud_t = c_null_ptr
  end type mom_gpu_bt_variant_t

  ! the tracer loop of update_ocean_tracer + vert_diffuse_implicit
  ! (ocean_tracers/ocean_tracer.F90:2295-2350,2432): host tables of device
  ! addresses (c_loc of type(c_ptr) arrays), dtype(n) = 8 (FP64) or 4 (FP32);
  ! scheme (c_loc of an integer array, or c_null_ptr: all MDFL_SWEBY) and,
  ! for QUICKER tracers, the tmask_limit ranges and flags
  type, bind(C), public :: mom_gpu_tracer_step_t
     integer(c_int) :: ntr
     type(c_ptr) :: T_taum1, T_taup1, dtype, th_tendency, stf, btf
     integer(c_int) :: index_salt
     type(c_ptr) :: uhrho_et, vhrho_nt, wrho_bt, rho_dzt_tau
     type(c_ptr) :: rho_dzt_taum1, rho_dztr, rho_dzt_taup1
     type(c_ptr) :: dzwt, diff_cbt
     real(c_double) :: rho0, dtime, aidif
     type(c_ptr) :: scheme = c_null_ptr
     type(c_ptr) :: min_limit = c_null_ptr, max_limit = c_null_ptr
     integer(c_int) :: index_temp = -1
     integer(c_int) :: tmask_limit_ts_same = 1
     integer(c_int) :: limit_with_upwind = 0
  end type mom_gpu_tracer_step_t

  public :: mom_gpu_create, mom_gpu_destroy, mom_gpu_sync, mom_gpu_last_error_f
  public :: mom_gpu_malloc, mom_gpu_free, mom_gpu_memcpy_htod, mom_gpu_memcpy_dtoh
  public :: mom_gpu_fill, mom_gpu_set_tuning, mom_gpu_update_domains
  public :: mom_gpu_horz_advect_tracer, mom_gpu_vert_advect_tracer
  public :: mom_gpu_advect_sweby_all, mom_gpu_compute_tmask_limit
  public :: mom_gpu_update_tracer_explicit, mom_gpu_invtri
  public :: mom_gpu_vert_diffuse_implicit, mom_gpu_update_ocean_barotropic
  public :: mom_gpu_tracer_content, mom_gpu_check
  public :: mom_gpu_step_tracers, mom_gpu_tracer_content_f32, mom_gpu_tracer_error
  public :: mom_gpu_copy, mom_gpu_memcpy_dtod, mom_gpu_get_tuning
  public :: mom_gpu_advection_velocity, mom_gpu_barotropic_forcing
  public :: mom_gpu_eta_and_pbot_update, mom_gpu_set_remap_metrics, mom_gpu_barotropic_closure
  public :: mom_gpu_update_ocean_barotropic_variant

  interface
     integer(c_int) function mom_gpu_create(dom, grid, ctx) bind(C, name="mom_gpu_create")
       import :: c_int, c_ptr, mom_gpu_domain_t, mom_gpu_grid_t
       type(mom_gpu_domain_t), intent(in) :: dom
       type(mom_gpu_grid_t), intent(in) :: grid
       type(c_ptr), intent(out) :: ctx
     end function mom_gpu_create

     integer(c_int) function mom_gpu_destroy(ctx) bind(C, name="mom_gpu_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function mom_gpu_destroy

     integer(c_int) function mom_gpu_sync(ctx) bind(C, name="mom_gpu_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function mom_gpu_sync

     type(c_ptr) function mom_gpu_last_error(ctx) bind(C, name="mom_gpu_last_error")
       import :: c_ptr
       type(c_ptr), value :: ctx
     end function mom_gpu_last_error

     integer(c_int) function mom_gpu_malloc(ptr, bytes) bind(C, name="mom_gpu_malloc")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), intent(out) :: ptr
       integer(c_size_t), value :: bytes
     end function mom_gpu_malloc

     integer(c_int) function mom_gpu_free(ptr) bind(C, name="mom_gpu_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: ptr
     end function mom_gpu_free

     integer(c_int) function mom_gpu_memcpy_htod(dst, src, bytes) bind(C, name="mom_gpu_memcpy_htod")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
     end function mom_gpu_memcpy_htod

     integer(c_int) function mom_gpu_memcpy_dtoh(dst, src, bytes) bind(C, name="mom_gpu_memcpy_dtoh")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
     end function mom_gpu_memcpy_dtoh

     integer(c_int) function mom_gpu_fill(ctx, ptr, val, n) bind(C, name="mom_gpu_fill")
       import :: c_int, c_ptr, c_double, c_size_t
       type(c_ptr), value :: ctx, ptr
       real(c_double), value :: val
       integer(c_size_t), value :: n
     end function mom_gpu_fill

     ! mpp_update_domains(field, Dom%domain2d) on one PE, device field of nz levels
     integer(c_int) function mom_gpu_update_domains(ctx, field, nz) &
          bind(C, name="mom_gpu_update_domains")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, field
       integer(c_int), value :: nz
     end function mom_gpu_update_domains

     integer(c_int) function mom_gpu_set_tuning(ctx, knob, val) bind(C, name="mom_gpu_set_tuning")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), dimension(*), intent(in) :: knob
       integer(c_int), value :: val
     end function mom_gpu_set_tuning

     ! horz_advect_tracer, ocean_tracers/ocean_tracer_advect.F90:1898
     integer(c_int) function mom_gpu_horz_advect_tracer(ctx, scheme, T_taum1, T_tau, uhrho_et, &
          vhrho_nt, wrho_bt, rho_dzt_tau, tmask_limit, limit_with_upwind, dtime, th_tendency, &
          wrk1, flux_x, flux_y, flux_z) bind(C, name="mom_gpu_horz_advect_tracer")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: scheme
       type(c_ptr), value :: T_taum1, T_tau, uhrho_et, vhrho_nt, wrho_bt, rho_dzt_tau, tmask_limit
       integer(c_int), value :: limit_with_upwind
       real(c_double), value :: dtime
       type(c_ptr), value :: th_tendency, wrk1, flux_x, flux_y, flux_z
     end function mom_gpu_horz_advect_tracer

     ! vert_advect_tracer, ocean_tracers/ocean_tracer_advect.F90:2095
     integer(c_int) function mom_gpu_vert_advect_tracer(ctx, scheme, T_taum1, T_tau, wrho_bt, &
          tmask_limit, th_tendency, wrk1, flux_z) bind(C, name="mom_gpu_vert_advect_tracer")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: scheme
       type(c_ptr), value :: T_taum1, T_tau, wrho_bt, tmask_limit, th_tendency, wrk1, flux_z
     end function mom_gpu_vert_advect_tracer

     ! advect_tracer_sweby_all, ocean_tracers/ocean_tracer_advect.F90:4104
     integer(c_int) function mom_gpu_advect_sweby_all(ctx, ntr, T_taum1, uhrho_et, vhrho_nt, &
          wrho_bt, rho_dzt_tau, dtime, th_tendency, wrk1) bind(C, name="mom_gpu_advect_sweby_all")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ntr
       type(c_ptr), dimension(*), intent(in) :: T_taum1
       type(c_ptr), value :: uhrho_et, vhrho_nt, wrho_bt, rho_dzt_tau
       real(c_double), value :: dtime
       type(c_ptr), dimension(*), intent(in) :: th_tendency
       type(c_ptr), value :: wrk1
     end function mom_gpu_advect_sweby_all

     ! compute_tmask_limit, ocean_tracers/ocean_tracer.F90:2788
     integer(c_int) function mom_gpu_compute_tmask_limit(ctx, ntr, T_tau, min_limit, max_limit, &
          index_temp, index_salt, ts_same, tmask_limit) bind(C, name="mom_gpu_compute_tmask_limit")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ntr
       type(c_ptr), dimension(*), intent(in) :: T_tau
       real(c_double), dimension(*), intent(in) :: min_limit, max_limit
       integer(c_int), value :: index_temp, index_salt, ts_same
       type(c_ptr), dimension(*), intent(in) :: tmask_limit
     end function mom_gpu_compute_tmask_limit

     ! explicit update, ocean_tracers/ocean_tracer.F90:2341-2350
     integer(c_int) function mom_gpu_update_tracer_explicit(ctx, T_taum1, th_tendency, &
          rho_dzt_taum1, rho_dztr, dtime, T_taup1) bind(C, name="mom_gpu_update_tracer_explicit")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, T_taum1, th_tendency, rho_dzt_taum1, rho_dztr
       real(c_double), value :: dtime
       type(c_ptr), value :: T_taup1
     end function mom_gpu_update_tracer_explicit

     ! invtri, ocean_core/ocean_util.F90:153
     integer(c_int) function mom_gpu_invtri(ctx, z, topbc, botbc, dcb, tdt, kmz, mask, dh, dhw, &
          aidif) bind(C, name="mom_gpu_invtri")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, z, topbc, botbc, dcb
       real(c_double), value :: tdt
       type(c_ptr), value :: kmz, mask, dh, dhw
       real(c_double), value :: aidif
     end function mom_gpu_invtri

     ! vert_diffuse_implicit, ocean_param/vertical/ocean_vert_mix.F90:3426
     integer(c_int) function mom_gpu_vert_diffuse_implicit(ctx, ntr, T_taup1, stf, btf, diff_cbt, &
          K33_implicit, index_salt, rho0, rho_dzt_taup1, dzwt, dtime_t, aidif) &
          bind(C, name="mom_gpu_vert_diffuse_implicit")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: ntr
       type(c_ptr), dimension(*), intent(in) :: T_taup1, stf, btf
       type(c_ptr), value :: diff_cbt, K33_implicit
       integer(c_int), value :: index_salt
       real(c_double), value :: rho0
       type(c_ptr), value :: rho_dzt_taup1, dzwt
       real(c_double), value :: dtime_t, aidif
     end function mom_gpu_vert_diffuse_implicit

     ! update_ocean_barotropic, ocean_core/ocean_barotropic.F90:2763
     integer(c_int) function mom_gpu_update_ocean_barotropic(ctx, p, eta_t_tau, eta_t_bar_tau, &
          udrho_tau, forcing_bt, thicku, patm, pme, river, source, eta_smooth, rho_g, &
          eta_t_bar_taup1, udrho_taup1, conv_rho_ud_t) &
          bind(C, name="mom_gpu_update_ocean_barotropic")
       import :: c_int, c_ptr, mom_gpu_bt_params_t
       type(c_ptr), value :: ctx
       type(mom_gpu_bt_params_t), intent(in) :: p
       type(c_ptr), value :: eta_t_tau, eta_t_bar_tau, udrho_tau, forcing_bt, thicku, patm
       type(c_ptr), value :: pme, river, source, eta_smooth, rho_g
       type(c_ptr), value :: eta_t_bar_taup1, udrho_taup1, conv_rho_ud_t
     end function mom_gpu_update_ocean_barotropic

     ! pred_corr_tropic_{depth,press}_{bgrid,cgrid}, ocean_barotropic.F90:3225-4693
     integer(c_int) function mom_gpu_update_ocean_barotropic_variant(ctx, p, v) &
          bind(C, name="mom_gpu_update_ocean_barotropic_variant")
       import :: c_int, c_ptr, mom_gpu_bt_params_t, mom_gpu_bt_variant_t
       type(c_ptr), value :: ctx
       type(mom_gpu_bt_params_t), intent(in) :: p
       type(mom_gpu_bt_variant_t), intent(in) :: v
     end function mom_gpu_update_ocean_barotropic_variant

     ! tracer_integrals content, ocean_diag/ocean_tracer_diag.F90:2648-2661
     integer(c_int) function mom_gpu_tracer_content(ctx, T, rho_dzt, content) &
          bind(C, name="mom_gpu_tracer_content")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, T, rho_dzt
       real(c_double), intent(out) :: content
     end function mom_gpu_tracer_content

     ! ocean_advection_velocity, ocean_core/ocean_advection_velocity.F90:568-672
     integer(c_int) function mom_gpu_advection_velocity(ctx, u, rho_dzu, pme, river, &
          rho_dzt_tendency, mass_source, uhrho_et, vhrho_nt, wrho_bt, diverge_t) &
          bind(C, name="mom_gpu_advection_velocity")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, u, rho_dzu, pme, river, rho_dzt_tendency, mass_source
       type(c_ptr), value :: uhrho_et, vhrho_nt, wrho_bt, diverge_t
     end function mom_gpu_advection_velocity

     ! ocean_barotropic_forcing, ocean_core/ocean_barotropic.F90:3058-3141
     integer(c_int) function mom_gpu_barotropic_forcing(ctx, accel, smf, bmf, forcing_bt) &
          bind(C, name="mom_gpu_barotropic_forcing")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, accel, smf, bmf, forcing_bt
     end function mom_gpu_barotropic_forcing

     ! eta_and_pbot_tendency + eta_and_pbot_update, ocean_barotropic.F90:2667-2726, 2290-2385
     integer(c_int) function mom_gpu_eta_and_pbot_update(ctx, pressure_based, rho0, grav, &
          dtime, conv_rho_ud_t_tau, pme, river, source, dpatm_dt, ht, x_taum1, x_taup1, &
          pbot_t_taup1, tendency) bind(C, name="mom_gpu_eta_and_pbot_update")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: pressure_based
       real(c_double), value :: rho0, grav, dtime
       type(c_ptr), value :: conv_rho_ud_t_tau, pme, river, source, dpatm_dt, ht
       type(c_ptr), value :: x_taum1, x_taup1, pbot_t_taup1, tendency
     end function mom_gpu_eta_and_pbot_update

     ! Grd%dte, dtw, dus, dun, daur (host arrays) for REMAP_BT_TO_BU
     integer(c_int) function mom_gpu_set_remap_metrics(ctx, dte, dtw, dus, dun, daur) &
          bind(C, name="mom_gpu_set_remap_metrics")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: dte(*), dtw(*), dus(*), dun(*), daur(*)
     end function mom_gpu_set_remap_metrics

     ! closure of update_ocean_barotropic, ocean_barotropic.F90:2906-2921, 2956-2985
     integer(c_int) function mom_gpu_barotropic_closure(ctx, pressure_based, rho0, grav, &
          eta_t_bar, anompb_bar, patm, rho_g, ht, tmasken, eta_u, ps, grad_ps, grad_anompb) &
          bind(C, name="mom_gpu_barotropic_closure")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: pressure_based
       real(c_double), value :: rho0, grav
       type(c_ptr), value :: eta_t_bar, anompb_bar, patm, rho_g, ht, tmasken
       type(c_ptr), value :: eta_u, ps, grad_ps, grad_anompb
     end function mom_gpu_barotropic_closure

     ! the batched step (all MDFL-Sweby tracers, FP64 or FP32 storage each)
     integer(c_int) function mom_gpu_step_tracers(ctx, s) bind(C, name="mom_gpu_step_tracers")
       import :: c_int, c_ptr, mom_gpu_tracer_step_t
       type(c_ptr), value :: ctx
       type(mom_gpu_tracer_step_t), intent(in) :: s
     end function mom_gpu_step_tracers

     integer(c_int) function mom_gpu_tracer_content_f32(ctx, T, rho_dzt, content) &
          bind(C, name="mom_gpu_tracer_content_f32")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, T, rho_dzt
       real(c_double), intent(out) :: content
     end function mom_gpu_tracer_content_f32

     ! error of a tracer against an FP64 run: out(1:7) = max|a-b|, max|b|,
     ! sum (a-b)**2, sum b**2, cells, cells differing, max relative
     integer(c_int) function mom_gpu_tracer_error(ctx, a, dtype, b, rel_floor, out) &
          bind(C, name="mom_gpu_tracer_error")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, a, b
       integer(c_int), value :: dtype
       real(c_double), value :: rel_floor
       real(c_double), intent(out) :: out(7)
     end function mom_gpu_tracer_error

     integer(c_int) function mom_gpu_copy(ctx, dst, src, n) bind(C, name="mom_gpu_copy")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: ctx, dst, src
       integer(c_size_t), value :: n
     end function mom_gpu_copy

     integer(c_int) function mom_gpu_memcpy_dtod(dst, src, bytes) &
          bind(C, name="mom_gpu_memcpy_dtod")
       import :: c_int, c_ptr, c_size_t
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
     end function mom_gpu_memcpy_dtod

     integer(c_int) function mom_gpu_get_tuning(ctx, knob, value) &
          bind(C, name="mom_gpu_get_tuning")
       import :: c_int, c_ptr, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char), intent(in) :: knob(*)
       integer(c_int), intent(out) :: value
     end function mom_gpu_get_tuning

     integer(c_size_t) function c_strlen(s) bind(C, name="strlen")
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  !> Message of the last failing call on ctx (c_null_ptr: context-less calls).
  function mom_gpu_last_error_f(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = mom_gpu_last_error(ctx)
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    enddo
  end function mom_gpu_last_error_f

  !> Stop on a non-zero return code, with the library's message
  !> (the counterpart of mpp_error(FATAL, ...)).
  subroutine mom_gpu_check(rc, ctx, what)
    integer(c_int), intent(in) :: rc
    type(c_ptr), intent(in) :: ctx
    character(len=*), intent(in) :: what
    if (rc /= MOM_GPU_OK) then
       write(*,'(a)') 'FATAL from mom_gpu_mod ('//what//'): '//mom_gpu_last_error_f(ctx)
       error stop 1
    endif
  end subroutine mom_gpu_check

end module mom_gpu_mod
