/*
 * mom5_gpu.h -- C-ABI of the MI355X implementation of MOM5's ocean hot path.
 *
 * Each entry point replaces one reference routine of mom-ocean/MOM5 and keeps
 * its argument list, with the derived-type fields (Grd%, Thickness%, Adv_vel%,
 * T_prog(n)%, Ext_mode%) flattened to plain pointers.  Paths in the comments
 * are relative to src/mom5/ of the reference.
 *
 * Conventions
 *  - Every array pointer is a DEVICE pointer, except the pointer tables
 *    (`const double *const *`), which are HOST arrays of device pointers.
 *  - Arrays are the Fortran memory images MOM passes around: column-major,
 *    i fastest, halos included.  A 2-D field is (isd:ied,jsd:jed); a 3-D
 *    field (isd:ied,jsd:jed,1:nk); wrho_bt and dzwt are (isd:ied,jsd:jed,0:nk).
 *    A time level is passed as the address of field(isd,jsd,1,tau).
 *  - Only a one-cell halo is supported (MOM's default ocean domain):
 *    isd = isc-1, ied = iec+1, jsd = jsc-1, jed = jec+1.
 *  - One context owns one domain (a whole global grid or one regional
 *    domain) and one HIP stream.  Calls are asynchronous on that stream;
 *    mom_gpu_sync() waits.  A context is not thread-safe; separate contexts
 *    may be driven from separate threads.
 *  - Every call returns MOM_GPU_OK (0) or a negative error code and leaves a
 *    message in mom_gpu_last_error(ctx).  No exception crosses the ABI.
 */
#ifndef MOM5_GPU_H
#define MOM5_GPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MOM_GPU_OK               0
#define MOM_GPU_ERR_ARG         -1   /* bad argument or unsupported option   */
#define MOM_GPU_ERR_HIP         -2   /* HIP runtime failure                  */
#define MOM_GPU_ERR_NOMEM       -3   /* device allocation failed             */
#define MOM_GPU_ERR_GRID        -4   /* inconsistent grid (e.g. tmask/kmt)   */

/* Advection scheme ids, ocean_core/ocean_parameters.F90:148-163. */
#define MOM_ADVECT_QUICKER       5
#define MOM_ADVECT_MDFL_SWEBY    9
#define MOM_ADVECT_DST_LINEAR   10

typedef struct mom_gpu_ctx mom_gpu_ctx_t;

/* Domain bounds (Fortran, 1-based) -- ocean_core/ocean_types.F90:922-935
 * (ocean_domain_type) with the cyclic flags of ocean_domains.F90:123-230. */
typedef struct {
  int isd, ied, jsd, jed;     /* data domain  */
  int isc, iec, jsc, jec;     /* compute domain */
  int nk;
  int cyclic_x, cyclic_y;
} mom_gpu_domain_t;

/* Grid fields (ocean_core/ocean_types.F90:747-920, ocean_grid_type) on the
 * data domain.  tmask/umask are (isd:ied,jsd:jed,nk); dzt is (1:nk).
 * mom_gpu_create() copies what it needs; the caller may free these after.
 * tmask/umask may both be NULL: the masks are then taken as k<=kmt and
 * k<=kmu (their definition, ocean_grids.F90:1374-1386) without a check. */
typedef struct {
  const double *dxt, *dyt, *dat, *datr;
  const double *dxte, *dyte, *dxtn, *dytn;
  const double *dxu, *dyu, *dxur, *dyur;
  const double *f;
  const int    *kmt, *kmu;
  const double *tmask, *umask;
  const double *dzt;
} mom_gpu_grid_t;

/* Barotropic options (ocean_barotropic.F90:600-630, ocean_barotropic_nml). */
typedef struct {
  int    nts;                 /* subcycles per baroclinic step, :1574     */
  double dtbt;                /* barotropic time step (s)                 */
  double pred_corr_gamma;     /* :614, default 0.2                        */
  double rho0;                /* ocean_parameters.F90:194, 1035           */
  int    stepping_a;          /* 1: barotropic_time_stepping_A, 0: B      */
  int    initsum_with_bar;    /* :609, default true                       */
} mom_gpu_bt_params_t;

/* ---- lifecycle -------------------------------------------------------- */

/* Replaces the grid/mask set-up of ocean_tracer_advect_init (quicker_init
 * :1442-1586, mdfl_init :1644-1691, ocean_tracer_advect.F90:507) and of
 * ocean_barotropic_init (ocean_barotropic.F90:946, :1255-1288).  Verifies
 * tmask == (k <= kmt) and umask == (k <= kmu) on the data domain
 * (ocean_grids.F90:1374-1386). */
int mom_gpu_create(const mom_gpu_domain_t *dom, const mom_gpu_grid_t *grid,
                   mom_gpu_ctx_t **ctx);
int mom_gpu_destroy(mom_gpu_ctx_t *ctx);
int mom_gpu_sync(mom_gpu_ctx_t *ctx);
/* Message of the last failing call on ctx (ctx == NULL: last failure of a
 * call that had no context, e.g. mom_gpu_create). */
const char *mom_gpu_last_error(const mom_gpu_ctx_t *ctx);

/* Device memory helpers for hosts without a GPU allocator (the Fortran
 * host keeps the returned addresses as type(c_ptr)).  The copies first wait
 * for all work queued on the device, on every context's stream. */
int mom_gpu_malloc(void **ptr, size_t bytes);
int mom_gpu_free(void *ptr);
int mom_gpu_memcpy_htod(void *dst, const void *src, size_t bytes);
int mom_gpu_memcpy_dtoh(void *dst, const void *src, size_t bytes);
int mom_gpu_memcpy_dtod(void *dst, const void *src, size_t bytes);
int mom_gpu_memset(void *ptr, int value, size_t bytes);
/* Fill n doubles with value on the context's stream (a kernel, so it is
 * ordered with the other calls on ctx and may be captured in a graph). */
int mom_gpu_fill(mom_gpu_ctx_t *ctx, double *ptr, double value, size_t n);
/* Copy n doubles device to device on the context's stream (a kernel that
 * reads and writes each byte once; also the reference point for measured
 * HBM traffic). */
int mom_gpu_copy(mom_gpu_ctx_t *ctx, double *dst, const double *src, size_t n);

/* Launch-shape knobs, set per context (the autotuner, mom5_amd/tune.py,
 * drives these; results do not depend on them, bit for bit, except FP32
 * results on vdiff_solver).  Knobs:
 *   "barotropic_subcycles_per_launch"  0: three launches per subcycle;
 *                                      n>0: n subcycles per launch (1..12)
 *   "mdfl_rows_per_wave"               1, 2 or 4: rows a wavefront of the MDFL kernel
 *                                      (16 wavefronts with 1 row, 8 otherwise)
 *   "step_group_f64"                   1, 2 or 4 FP64 tracers per batched pass
 *                                      (4 with 1 row x 16 or 2 x 8 wavefronts)
 *   "step_group_f32"                   1, 2, 4, 5, 6, 8 or 10 FP32 tracers per
 *                                      batched pass (at most what the tile
 *                                      shape below is built for: 4, or 10 with
 *                                      1 row x 16 wavefronts, 5 with 2 x 12)
 *   "step_rows_per_wave"               1, 2 or 4: tile height of the FP64 passes
 *   "step_rows_per_wave_f32"           1, 2 or 4: tile height of the FP32 passes
 *   "step_waves_f32"                   8, 12 or 16 wavefronts per FP32 tile (16
 *                                      with 1 or 2 rows a wavefront, 12 with 2;
 *                                      other pairs use 8)
 *   "step_waves_f64"                   8, 12 or 16 wavefronts per FP64 tile (12
 *                                      with 2 rows a wavefront and groups of up
 *                                      to 2 tracers, 16 with 1 row)
 *   "vdiff_prefetch"                   0 or 1: the FP32 implicit diffusion loads
 *                                      one level ahead
 *   "vdiff_block"                      64 or 256 threads a block of the FP32
 *                                      implicit diffusion
 *   "vdiff_merge"                      0 or 1: a step with FP32 tracers solves
 *                                      the implicit diffusion of every tracer
 *                                      in one launch (up to 12 FP32 tracers of
 *                                      mixing class 1 and two FP64 tracers of
 *                                      each class; other mixes: per group)
 *   "vdiff_solver"                     0 or 1: the FP32 tracers' implicit
 *                                      diffusion.  0: one lane per column, the
 *                                      forward sweep through a scratch plane
 *                                      per tracer; 1: two wavefronts per 64
 *                                      columns, every level in registers (the
 *                                      two halves of a column eliminated from
 *                                      both ends), no scratch (2 <= nk <= 80;
 *                                      other nk use 0).  The two round
 *                                      differently: FP32 results depend on
 *                                      this knob, within the same error bound;
 *                                      FP64 results never do.  With 1,
 *                                      vdiff_prefetch, vdiff_block and
 *                                      vdiff_merge have no effect.  Single
 *                                      contexts only (ensembles use 0).
 *   "step_streams"                     1 or 2: with 2, mom_gpu_step_tracers
 *                                      runs its FP32 tracers (advection,
 *                                      update, diffusion) on a second stream
 *                                      of the context, overlapping the FP64
 *                                      tracers and the barotropic calls that
 *                                      follow; needs the register-resident
 *                                      FP32 solver when diffusing.  Every
 *                                      other entry point except the
 *                                      barotropic ones, mom_gpu_sync and
 *                                      mom_gpu_mark first wait for that
 *                                      stream, so results are the same.
 *   "barotropic_land_tiles"            0 or 1: with 1, a region of the blocked
 *                                      barotropic whose every point is inert
 *                                      (both masks 0, forcing and initial
 *                                      velocity zeros, inputs bounded) only
 *                                      accumulates its height, which the
 *                                      subcycles leave as it is (bit for bit
 *                                      the full computation, signed zeros
 *                                      included, while the ocean's state
 *                                      stays finite)
 * Defaults (the fastest measured at 3600x2700x75 with T, S and ten FP32
 * tracers): 5, 1, 2, 10, 1, 1, 16, 16, 1, 256, 0, 1, 1, 1 in the order above.  Unknown knobs and
 * out-of-range values return MOM_GPU_ERR_ARG. */
int mom_gpu_set_tuning(mom_gpu_ctx_t *ctx, const char *knob, int value);
int mom_gpu_get_tuning(const mom_gpu_ctx_t *ctx, const char *knob, int *value);

/* Elapsed time in ms between two marks recorded on the context's stream.
 * mom_gpu_mark(ctx, slot) records slot 0..7; mom_gpu_elapsed waits. */
int mom_gpu_mark(mom_gpu_ctx_t *ctx, int slot);
int mom_gpu_elapsed(mom_gpu_ctx_t *ctx, int slot0, int slot1, float *ms);

/* ---- tracer advection ------------------------------------------------ */

/* Replaces horz_advect_tracer, ocean_tracer_advect.F90:1898-2083, for the
 * MDFL_SWEBY (advect_tracer_mdfl_sweby :3806-4066, sweby_limiter=1),
 * DST_LINEAR (the same routine with sweby_limiter=0) and QUICKER
 * (horz_advect_tracer_quicker :2538-2653) schemes, with
 * advect_sweby_all=.false.:
 *   wrk1 = 0 on the data domain; wrk1(comp) = -scheme(...);
 *   th_tendency(comp) += wrk1.
 * T_taum1 = Tracer%field(:,:,:,taum1); T_tau = Tracer%field(:,:,:,tau)
 * (QUICKER only); rho_dzt_tau = Thickness%rho_dzt(:,:,:,tau) (MDFL only);
 * tmask_limit = Tracer%tmask_limit (QUICKER with limit_with_upwind only).
 * wrk1 and flux_x/flux_y/flux_z (the module diagnostics arrays, data
 * domain) may be NULL. */
int mom_gpu_horz_advect_tracer(mom_gpu_ctx_t *ctx, int scheme,
    const double *T_taum1, const double *T_tau,
    const double *uhrho_et, const double *vhrho_nt, const double *wrho_bt,
    const double *rho_dzt_tau, const double *tmask_limit,
    int limit_with_upwind, double dtime,
    double *th_tendency, double *wrk1,
    double *flux_x, double *flux_y, double *flux_z);

/* Replaces vert_advect_tracer, ocean_tracer_advect.F90:2095-2226: QUICKER
 * (vert_advect_tracer_quicker :2981-3031; tmask_limit is always read, :3006)
 * adds its tendency to th_tendency; the three-dimensional MDFL schemes are
 * a no-op here. */
int mom_gpu_vert_advect_tracer(mom_gpu_ctx_t *ctx, int scheme,
    const double *T_taum1, const double *T_tau, const double *wrho_bt,
    const double *tmask_limit, double *th_tendency, double *wrk1,
    double *flux_z);

/* Replaces advect_tracer_sweby_all, ocean_tracer_advect.F90:4104-4511
 * (advect_sweby_all=.true.): all ntr tracers in one call, with that
 * routine's own rounding order; th_tendency[n](comp) += wrk1(n).
 * wrk1 may be NULL, or hold NULL entries. */
int mom_gpu_advect_sweby_all(mom_gpu_ctx_t *ctx, int ntr,
    const double *const *T_taum1, const double *uhrho_et,
    const double *vhrho_nt, const double *wrho_bt, const double *rho_dzt_tau,
    double dtime, double *const *th_tendency, double *const *wrk1);

/* Replaces compute_tmask_limit, ocean_tracer.F90:2788-2892.
 * index_temp/index_salt are 0-based positions in the tables (-1: absent).
 * Writes (isc-1:iec, jsc-1:jec, 1:nk) of each tmask_limit, like the
 * reference; other points are left untouched. */
int mom_gpu_compute_tmask_limit(mom_gpu_ctx_t *ctx, int ntr,
    const double *const *T_tau, const double *min_limit,
    const double *max_limit, int index_temp, int index_salt, int ts_same,
    double *const *tmask_limit);

/* Replaces the explicit update of update_ocean_tracer, ocean_tracer.F90:
 * 2341-2350: T_taup1 = (rho_dzt_taum1*T_taum1 + dtime*th_tendency)*rho_dztr
 * on the compute domain. */
int mom_gpu_update_tracer_explicit(mom_gpu_ctx_t *ctx, const double *T_taum1,
    const double *th_tendency, const double *rho_dzt_taum1,
    const double *rho_dztr, double dtime, double *T_taup1);

/* ---- implicit vertical diffusion -------------------------------------- */

/* Replaces invtri(z, topbc, botbc, dcb, tdt, kmz, mask, dh, dhw, aidif, nk),
 * ocean_core/ocean_util.F90:153-222 (nk from the context). */
int mom_gpu_invtri(mom_gpu_ctx_t *ctx, double *z, const double *topbc,
    const double *botbc, const double *dcb, double tdt, const int *kmz,
    const double *mask, const double *dh, const double *dhw, double aidif);

/* Replaces the tracer loop of vert_diffuse_implicit,
 * ocean_param/vertical/ocean_vert_mix.F90:3426-3513: for each tracer
 * dcb = rho0*diff_cbt(:,:,:,nmix) + K33_implicit(n) (nmix = 2 for
 * index_salt, else 1; diff_cbt holds both planes back to back) and invtri
 * with kmz = Grd%kmt, mask = Grd%tmask, dh = rho_dzt_taup1, dhw = dzwt.
 * K33_implicit may be NULL or hold NULL entries (zero).  Tracers that share
 * nmix and have no K33 share one factorisation (same arithmetic). */
int mom_gpu_vert_diffuse_implicit(mom_gpu_ctx_t *ctx, int ntr,
    double *const *T_taup1, const double *const *stf,
    const double *const *btf, const double *diff_cbt,
    const double *const *K33_implicit, int index_salt, double rho0,
    const double *rho_dzt_taup1, const double *dzwt, double dtime_t,
    double aidif);

/* ---- advection velocity ------------------------------------------------ */

/* Replaces mpp_update_domains(field, Dom%domain2d) on a single PE
 * (shared/mpp/include/mpp_do_update.h:186-288, ocean_core/ocean_domains.F90:
 * 123-230) for a device-resident field of nz levels on the data domain: a
 * halo column of a compute row takes its cyclic-x image, a halo row of a
 * compute column its cyclic-y image, a corner only when both axes are
 * cyclic; halo points on a closed axis are left as they are.  MOM calls it
 * on T(taup1) after the tracer update (ocean_model.F90:2049). */
int mom_gpu_update_domains(mom_gpu_ctx_t *ctx, double *field, int nz);

/* Replaces the B-grid branch of ocean_advection_velocity
 * (ocean_core/ocean_advection_velocity.F90:568-672; no blobs, OBC or
 * prescribed inflow): the T-cell advection transports from the U-cell
 * velocities and continuity,
 *   uhrho_et = BAY(u(:,:,k,1)*dyu*rho_dzu)/dyte,
 *   vhrho_nt = BAX(u(:,:,k,2)*dxu*rho_dzu)/dxtn,
 *   diverge_t = tmask*(BDX_ET(uhrho_et) + BDY_NT(vhrho_nt)),
 *   wrho_bt(0) = -(pme+river),
 *   wrho_bt(k) = (rho_dzt_tendency - mass_source + diverge_t + wrho_bt(k-1))*tmask.
 * u holds Velocity%u(:,:,:,1:2,tau) back to back, rho_dzu
 * Thickness%rho_dzu(:,:,:,tau).  rho_dzt_tendency, mass_source (zero) and
 * diverge_t (not written) may be NULL.  The U-cell components (uhrho_eu,
 * vhrho_nu, wrho_bu) feed the momentum equations and are not computed. */
int mom_gpu_advection_velocity(mom_gpu_ctx_t *ctx, const double *u,
    const double *rho_dzu, const double *pme, const double *river,
    const double *rho_dzt_tendency, const double *mass_source,
    double *uhrho_et, double *vhrho_nt, double *wrho_bt, double *diverge_t);

/* ---- barotropic -------------------------------------------------------- */

/* Replaces ocean_barotropic_forcing (ocean_core/ocean_barotropic.F90:
 * 3058-3141) with the default namelist (no zeroed forcing, no barotropic
 * friction, no OBC) on the B-grid: forcing_bt(:,:,n) = sum over the wet
 * U-levels of accel(:,:,k,n) + umask(k=1)*(smf(:,:,n) - bmf(:,:,n)) on the
 * compute domain, halos as the single-PE update leaves them.  accel holds
 * Velocity%accel (data domain, nk, 2 components); smf, bmf and forcing_bt
 * (data domain, 2 components). */
int mom_gpu_barotropic_forcing(mom_gpu_ctx_t *ctx, const double *accel,
    const double *smf, const double *bmf, double *forcing_bt);

/* Replaces update_ocean_barotropic (ocean_barotropic.F90:2763-2894) for the
 * B-grid depth-based case: pred_corr_tropic_depth_bgrid (:3225-3579, with
 * barotropic_halo=1 and use_legacy_barotropic_halos=.true., the defaults;
 * no tides, geoid, friction, smoothing or OBC) and the time-averaging tail
 * (:2804-2811, :2846-2888).  udrho_* and forcing_bt hold the two velocity
 * components back to back.  conv_rho_ud_t (= -DIV_UD of the averaged
 * transport) may be NULL.  Supports cyclic_y = 0 only.  The outputs must
 * not alias eta_t_tau, eta_t_bar_tau or udrho_tau (MOM passes other time
 * levels).
 * Precondition (what mpp_update_domains leaves in MOM on a cyclic-x
 * domain): halo column isd equals column iec and ied equals isc, on every
 * row, for the state and for the 2-D inputs. */
int mom_gpu_update_ocean_barotropic(mom_gpu_ctx_t *ctx,
    const mom_gpu_bt_params_t *p, const double *eta_t_tau,
    const double *eta_t_bar_tau, const double *udrho_tau,
    const double *forcing_bt, const double *thicku, const double *patm,
    const double *pme, const double *river, const double *source,
    const double *eta_smooth, const double *rho_g,
    double *eta_t_bar_taup1, double *udrho_taup1, double *conv_rho_ud_t);

/* Replaces eta_and_pbot_tendency (ocean_barotropic.F90:2667-2726, no
 * blobs, no zero_tendency) followed by eta_and_pbot_update (:2290-2385, no
 * OBC, truncate_eta false, split time stepping).  Depth-based
 * (pressure_based = 0), on the data domain:
 *   tendency = tmask(k=1)*(1/rho0)*(conv_rho_ud_t_tau + pme + river + source)
 *   (Ext_mode%deta_dt); x_taup1 = x_taum1 + dtime*tendency on the compute
 *   domain (eta_t), halos as the single-PE update leaves them (closed halo
 *   points keep x_taup1's previous contents).
 * Pressure-based: tendency = dpatm_dt + grav*tmask(k=1)*(...) (dpbot_dt),
 * x is anompb, and pbot_t_taup1 = x_taup1 + rho0*grav*ht on the data domain.
 * dtime is Time_steps%dtime_e; grav defaults to 9.80 in the reference
 * (ocean_parameters.F90:176).  x_taup1 and tendency must not alias x_taum1;
 * dpatm_dt, ht and pbot_t_taup1 may be NULL when depth-based. */
int mom_gpu_eta_and_pbot_update(mom_gpu_ctx_t *ctx, int pressure_based, double rho0,
    double grav, double dtime, const double *conv_rho_ud_t_tau, const double *pme,
    const double *river, const double *source, const double *dpatm_dt, const double *ht,
    const double *x_taum1, double *x_taup1, double *pbot_t_taup1, double *tendency);

/* Grid metrics of REMAP_BT_TO_BU (ocean_operators.F90:462-477): Grd%dte,
 * dtw, dus, dun, daur (ocean_types.F90:281-293), HOST arrays on the data
 * domain, copied into the context (the part of ocean_grids_init the
 * closure needs; mom_gpu_barotropic_closure fails until they are set). */
int mom_gpu_set_remap_metrics(mom_gpu_ctx_t *ctx, const double *dte, const double *dtw,
    const double *dus, const double *dun, const double *daur);

/* Replaces the closure of update_ocean_barotropic (ocean_barotropic.F90:
 * 2906-2921 and 2956-2985; no OBC, tidal or geoid forcing): depth-based,
 * eta_u = umask(k=1)*REMAP_BT_TO_BU(eta_t_bar); pressure-based, eta_u is
 * pbot_u = umask(k=1)*REMAP_BT_TO_BU(rho0*grav*ht + anompb_bar); both then
 * halo-updated.  ps = rho_g*eta_t_bar + patm (Ext_mode%ps), grad_ps =
 * GRAD_BAROTROPIC_P(ps) and grad_anompb = GRAD_BAROTROPIC_P(anompb_bar)
 * (ocean_operators.F90:631-676; two components back to back): B-grid when
 * tmasken is NULL, C-grid with tmasken = Grd%tmasken(:,:,1,1:2) (device).
 * eta_t_bar and anompb_bar are the itime level (taup1 with two-level
 * stepping, :2797-2801).  anompb_bar and grad_anompb may both be NULL when
 * depth-based; ht is read only when pressure-based. */
int mom_gpu_barotropic_closure(mom_gpu_ctx_t *ctx, int pressure_based, double rho0,
    double grav, const double *eta_t_bar, const double *anompb_bar, const double *patm,
    const double *rho_g, const double *ht, const double *tmasken, double *eta_u, double *ps,
    double *grad_ps, double *grad_anompb);

/* The four split-explicit barotropic schemes update_ocean_barotropic
 * selects (ocean_barotropic.F90:2829-2843), with barotropic_halo = 1 and no
 * tides, geoid, friction, smoothing, blobs or OBC, each followed by the
 * time-mean tail (:2846-2894):
 *   cgrid = 0, pressure_based = 0: pred_corr_tropic_depth_bgrid :3225-3579
 *   cgrid = 1, pressure_based = 0: pred_corr_tropic_depth_cgrid :3601-3965
 *   cgrid = 0, pressure_based = 1: pred_corr_tropic_press_bgrid :3989-4329
 *   cgrid = 1, pressure_based = 1: pred_corr_tropic_press_cgrid :4355-4693
 * x is eta_t (depth-based) or anompb (pressure-based).  thick is thicku
 * (B, depth), Thickness%thicken (C, depth; 2 planes), Thickness%mass_u
 * (B, pressure) or Thickness%mass_en (C, pressure; 2 planes).  tmasken
 * (C-grid) = Grd%tmasken(:,:,1,1:2).  smooth is eta_smooth (depth) or
 * pbot_smooth (pressure).  x_tau is read only when initsum_with_bar = 0;
 * conv_rho_ud_t may be NULL.  All 2-D fields on the data domain, device
 * pointers.  The B-grid depth-based case is also served (faster, blocked in
 * time) by mom_gpu_update_ocean_barotropic; both give the same bits. */
typedef struct {
  int cgrid;
  int pressure_based;
  double grav;                      /* ocean_parameters.F90:176 (9.80) */
  const double *x_tau, *x_bar_tau, *udrho_tau, *forcing_bt, *thick, *tmasken;
  const double *patm, *rho_g;       /* depth-based */
  const double *dpatm_dt;           /* pressure-based */
  const double *pme, *river, *source, *smooth;
  double *x_bar_taup1, *udrho_taup1, *conv_rho_ud_t;
} mom_gpu_bt_variant_t;

int mom_gpu_update_ocean_barotropic_variant(mom_gpu_ctx_t *ctx, const mom_gpu_bt_params_t *p,
    const mom_gpu_bt_variant_t *v);

/* ---- batched tracer step ------------------------------------------------ */

/* All tracers of one baroclinic step: replaces the tracer loop of
 * update_ocean_tracer (ocean_tracers/ocean_tracer.F90:2295-2350: advection
 * into th_tendency and the explicit update to taup1) followed by
 * vert_diffuse_implicit (:2432, ocean_param/vertical/ocean_vert_mix.F90:
 * 3426-3513), for tracers whose advection scheme is MDFL_SWEBY or QUICKER.
 *
 * Tracers are advanced in groups that share one pass over the shared
 * fields.  dtype[n] = 8: FP64 storage, reference operation order --
 * T_taup1 is bit-for-bit what mom_gpu_horz_advect_tracer (into a zeroed
 * th_tendency plus th_tendency[n]), mom_gpu_update_tracer_explicit and
 * mom_gpu_vert_diffuse_implicit produce.  dtype[n] = 4: FP32 storage and
 * FP32 advection arithmetic with the division-free limiter; the implicit
 * vertical solve stays FP64 (the forward-sweep values are stored in FP32).
 *
 * Optional pointers: th_tendency (table or entries; other tendency terms),
 * rho_dzt_taum1 (default rho_dzt_tau, two-level stepping), rho_dztr
 * (default 1/(rho_dzt_tau+epsln), ocean_thickness.F90:1567),
 * rho_dzt_taup1 (default rho_dzt_tau), stf/btf (NULL: no implicit vertical
 * diffusion).
 *
 * scheme (host array, NULL: every tracer MOM_ADVECT_MDFL_SWEBY) selects
 * MOM_ADVECT_QUICKER per tracer (FP64 storage only).  A Quicker tracer gets
 * horz_advect_tracer and vert_advect_tracer (ocean_tracer_advect.F90:
 * 2538-2653, 2981-3031) into th_tendency, with its tmask_limit
 * (compute_tmask_limit, ocean_tracer.F90:2788-2892) formed inside the
 * kernel from T_taum1 (= T(tau)) and the range min_limit[n]..max_limit[n];
 * with tmask_limit_ts_same and index_temp, index_salt >= 0 the temperature
 * and salinity flags are OR'ed (the partner's T_taum1 is read);
 * limit_with_upwind applies the upwind fallback to the horizontal fluxes
 * (the vertical one always uses it).  Values beyond the compute domain are
 * read as the halo update leaves them (cyclic wrap; nothing beyond a closed
 * boundary), so T_taup1 equals bit for bit mom_gpu_compute_tmask_limit,
 * mom_gpu_horz/vert_advect_tracer into a zeroed th_tendency (plus
 * th_tendency[n]) and mom_gpu_update_tracer_explicit on halo-updated
 * inputs.  Ensembles take MDFL-Sweby tracers only. */
typedef struct {
  int ntr;
  const void *const *T_taum1;       /* host table of device pointers */
  void *const *T_taup1;
  const int *dtype;                 /* host array: 8 (FP64) or 4 (FP32) */
  const double *const *th_tendency;
  const double *const *stf;
  const double *const *btf;
  int index_salt;                   /* 0-based, -1: none */
  const double *uhrho_et, *vhrho_nt, *wrho_bt, *rho_dzt_tau;
  const double *rho_dzt_taum1, *rho_dztr, *rho_dzt_taup1;
  const double *dzwt, *diff_cbt;
  double rho0, dtime, aidif;
  const int *scheme;                /* host array or NULL */
  const double *min_limit, *max_limit;  /* host arrays, read for Quicker tracers */
  int index_temp;                   /* 0-based, -1: none */
  int tmask_limit_ts_same;
  int limit_with_upwind;
} mom_gpu_tracer_step_t;

int mom_gpu_step_tracers(mom_gpu_ctx_t *ctx, const mom_gpu_tracer_step_t *s);

/* ---- ensembles of domains ------------------------------------------------
 * No reference counterpart: several independent domains (each its own
 * context, so its own grid, bathymetry, time step and subcycle count) are
 * advanced together.  Each phase of a step is one launch over all domains:
 * the tracer passes use a persistent grid whose work-groups take (domain,
 * tile) items from a device-side queue, largest domain first; the implicit
 * diffusion and each block of barotropic subcycles cover every domain in one
 * grid.  Launches go to the stream of ctxs[0] (capture that context to graph
 * a whole ensemble step).  Outside a capture an ensemble call is ordered
 * after work already queued on every member context and before work queued
 * on them afterwards; while ctxs[0] is being captured the other members must
 * be idle (mom_gpu_sync them first), since their streams take no part in the
 * graph.  Calls that overlap with an ensemble call from another host thread
 * are not ordered.  Launch tables used while capturing stay allocated until
 * mom_gpu_ensemble_destroy (at most 64 tables are held).  Results are
 * bit-for-bit those of the per-domain calls. */
typedef struct mom_gpu_ensemble mom_gpu_ensemble_t;

/* The 2-D barotropic fields of one domain (as mom_gpu_update_ocean_barotropic). */
typedef struct {
  const double *eta_t_tau, *eta_t_bar_tau, *udrho_tau, *forcing_bt, *thicku;
  const double *patm, *pme, *river, *source, *eta_smooth, *rho_g;
  double *eta_t_bar_taup1, *udrho_taup1, *conv_rho_ud_t;
} mom_gpu_bt_fields_t;

int mom_gpu_ensemble_create(int ndom, mom_gpu_ctx_t *const *ctxs, mom_gpu_ensemble_t **out);
int mom_gpu_ensemble_destroy(mom_gpu_ensemble_t *ens);
/* steps[d]: domain d's batched tracer step (see mom_gpu_step_tracers). */
int mom_gpu_ensemble_step_tracers(mom_gpu_ensemble_t *ens, const mom_gpu_tracer_step_t *steps);
/* p[d], f[d]: domain d's barotropic parameters and fields (ndom <= 64,
 * cyclic_y = 0 for every domain; outputs must not alias eta_t_tau,
 * eta_t_bar_tau or udrho_tau). */
int mom_gpu_ensemble_update_barotropic(mom_gpu_ensemble_t *ens, const mom_gpu_bt_params_t *p,
                                       const mom_gpu_bt_fields_t *f);

/* ---- hipGraph capture of whole steps -------------------------------------
 * No reference counterpart (MOM launches nothing): a caller captures the
 * entry points of one or more steps once and replays them.  Between begin
 * and end only kernel-launching entry points may be called on ctx
 * (advection, explicit update, tmask limit, invtri / vert_diffuse_implicit,
 * step_tracers, barotropic, fill, copy); copies, syncs, marks, content and
 * error reports end the capture with an error.  Arguments (addresses,
 * scalars, knobs) are frozen at capture: with two-level time stepping,
 * capture an even number of steps so a replay leaves the time levels where
 * it found them.  A replay runs asynchronously on ctx's stream. */
typedef struct mom_gpu_graph mom_gpu_graph_t;
int mom_gpu_graph_begin(mom_gpu_ctx_t *ctx);
int mom_gpu_graph_end(mom_gpu_ctx_t *ctx, mom_gpu_graph_t **graph);
int mom_gpu_graph_launch(mom_gpu_ctx_t *ctx, const mom_gpu_graph_t *graph);
int mom_gpu_graph_nodes(const mom_gpu_graph_t *graph);
int mom_gpu_graph_destroy(mom_gpu_graph_t *graph);

/* ---- diagnostics -------------------------------------------------------- */

/* Masked tracer content sum(T*rho_dzt*dat*tmask) over the compute domain,
 * the quantity of tracer_integrals (ocean_diag/ocean_tracer_diag.F90:
 * 2648-2661), accumulated with compensated (double-double) sums.  Waits
 * for the stream and writes the result to a HOST double. */
int mom_gpu_tracer_content(mom_gpu_ctx_t *ctx, const double *T,
    const double *rho_dzt, double *content);

/* The same sum for a tracer stored in FP32 (each value widened to FP64,
 * products and sums in FP64 with the same compensation). */
int mom_gpu_tracer_content_f32(mom_gpu_ctx_t *ctx, const float *T,
    const double *rho_dzt, double *content);

/* The masked content of mom_gpu_tracer_content (dtype 8) or
 * mom_gpu_tracer_content_f32 (dtype 4), bit for bit the same value, written
 * to DEVICE memory (content_dev: one double) and computed without allocating
 * anything: no hipMalloc, hipFree, host copy or stream sync.  No counterpart
 * in the reference: for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops).  The partial sums live in one fixed array of
 * the library per GPU (1024 blocks' partials, 16 KiB, held until the process
 * ends); a context whose reduction uses more blocks is refused.  Runs
 * asynchronously on ctx's stream.  Calls that overlap on one GPU (other
 * streams, other contexts on that GPU) share that array and race. */
int mom_gpu_tracer_content_noalloc(mom_gpu_ctx_t *ctx, const void *T, int dtype,
    const double *rho_dzt, double *content_dev);

/* Error of a tracer a (dtype 8: FP64, 4: FP32) against an FP64 result b
 * over the wet compute cells (k <= kmt).  No reference counterpart: the
 * report of each result's error against an FP64 run.  Writes 7 doubles to
 * the HOST array out: max|a-b|, max|b|, sum (a-b)^2, sum b^2, number of
 * cells, number of cells whose FP64 bit patterns differ, and
 * max |a-b|/|b| over cells with |b| >= rel_floor.  Waits for the stream. */
int mom_gpu_tracer_error(mom_gpu_ctx_t *ctx, const void *a, int dtype,
    const double *b, double rel_floor, double *out);

#ifdef __cplusplus
}
#endif
#endif /* MOM5_GPU_H */
