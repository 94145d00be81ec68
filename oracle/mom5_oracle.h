/*
 * mom5_oracle.h -- CPU restatement of the MOM5 ocean hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * kernels are compared against, and the "port" CPU baseline timed beside
 * them.  Nothing in the product (mom5_amd/) links, loads or calls it.
 *
 * Every routine keeps the reference's loop structure (per tracer, i
 * innermost) and its left-to-right floating-point evaluation order, and is
 * compiled with -ffp-contract=off so that no multiply-add is fused.
 *
 * Array convention (same as the Fortran memory images MOM passes around):
 *   2-D   (isd:ied, jsd:jed)          i fastest
 *   3-D   (isd:ied, jsd:jed, 1:nk)
 *   3-Dw  (isd:ied, jsd:jed, 0:nk)    wrho_bt, dzwt
 *   comp  (isc:iec, jsc:jec, 1:nk)    Fortran function results
 * Indices in the domain struct are the Fortran (1-based) bounds.
 */
#ifndef MOM5_ORACLE_H
#define MOM5_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int isd, ied, jsd, jed;   /* data domain  */
  int isc, iec, jsc, jec;   /* compute domain */
  int nk;
  int cyclic_x, cyclic_y;
} orc_dom;

typedef struct {
  const double *dxt, *dyt, *dat, *datr;
  const double *dxte, *dyte, *dxtn, *dytn;
  const double *dxu, *dyu, *dxur, *dyur;
  const double *f;
  const int    *kmt;
  const double *tmask;      /* 3-D */
  const double *umask;      /* 3-D */
  const double *dzt;        /* (1:nk) */
} orc_grid;

typedef struct {
  int    nts;               /* barotropic subcycles */
  double dtbt;
  double pred_corr_gamma;
  double rho0;
  int    stepping_a;        /* 1: barotropic_time_stepping_A, 0: _B */
  int    initsum_with_bar;
} orc_bt_params;

/* ocean_tracer_advect.F90:3806-4066 advect_tracer_mdfl_sweby; result on the
 * compute domain.  flux_x/y/z (data domain) may be NULL. */
void orc_advect_tracer_mdfl_sweby(const orc_dom *d, const orc_grid *g,
    const double *T, const double *uhrho_et, const double *vhrho_nt,
    const double *wrho_bt, const double *rho_dzt, double dtime,
    double sweby_limiter, double *advect,
    double *flux_x, double *flux_y, double *flux_z);

/* ocean_tracer_advect.F90:1898-2083 horz_advect_tracer for the MDFL-Sweby
 * (lim=1) / DST-linear (lim=0) schemes: wrk1 = -advect over the data domain
 * (zero outside the compute domain), th_tendency += wrk1.  wrk1 may be NULL. */
void orc_horz_advect_tracer_mdfl(const orc_dom *d, const orc_grid *g,
    const double *T, const double *uhrho_et, const double *vhrho_nt,
    const double *wrho_bt, const double *rho_dzt, double dtime,
    double sweby_limiter, double *th_tendency, double *wrk1);

/* ocean_tracer_advect.F90:4104-4511 advect_tracer_sweby_all: all tracers,
 * th_tendency(n) += wrk1(n); wrk1 (data domain) may be NULL. */
void orc_advect_tracer_sweby_all(const orc_dom *d, const orc_grid *g, int ntr,
    const double *const *T, const double *uhrho_et, const double *vhrho_nt,
    const double *wrho_bt, const double *rho_dzt, double dtime,
    double *const *th_tendency, double *const *wrk1);

/* ocean_tracer_advect.F90:1442-1586 quicker_init: weights on the data
 * domain.  quick_x/y: 2 planes, curv_*: 3 planes, quick_z: (nk,2),
 * curv_zp/zn: (nk,3), all column-major like the reference. */
void orc_quicker_init(const orc_dom *d, const orc_grid *g,
    double *quick_x, double *quick_y, double *curv_xp, double *curv_xn,
    double *curv_yp, double *curv_yn, double *quick_z, double *curv_zp,
    double *curv_zn);

/* ocean_tracer_advect.F90:2538-2653 horz_advect_tracer_quicker (result on
 * the compute domain). tmask_limit is read only if limit_with_upwind. */
void orc_horz_advect_tracer_quicker(const orc_dom *d, const orc_grid *g,
    const double *T_taum1, const double *T_tau, const double *uhrho_et,
    const double *vhrho_nt, const double *tmask_limit, int limit_with_upwind,
    double *result, double *flux_x, double *flux_y);

/* ocean_tracer_advect.F90:2981-3031 vert_advect_tracer_quicker. */
void orc_vert_advect_tracer_quicker(const orc_dom *d, const orc_grid *g,
    const double *T_taum1, const double *T_tau, const double *wrho_bt,
    const double *tmask_limit, double *result, double *flux_z);

/* ocean_tracer.F90:2788-2892 compute_tmask_limit. index_temp/index_salt are
 * 0-based (-1 when absent). */
void orc_compute_tmask_limit(const orc_dom *d, const orc_grid *g, int ntr,
    const double *const *T_tau, const double *min_limit,
    const double *max_limit, int index_temp, int index_salt, int ts_same,
    double *const *tmask_limit);

/* ocean_tracer.F90:2341-2350 explicit update. */
void orc_update_tracer_explicit(const orc_dom *d, const double *T_taum1,
    const double *th_tendency, const double *rho_dzt_taum1,
    const double *rho_dztr, double dtime, double *T_taup1);

/* ocean_util.F90:153-222 invtri. */
void orc_invtri(const orc_dom *d, double *z, const double *topbc,
    const double *botbc, const double *dcb, double tdt, const int *kmz,
    const double *mask, const double *dh, const double *dhw, double aidif);

/* ocean_vert_mix.F90:3426-3513 vert_diffuse_implicit (the invtri part):
 * dcb = rho0*diff_cbt(:,:,:,nmix) + K33_implicit(n); nmix=2 for index_salt.
 * diff_cbt holds 2 planes of 3-D; K33 entries may be NULL (zero). */
void orc_vert_diffuse_implicit(const orc_dom *d, const orc_grid *g, int ntr,
    double *const *T_taup1, const double *const *stf,
    const double *const *btf, const double *diff_cbt,
    const double *const *K33_implicit, int index_salt, double rho0,
    const double *rho_dzt_taup1, const double *dzwt, double dtime_t,
    double aidif);

/* ocean_barotropic.F90:3225-3579 pred_corr_tropic_depth_bgrid followed by
 * the averaging tail of update_ocean_barotropic (:2804-2811, :2846-2894).
 * udrho_* and forcing_bt hold 2 planes.  conv_rho_ud_t may be NULL. */
void orc_update_ocean_barotropic(const orc_dom *d, const orc_grid *g,
    const orc_bt_params *p, const double *eta_t_tau,
    const double *eta_t_bar_tau, const double *udrho_tau,
    const double *forcing_bt, const double *thicku, const double *patm,
    const double *pme, const double *river, const double *source,
    const double *eta_smooth, const double *rho_g,
    double *eta_t_bar_taup1, double *udrho_taup1, double *conv_rho_ud_t);

/* ocean_tracer_diag.F90:2648-2661 masked tracer content
 * sum(T*rho_dzt*dat*tmask) over the compute domain, accumulated in long
 * double, column by column as the reference's wrk1_2d. */
double orc_tracer_content(const orc_dom *d, const orc_grid *g,
    const double *T, const double *rho_dzt);

/* Single-PE halo semantics of mpp_update_domains on an array whose index
 * space is (isc-h:iec+h, jsc-h:jec+h, nz).  flags: 1 = x, 2 = y, 3 = both. */
void orc_halo_update(const orc_dom *d, int h, double *a, int nz, int flags);

/* ocean_core/ocean_advection_velocity.F90:568-672, B-grid branch: T-cell
 * advection transports from the U-cell velocities u (two 3-D components
 * back to back) and continuity.  rho_dzt_tendency, mass_source (3-D) and
 * diverge_t (output) may be NULL.  All arrays on the data domain. */
void orc_advection_velocity(const orc_dom *d, const orc_grid *g, const double *u,
    const double *rho_dzu, const double *pme, const double *river,
    const double *rho_dzt_tendency, const double *mass_source,
    double *uhrho_et, double *vhrho_nt, double *wrho_bt, double *diverge_t);

/* ocean_core/ocean_barotropic.F90:3058-3141 (default namelist, B-grid):
 * forcing_bt(:,:,n) = sum_k umask*accel(:,:,k,n) + umask(1)*(smf - bmf) on
 * the compute domain, halo-updated.  accel (data domain, nk, 2); smf, bmf,
 * forcing_bt (data domain, 2). */
void orc_barotropic_forcing(const orc_dom *d, const orc_grid *g, const double *accel,
    const double *smf, const double *bmf, double *forcing_bt);

/* ocean_barotropic.F90:2667-2700 + :2290-2363: surface height (or, when
 * pressure_based, bottom pressure anomaly) tendency and "big" time step;
 * all arrays 2-D on the data domain.  dpatm_dt, ht and pbot_taup1 are read
 * or written only when pressure_based. */
void orc_eta_and_pbot_update(const orc_dom *d, const orc_grid *g, int pressure_based,
    double rho0, double grav, double dtime, const double *conv_tau, const double *pme,
    const double *river, const double *source, const double *dpatm_dt, const double *ht,
    const double *x_taum1, double *x_taup1, double *pbot_taup1, double *tendency);

/* ocean_barotropic.F90:2906-2921, :2956-2985: eta_u (pbot_u) closure by
 * REMAP_BT_TO_BU, surface pressure ps and the gradients grad_ps,
 * grad_anompb ((isd:ied,jsd:jed,2)); B-grid, or C-grid when tmasken is
 * given. */
void orc_barotropic_closure(const orc_dom *d, const orc_grid *g, int pressure_based,
    double rho0, double grav, const double *eta_t_bar, const double *anompb_bar,
    const double *patm, const double *rho_g, const double *ht, const double *tmasken,
    const double *dte, const double *dtw, const double *dus, const double *dun,
    const double *daur, double *eta_u, double *ps, double *grad_ps, double *grad_anompb);

/* The four split-explicit barotropic schemes with barotropic_halo = 1 (no
 * tides, geoid, friction, smoothing, blobs or OBC) and the tail of
 * update_ocean_barotropic: cgrid selects pred_corr_tropic_*_cgrid
 * (ocean_barotropic.F90:3601-3965, 4355-4693) over *_bgrid (:3225-3579,
 * 3989-4329), press the pressure-based form (x = anompb) over the
 * depth-based one (x = eta_t).  thick: thicku (B, depth), thicken (C, depth,
 * 2 planes), mass_u (B, press), mass_en (C, press, 2 planes); tmasken
 * (C-grid) = Grd%tmasken(:,:,1,1:2); patm, rho_g (depth); dpatm_dt (press);
 * smooth = eta_smooth (depth) or pbot_smooth (press).  x_tau is read only
 * when initsum_with_bar = 0. */
void orc_update_ocean_barotropic_variant(const orc_dom *d, const orc_grid *g,
    const orc_bt_params *p, int cgrid, int press, double grav, const double *x_tau,
    const double *xbar_tau, const double *udrho_tau, const double *forcing_bt,
    const double *thick, const double *tmasken, const double *patm, const double *rho_g,
    const double *dpatm_dt, const double *pme, const double *river, const double *source,
    const double *smooth, double *xbar, double *udrho, double *conv_rho_ud_t);

#ifdef __cplusplus
}
#endif
#endif
